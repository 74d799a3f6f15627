"""HIP path against the CPU oracle, through the C ABI.

Sub-steps are compared one by one from the same random state, then whole
runs.  Both sides compute in float32 without fused multiply-add and in the
same association order, so the expected agreement is bit for bit
(np.array_equal, which treats -0 and +0 as equal).
"""

import numpy as np
import pytest

import oracle
from oracle import solver as osolver
from gpu_models import plane_model, randomise_tables, rich_model
from models import BASIC_MODELS, PLANE_MODELS

pytestmark = pytest.mark.gpu

COMPS = ('Ex', 'Ey', 'Ez', 'Hx', 'Hy', 'Hz')


def _hip():
    from gprmax_amd.hip import HipSolver
    return HipSolver


def _random_state(G, seed):
    """Random fields, Phi and T on the oracle grid; returns the arrays."""
    rng = np.random.default_rng(seed)
    osolver.initialise_state(G)
    for c in COMPS:
        getattr(G, c)[...] = rng.standard_normal(getattr(G, c).shape).astype(np.float32)
    for pml in G.pmls:
        for name in ('EPhi1', 'EPhi2', 'HPhi1', 'HPhi2'):
            a = getattr(pml, name)
            a[...] = rng.standard_normal(a.shape).astype(np.float32)
    if G.maxpoles:
        real = np.all(G.updatecoeffsdispersive.imag == 0)
        # T stays exactly zero on edges whose material has no poles; a
        # physical state never holds anything else there.
        has_poles = np.any(G.updatecoeffsdispersive != 0, axis=1)
        for c, name in enumerate(('Tx', 'Ty', 'Tz')):
            a = getattr(G, name)
            re = rng.standard_normal(a.shape)
            a[...] = re if real else re + 1j * rng.standard_normal(a.shape)
            a[:, ~has_poles[G.ID[c]]] = 0


def _push_state(s, G):
    s.upload_fields(np.stack([getattr(G, c) for c in COMPS]))
    for n, pml in enumerate(G.pmls):
        for w, name in enumerate(('EPhi1', 'EPhi2', 'HPhi1', 'HPhi2')):
            s.upload_phi(n, w, getattr(pml, name))
    if G.maxpoles:
        s.upload_T(np.stack([G.Tx, G.Ty, G.Tz]))


def _assert_state_equal(s, G, what):
    F = s.download_fields()
    for n, c in enumerate(COMPS):
        assert np.array_equal(F[n], getattr(G, c)), (what, c, np.abs(F[n] - getattr(G, c)).max())
    for n, pml in enumerate(G.pmls):
        for w, name in enumerate(('EPhi1', 'EPhi2', 'HPhi1', 'HPhi2')):
            assert np.array_equal(s.download_phi(n, w), getattr(pml, name)), (what, pml.ID, name)
    if G.maxpoles:
        T = s.download_T()
        for n, name in enumerate(('Tx', 'Ty', 'Tz')):
            assert np.array_equal(T[n], getattr(G, name)), (what, name)


CASES = {
    'horipml1': dict(order=1, formulation='HORIPML'),
    'horipml2': dict(order=2, formulation='HORIPML'),
    'mripml1': dict(order=1, formulation='MRIPML'),
    'mripml2': dict(order=2, formulation='MRIPML'),
    'debye3': dict(poles=3, kind='debye'),
    'debye1': dict(poles=1, kind='debye'),
    'lorentz2': dict(poles=2, kind='lorentz'),
    'drude1': dict(poles=1, kind='drude'),
}


@pytest.mark.parametrize('case', list(CASES))
def test_each_substep_matches_oracle(case):
    """update_magnetic, every H-PML slab, magnetic sources, update_electric /
    dispersive A, every E-PML slab, electric sources, dispersive B, receivers
    and snapshots, each from the same random state."""
    _check_substeps(rich_model(seed=7, **CASES[case]))


# Cells of 1, 1.5 and 2 mm: the three CB (DB) columns of a physical table and
# the three slab spacings all differ.
ANISO = (1e-3, 1.5e-3, 2e-3)


@pytest.mark.parametrize('case', list(CASES))
def test_each_substep_matches_oracle_anisotropic_random_tables(case):
    """The sub-steps on cells of 1 x 1.5 x 2 mm with every coefficient
    column and every PML R table drawn at random: a swapped CBx/CBy/CBz or
    DBx/DBy/DBz column, a wrong source column or a slab spacing or profile
    taken from the wrong axis breaks the match."""
    G = randomise_tables(rich_model(seed=17, cell=ANISO, **CASES[case]), 23)
    assert len({G.dx, G.dy, G.dz}) == 3
    _check_substeps(G)


def _check_substeps(G):
    lib = oracle.load(np.float32)
    f = lib.funcs
    it = 9
    with _hip()(G) as s:
        steps = [('h_main', 0, lambda: f['update_magnetic'](G.nx, G.ny, G.nz, 4, oracle.ptr(G.updatecoeffsH),
                                                            oracle.ptr(G.ID), *(oracle.ptr(getattr(G, c)) for c in COMPS)))]
        for n, pml in enumerate(G.pmls):
            steps.append(('h_pml', n, (lambda p=pml: osolver._pml_call(lib, G, p, 1, 4))))
        steps.append(('h_src', 0, lambda: [osolver.magnetic_dipole_update(m, it, G) for m in G.magneticdipoles]))
        if G.maxpoles:
            steps.append(('e_main', 0, lambda: f['update_electric_dispersive_A'](
                G.nx, G.ny, G.nz, 4, G.maxpoles, oracle.ptr(G.updatecoeffsE), oracle.ptr(G.updatecoeffsdispersive),
                oracle.ptr(G.ID), oracle.ptr(G.Tx), oracle.ptr(G.Ty), oracle.ptr(G.Tz),
                *(oracle.ptr(getattr(G, c)) for c in COMPS))))
        else:
            steps.append(('e_main', 0, lambda: f['update_electric'](G.nx, G.ny, G.nz, 4, oracle.ptr(G.updatecoeffsE),
                                                                    oracle.ptr(G.ID), *(oracle.ptr(getattr(G, c)) for c in COMPS))))
        for n, pml in enumerate(G.pmls):
            steps.append(('e_pml', n, (lambda p=pml: osolver._pml_call(lib, G, p, 0, 4))))

        def esrc():
            for v in G.voltagesources:
                osolver.voltage_source_update(v, it, G)
            for h in G.hertziandipoles:
                osolver.hertzian_dipole_update(h, it, G)
        steps.append(('e_src', 0, esrc))
        if G.maxpoles:
            steps.append(('e_disp_b', 0, lambda: f['update_electric_dispersive_B'](
                G.nx, G.ny, G.nz, 4, G.maxpoles, oracle.ptr(G.updatecoeffsdispersive), oracle.ptr(G.ID),
                oracle.ptr(G.Tx), oracle.ptr(G.Ty), oracle.ptr(G.Tz), oracle.ptr(G.Ex), oracle.ptr(G.Ey),
                oracle.ptr(G.Ez))))
        for n, (part, arg, cpu) in enumerate(steps):
            _random_state(G, 100 + n)
            _push_state(s, G)
            s.part(part, arg, it)
            cpu()
            _assert_state_equal(s, G, '{} {}'.format(part, arg))

        # receivers: the sample at iteration it is the current state
        _random_state(G, 999)
        _push_state(s, G)
        s.part('store_rx', 0, it)
        buf = s.download_rx()
        for r, rx in enumerate(G.rxs):
            for c, name in enumerate(COMPS):
                assert buf[c, it, r] == getattr(G, name)[rx.xcoord, rx.ycoord, rx.zcoord]
        # snapshots
        for n, snap in enumerate(G.snapshots):
            s.part('snapshot', n, it)
            osolver.store_snapshot(lib, snap, G)
            out = s.download_snapshot(n)
            for c in range(6):
                assert np.array_equal(out[c], snap.fields[c]), ('snapshot', n, c)


def _compare_runs(F, rxb, snaps, phis, T, Gc):
    for n, c in enumerate(COMPS):
        assert np.array_equal(F[n], getattr(Gc, c)), c
    for r, rx in enumerate(Gc.rxs):
        for c, name in enumerate(COMPS):
            assert np.array_equal(rxb[c, :, r], rx.outputs[name]), (r, name)
    for n, snap in enumerate(Gc.snapshots):
        for c in range(6):
            assert np.array_equal(snaps[n][c], snap.fields[c]), ('snapshot', n, c)
    for n, pml in enumerate(Gc.pmls):
        for w, name in enumerate(('EPhi1', 'EPhi2', 'HPhi1', 'HPhi2')):
            assert np.array_equal(phis[n][w], getattr(pml, name)), (pml.ID, name)
    if T is not None:
        for n, name in enumerate(('Tx', 'Ty', 'Tz')):
            assert np.array_equal(T[n], getattr(Gc, name)), name


FORMS = dict(argvalues=[1, 0, 2, 6, 3, 5, 7, 4],
             ids=['fused-step', 'two-pass', 'side-by-side', 'merged', 'split', 'split-beside', 'split-merged',
                  'timed'])
CHOSEN_FORM = {0: 0, 1: 3, 2: 5, 3: 2, 4: 6, 5: 7}    # form 4's choice -> the form it runs


def _expected_form(fused, G):
    # form 4 runs the split step until its first long run has timed it
    if fused in (3, 4, 5, 7):
        return (fused if fused in (5, 7) else 3) if G.maxpoles == 0 else 0
    return fused


@pytest.mark.parametrize('fused', **FORMS)
@pytest.mark.parametrize('case', list(CASES))
def test_fused_run_matches_oracle(case, fused):
    """All iterations, as one fused kernel per step or as an H and an E pass:
    final state, receiver traces and snapshots equal to the CPU solve."""
    G = rich_model(seed=3, **CASES[case])
    Gc = rich_model(seed=3, **CASES[case])
    with _hip()(G) as s:
        s.set_step(fused)
        assert (s.step_kind() > 0) == (fused == 1)
        assert s.step_form() == _expected_form(fused, G)
        assert s.step_choice() == -1
        s.run(0, G.iterations)
        # 40 iterations: form 4 timed six forms on iterations 2-25 (three,
        # the passes alone, side by side and merged, on iterations 2-13 of
        # dispersive models) and went on with the fastest, switching forms in
        # the middle of the run
        if fused == 4:
            assert s.step_choice() in ((0, 3, 4) if G.maxpoles else (0, 1, 2, 3, 4, 5))
            assert s.step_form() == CHOSEN_FORM[s.step_choice()]
        else:
            assert s.step_choice() == -1
        F = s.download_fields()
        rxb = s.download_rx()
        snaps = [s.download_snapshot(n) for n in range(len(G.snapshots))]
        phis = [[s.download_phi(n, w) for w in range(4)] for n in range(len(G.pmls))]
        T = s.download_T() if G.maxpoles else None
    osolver.solve_cpu(1, 1, Gc, nthreads=8)
    _compare_runs(F, rxb, snaps, phis, T, Gc)


@pytest.mark.parametrize('form', [3, 5, 7], ids=['split', 'split-beside', 'split-merged'])
@pytest.mark.parametrize('xchunk', [0, 1, 2, 5])
@pytest.mark.parametrize('case', ['horipml1', 'mripml2'])
def test_split_step_from_random_state(case, xchunk, form):
    """The split step (its boundary passes after the interior kernel, or
    the H boundary pass beside it) from a random state (fields and Phi) on a
    grid with several blocks per plane and short chunks: state and receiver
    samples equal to the oracle stepped the same number of iterations."""
    G = rich_model(seed=22, shape=WIDE, iterations=12, snapshots=False, **CASES[case])
    _random_state(G, 6)
    lib = oracle.load(np.float32)
    with _hip()(G) as s:
        s.set_step(form)
        s.set_launch(xchunk, 0)
        assert s.step_form() == form
        _push_state(s, G)
        s.run(2, 6)
        for it in range(2, 8):
            osolver.step(lib, G, it, 8)
        _assert_state_equal(s, G, 'split form {} xchunk {}'.format(form, xchunk))
        buf = s.download_rx()
    for r, rx in enumerate(G.rxs):
        for c, name in enumerate(COMPS):
            assert np.array_equal(buf[c, 2:8, r], rx.outputs[name][2:8]), (r, name)


@pytest.mark.parametrize('fused', **FORMS)
@pytest.mark.parametrize('case', list(CASES))
def test_run_from_random_state_anisotropic_random_tables(case, fused):
    """Every step form from a random state (fields, Phi, T) on 1 x 1.5 x 2 mm
    cells with every coefficient and R table entry scaled by its own random
    factor, on a grid with several blocks per plane: state and receiver
    samples equal to the oracle's (and finite).  Forms the model cannot take
    run the passes, as gpx_run does."""
    G = randomise_tables(rich_model(seed=27, shape=WIDE, iterations=12, snapshots=False, cell=ANISO,
                                    **CASES[case]), 29, scaled=True)
    _random_state(G, 31)
    assert all(np.isfinite(getattr(G, c)).all() for c in COMPS)
    lib = oracle.load(np.float32)
    with _hip()(G) as s:
        s.set_step(fused)
        s.set_launch(3, 0)
        _push_state(s, G)
        s.run(2, 8)
        for it in range(2, 10):
            osolver.step(lib, G, it, 8)
        assert all(np.isfinite(getattr(G, c)).all() for c in COMPS), 'the oracle run overflowed'
        _assert_state_equal(s, G, 'form {} from a random state'.format(fused))
        buf = s.download_rx()
    for r, rx in enumerate(G.rxs):
        for c, name in enumerate(COMPS):
            assert np.array_equal(buf[c, 2:10, r], rx.outputs[name][2:10]), (r, name)


# Per-face PML thicknesses, some faces open (#pml_cells x0 y0 z0 xmax ymax zmax).
FACES = {
    'uneven': {'x0': 3, 'y0': 5, 'z0': 2, 'xmax': 6, 'ymax': 4, 'zmax': 7},
    'open': {'x0': 0, 'y0': 4, 'z0': 5, 'xmax': 3, 'ymax': 0, 'zmax': 0},
    'none': {},
}


@pytest.mark.parametrize('fused', **FORMS)
@pytest.mark.parametrize('faces', list(FACES))
@pytest.mark.parametrize('case', ['horipml2', 'mripml1', 'debye3'])
def test_run_from_random_state_uneven_pml(case, faces, fused):
    """Every step form from a random state on a grid whose PML has a
    different thickness on every face, some faces open, or no PML at all:
    the interior box, the boundary frame and the merged block order follow
    the slabs, and state and receiver samples equal the oracle's."""
    G = rich_model(seed=41, shape=WIDE, iterations=12, snapshots=False, pml=FACES[faces], **CASES[case])
    assert len(G.pmls) == sum(1 for v in FACES[faces].values() if v > 0)
    _random_state(G, 43)
    lib = oracle.load(np.float32)
    with _hip()(G) as s:
        s.set_step(fused)
        _push_state(s, G)
        s.run(2, 8)
        for it in range(2, 10):
            osolver.step(lib, G, it, 8)
        _assert_state_equal(s, G, 'form {} faces {}'.format(fused, faces))
        buf = s.download_rx()
    for r, rx in enumerate(G.rxs):
        for c, name in enumerate(COMPS):
            assert np.array_equal(buf[c, 2:10, r], rx.outputs[name][2:10]), (r, name)


PLANE_CASES = {
    'horipml1': dict(order=1),
    'mripml2': dict(order=2, formulation='MRIPML'),
    'debye2': dict(poles=2, kind='debye'),
    'lorentz1': dict(poles=1, kind='lorentz', order=2),
}


@pytest.mark.parametrize('fused', **FORMS)
@pytest.mark.parametrize('axis', [0, 1, 2], ids=['TMx', 'TMy', 'TMz'])
@pytest.mark.parametrize('case', list(PLANE_CASES))
def test_2d_run_matches_oracle(case, axis, fused):
    """2-D models (one cell along `axis`): the GPU runs the 3-D update ranges,
    the oracle the reference's 2-D branches; final state, Phi, T, receiver
    traces and snapshots are the same bits."""
    G = plane_model(seed=13 + axis, axis=axis, **PLANE_CASES[case])
    Gc = plane_model(seed=13 + axis, axis=axis, **PLANE_CASES[case])
    with _hip()(G) as s:
        s.set_step(fused)
        assert (s.step_kind() > 0) == (fused == 1)
        assert s.step_form() == (0 if fused in (3, 4, 5, 7) else fused)   # a 2-D grid has no fused box
        s.run(0, G.iterations)
        F = s.download_fields()
        rxb = s.download_rx()
        snaps = [s.download_snapshot(n) for n in range(len(G.snapshots))]
        phis = [[s.download_phi(n, w) for w in range(4)] for n in range(len(G.pmls))]
        T = s.download_T() if G.maxpoles else None
    osolver.solve_cpu(1, 1, Gc, nthreads=8)
    assert np.abs(np.stack([rx.outputs[COMPS[axis]] for rx in Gc.rxs])).max() > 0
    _compare_runs(F, rxb, snaps, phis, T, Gc)


@pytest.mark.parametrize('name', list(PLANE_MODELS))
def test_reference_2d_models_match_oracle(name):
    """The reference's 2-D test models through solve_hip: receiver traces equal
    to the oracle's (which the CPU tests pin to the reference's files)."""
    from gprmax_amd.solver import solve_hip
    G = PLANE_MODELS[name]()
    solve_hip(1, 1, G)
    Gc = PLANE_MODELS[name]()
    osolver.solve_cpu(1, 1, Gc, nthreads=8)
    for c in COMPS:
        assert np.array_equal(G.rxs[0].outputs[c], Gc.rxs[0].outputs[c]), c


@pytest.mark.parametrize('xchunk,block', [(1, 64), (7, 128), (0, 256), (300, 512), (5, 1024)])
def test_launch_shapes_do_not_change_results(xchunk, block):
    G = rich_model(seed=5, poles=2, order=2)
    ref = None
    with _hip()(G) as s:
        s.set_step(0)
        s.set_launch(xchunk, block)
        s.run(0, G.iterations)
        out = s.download_fields()
    with _hip()(G) as s:
        s.run(0, G.iterations)
        ref = s.download_fields()
    assert np.array_equal(out, ref)


# A cross-section of 46 x 38 = 1748 cells spans two to three tiles of the fused
# step, so tile halos and x-chunk halo planes are crossed everywhere.
WIDE = (23, 45, 37)


@pytest.mark.parametrize('tile,xchunk', [(1024, 3), (512, 2), (1024, 0), (256, 5)])
@pytest.mark.parametrize('case', ['horipml1', 'mripml2', 'debye3', 'lorentz2'])
def test_fused_step_from_random_state(case, tile, xchunk):
    """Fused iterations from a random state (fields, Phi, T) with several tiles
    per plane and short x chunks: state and receiver samples equal to the
    oracle stepped the same number of iterations."""
    G = rich_model(seed=21, shape=WIDE, iterations=12, snapshots=False, **CASES[case])
    _random_state(G, 5)
    lib = oracle.load(np.float32)
    with _hip()(G) as s:
        s.set_step(1, tile)
        s.set_launch(xchunk, 0)
        assert s.step_kind() == tile
        _push_state(s, G)
        s.run(2, 6)
        for it in range(2, 8):
            osolver.step(lib, G, it, 8)
        _assert_state_equal(s, G, 'fused tile {} xchunk {}'.format(tile, xchunk))
        buf = s.download_rx()
    for r, rx in enumerate(G.rxs):
        for c, name in enumerate(COMPS):
            assert np.array_equal(buf[c, 2:8, r], rx.outputs[name][2:8]), (r, name)


@pytest.mark.parametrize('poles', [1, 0])
def test_step_forms_agree_on_a_wide_grid(poles):
    """All step forms give the same bits on a grid with several tiles and
    several blocks per plane."""
    out = []
    for fused in (1, 0, 2, 6, 3, 7, 4):
        G = rich_model(seed=8, shape=WIDE, iterations=30, poles=poles, order=2)
        with _hip()(G) as s:
            s.set_step(fused)
            s.run(0, G.iterations)
            out.append((s.download_fields(), s.download_rx()))
    for F, rx in out[1:]:
        assert np.array_equal(out[0][0], F)
        assert np.array_equal(out[0][1], rx)


@pytest.mark.parametrize('name', ['hertzian_dipole_fs', 'hertzian_dipole_dispersive', 'magnetic_dipole_fs'])
def test_reference_models_match_oracle_and_golden(name):
    """The reference's 100^3 test models, 1559 iterations, through solve_hip."""
    from gprmax_amd.solver import solve_hip
    G = BASIC_MODELS[name]()
    tsolve, mem = solve_hip(1, 1, G)
    assert tsolve > 0 and mem > 0
    ours = np.stack([G.rxs[0].outputs[c] for c in COMPS])
    Gc = BASIC_MODELS[name]()
    osolver.solve_cpu(1, 1, Gc, nthreads=8)
    cpu = np.stack([Gc.rxs[0].outputs[c] for c in COMPS])
    assert np.array_equal(ours, cpu)
    ref = np.load('tests/golden/{}_ref.npz'.format(name))['traces']
    if name == 'hertzian_dipole_dispersive':
        for c in range(5):
            shift = 1 if c >= 3 else 0
            a, b = (ours[c][shift:], ref[c][:len(ref[c]) - shift])
            assert np.abs(a - b).max() / np.abs(b).max() < 1e-6


def test_config1_freespace_ricker_within_tolerance():
    """100^3 free space, 10-cell PML, z-dipole with a 1.5 GHz Ricker, Ez receiver
    20 cells along x, 500 iterations: |GPU - CPU| <= 1e-4 of the peak |Ez| (the
    stated tolerance; the expected result is exact equality)."""
    from gprmax_amd.workloads import freespace_ricker
    G = freespace_ricker()
    Gc = freespace_ricker()
    from gprmax_amd.solver import solve_hip
    solve_hip(1, 1, G)
    osolver.solve_cpu(1, 1, Gc, nthreads=8)
    a, b = G.rxs[0].outputs['Ez'], Gc.rxs[0].outputs['Ez']
    peak = np.abs(b).max()
    assert peak > 0
    assert np.abs(a - b).max() <= 1e-4 * peak
    assert np.array_equal(a, b)


def test_hip_rejects_bad_input():
    from gprmax_amd.utilities import GeneralError
    G = rich_model(seed=1, iterations=4, snapshots=False)
    G.ID[0, 0, 0, 0] = len(G.updatecoeffsE) + 5
    with pytest.raises(GeneralError, match='material index'):
        _hip()(G)


def _single_runs(models, n):
    out = []
    for G in models:
        with _hip()(G) as s:
            s.run(0, n)
            out.append((s.download_fields(), s.download_rx()))
    return out


def _batch_models(iterations):
    shapes = [WIDE, (26, 22, 19), (30, 17, 25)]
    return [rich_model(seed=30 + n, shape=sh, iterations=iterations, snapshots=False, poles=1, order=1)
            for n, sh in enumerate(shapes)]


@pytest.mark.parametrize('tile', [0, 1024, 512], ids=['passes', 'fused-1024', 'fused-512'])
def test_batch_matches_single_runs(tile):
    """Three models of different sizes advanced by shared launches (the passes
    over all models, or the fused step over all models), first eagerly and
    then as one captured HIP graph: every model ends in the state and with the
    receiver traces of its own run."""
    from gprmax_amd.hip import HipBatch
    models = _batch_models(20)
    ref = _single_runs(models, 20)
    solvers = [_hip()(G) for G in models]
    try:
        with HipBatch(solvers) as b:
            b.set_shape(tile, 0)
            b.run(0, 7, graph=False)
            b.run(7, 13, graph=True)
        for s, (F, rx) in zip(solvers, ref):
            assert np.array_equal(s.download_fields(), F)
            assert np.array_equal(s.download_rx(), rx)
    finally:
        for s in solvers:
            s.close()


@pytest.mark.parametrize('group_bytes', [0, 1], ids=['one-group', 'each-alone'])
def test_bscan_solve_matches_single_runs(group_bytes):
    """The pipelined B-scan solve (groups built on a second host thread while
    the group before runs): the three models in one group (shared launches)
    or each alone (plain solves with their own step forms, timed on the
    30-iteration run) give every model the receiver traces of its own run."""
    from gprmax_amd.hip import solve_bscan
    models = _batch_models(30)
    ref = _single_runs(models, 30)
    rx, ms = solve_bscan(models, group_bytes=group_bytes, tune_steps=2)
    assert ms > 0 and len(rx) == len(models)
    for got, (_, want) in zip(rx, ref):
        assert np.array_equal(got, want)


def test_batch_autotuner_picks_a_shape_and_resets_the_state():
    from gprmax_amd.hip import HipBatch
    models = _batch_models(16)
    ref = _single_runs(models, 16)
    solvers = [_hip()(G) for G in models]
    try:
        with HipBatch(solvers) as b:
            tile, xchunk, ms = b.tune(steps=2)
            assert (tile == 0 or (tile in (256, 512, 1024) and xchunk >= 2)) and ms > 0
            for s in solvers:
                assert not s.download_fields().any() and not s.download_rx().any()
            b.run(0, 16, graph=True)
        for s, (F, rx) in zip(solvers, ref):
            assert np.array_equal(s.download_fields(), F)
            assert np.array_equal(s.download_rx(), rx)
    finally:
        for s in solvers:
            s.close()


def test_pass_batch_of_unlike_models():
    """The passes batch needs only one kernel: 3-D and 2-D models, different
    PML orders and slab sets, different receiver counts, advanced together."""
    from gprmax_amd.hip import HipBatch
    models = [rich_model(seed=40, shape=(24, 20, 18), iterations=18, snapshots=False, poles=1, order=2),
              plane_model(seed=41, axis=2, iterations=18, snapshots=False, poles=1, order=1),
              rich_model(seed=42, shape=(21, 26, 17), iterations=18, snapshots=False, poles=1, pml=0),
              plane_model(seed=43, axis=0, iterations=18, snapshots=False, poles=1, order=2)]
    ref = _single_runs(models, 18)
    for groups in ('one', 'each', 'pairs'):
        solvers = [_hip()(G) for G in models]
        # a model's grouped state: its bytes (a model holds a second field
        # state only once a double-buffered form has run)
        state = [s.memory_bytes() for s in solvers]
        group_bytes = {'one': 0, 'each': 1, 'pairs': max(state[0] + state[1], state[2] + state[3]) + 4096}[groups]
        try:
            with HipBatch(solvers) as b:
                b.set_group_bytes(group_bytes)
                b.run(0, 9, graph=False)
                b.run(9, 9, graph=True)
            for s, (F, rx) in zip(solvers, ref):
                assert np.array_equal(s.download_fields(), F), groups
                assert np.array_equal(s.download_rx(), rx), groups
        finally:
            for s in solvers:
                s.close()



def test_default_form_times_once_and_resets():
    """The default form times six forms on the first run of at least 26
    iterations only: short runs and profiled runs leave
    the choice open, gpx_set_step clears it, and the results stay the
    oracle's."""
    G = rich_model(seed=31, shape=WIDE, iterations=60, snapshots=False)
    Gc = rich_model(seed=31, shape=WIDE, iterations=60, snapshots=False)
    with _hip()(G) as s:
        assert s.step_choice() == -1
        s.run(0, 10)                       # too short to time
        assert s.step_choice() == -1
        s.run_profiled(10, 5)              # profiled runs never time
        assert s.step_choice() == -1
        s.run(15, 30)                      # times iterations 17-40, then keeps the fastest
        assert s.step_choice() in (0, 1, 2, 3, 4, 5)
        s.set_step(4)
        assert s.step_choice() == -1
        s.run(45, 15)
        F = s.download_fields()
        rxb = s.download_rx()
    osolver.solve_cpu(1, 1, Gc, nthreads=8)
    for n, c in enumerate(COMPS):
        assert np.array_equal(F[n], getattr(Gc, c)), c
    for r, rx in enumerate(Gc.rxs):
        for c, name in enumerate(COMPS):
            assert np.array_equal(rxb[c, :, r], rx.outputs[name]), (r, name)


@pytest.mark.parametrize('form', [3, 5, 7], ids=['split', 'split-beside', 'split-merged'])
@pytest.mark.parametrize('xchunk', [0, 3])
def test_split_step_sources_on_recomputed_cells(xchunk, form):
    """Magnetic dipoles on the cells whose new H the split step recomputes
    in-lane (the row before every block, the column before the box, the plane
    before every chunk) and on the box's own first row, plane and column: the
    recomputed values carry the dipole, as the owner's do."""
    from gprmax_amd.grid import ModelBuilder
    def model():
        nx, ny, nz = WIDE
        bld = ModelBuilder(nx, ny, nz, 1e-3, 1e-3, 1e-3, 14, pml_cells=4)
        soil = bld.add_material(4.0, 0.005, 1, 0, 'soil')
        bld.box(0, nx, 0, ny, 0, nz // 2, soil)
        bld.add_waveform('ricker', 1, 20e9, 'w')
        # box of the passes: cells 5..n-5; the split step's box starts one later
        i0, j0, k0 = 5, 5, 5
        for j in range(j0, ny - 5):                      # every row: block starts fall on some
            bld.add_magnetic_dipole('x', (i0 + 3, j, k0 + 2), 'w')
            bld.add_magnetic_dipole('z', (i0 + 6, j, k0 + 7), 'w')
        for i in range(i0, nx - 5):                      # every plane: chunk starts fall on some
            bld.add_magnetic_dipole('y', (i, j0 + 4, k0 + 3), 'w')
            bld.add_magnetic_dipole('z', (i, j0 + 9, k0 + 1), 'w')
        for k in range(k0, nz - 5):                      # the column before the box and the box's first
            bld.add_magnetic_dipole('x', (i0 + 2, j0, k), 'w')
            bld.add_magnetic_dipole('y', (i0 + 4, j0 + 1, k0), 'w')
        bld.add_magnetic_dipole('y', (i0 + 5, j0 + 6, k0), 'w')
        bld.add_hertzian_dipole('z', (nx // 2, ny // 2, nz // 2), 'w')
        for r in range(6):
            bld.add_rx((i0 + 1 + r, j0 + 2 * r, k0 + 1 + r))
        return bld.finalise()
    G, Gc = model(), model()
    with _hip()(G) as s:
        s.set_step(form)
        s.set_launch(xchunk, 0)
        assert s.step_form() == form
        box = s.interior_box()
        assert box[0] == 5 and box[2] == 5 and box[4] == 5
        s.run(0, G.iterations)
        F = s.download_fields()
        rxb = s.download_rx()
    osolver.solve_cpu(1, 1, Gc, nthreads=8)
    for n, c in enumerate(COMPS):
        assert np.array_equal(F[n], getattr(Gc, c)), c
    for r, rx in enumerate(Gc.rxs):
        for c, name in enumerate(COMPS):
            assert np.array_equal(rxb[c, :, r], rx.outputs[name]), (r, name)
