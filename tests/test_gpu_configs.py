"""The benchmark configurations at their stated sizes on the GPU, against the
CPU oracle, through the C ABI.

* 300x300x200 heterogeneous half-space (256 soil materials, PEC cylinder,
  Hz magnetic dipole, nine receivers): from the zero state with the default
  step form, long enough for it to time the forms and switch, and from a
  random state;
* 400x400x300 dispersive soil (3-pole Debye on 30 % of the cells, 20-cell
  PML): from a random state of fields, Phi and T;
* the uneven B-scan: five traces of 128-256 cells per side advanced by the
  batched launches (autotuned shape, HIP graph), each trace against the oracle;
* the differentiable solve: GPU gradient against a float64 PyTorch autograd
  reference at 40^3, and against central differences at 32 random cells of
  the 160x160x120 model with K = 64 over 2000 iterations.

Field runs are compared bit for bit (both sides compute in float32 without
contraction, in the reference's association order), which is stricter than
the stated tolerance of 1e-4 of the peak |E|.  Iteration counts are cut to
tens so that the oracle finishes in seconds; the size-independent parts of
the check (every cell, every slab, every receiver) are whole.
"""

import numpy as np
import pytest

import oracle
from oracle import solver as osolver

pytestmark = pytest.mark.gpu

COMPS = ('Ex', 'Ey', 'Ez', 'Hx', 'Hy', 'Hz')


def _threads():
    return osolver.default_threads()


def _random_state(G, seed):
    """Random fields and Phi (and T on the edges with poles) on the oracle
    grid, generated in float32."""
    rng = np.random.default_rng(seed)
    osolver.initialise_state(G)
    for c in COMPS:
        a = getattr(G, c)
        a[...] = rng.standard_normal(a.shape, dtype=np.float32)
    for pml in G.pmls:
        for name in ('EPhi1', 'EPhi2', 'HPhi1', 'HPhi2'):
            a = getattr(pml, name)
            a[...] = rng.standard_normal(a.shape, dtype=np.float32)
    if G.maxpoles:
        assert np.all(G.updatecoeffsdispersive.imag == 0)     # Debye: real T
        has_poles = np.any(G.updatecoeffsdispersive != 0, axis=1)
        for c, name in enumerate(('Tx', 'Ty', 'Tz')):
            a = getattr(G, name)
            for p in range(G.maxpoles):
                v = rng.standard_normal(a.shape[1:], dtype=np.float32)
                v[~has_poles[G.ID[c]]] = 0
                a[p] = v


def _push_state(s, G):
    s.upload_fields(np.stack([getattr(G, c) for c in COMPS]))
    for n, pml in enumerate(G.pmls):
        for w, name in enumerate(('EPhi1', 'EPhi2', 'HPhi1', 'HPhi2')):
            s.upload_phi(n, w, getattr(pml, name))
    if G.maxpoles:
        s.upload_T(np.stack([G.Tx, G.Ty, G.Tz]))


def _assert_state_equal(s, G, what, T=True):
    F = s.download_fields()
    for n, c in enumerate(COMPS):
        assert np.array_equal(F[n], getattr(G, c)), (what, c, float(np.abs(F[n] - getattr(G, c)).max()))
    del F
    for n, pml in enumerate(G.pmls):
        for w, name in enumerate(('EPhi1', 'EPhi2', 'HPhi1', 'HPhi2')):
            assert np.array_equal(s.download_phi(n, w), getattr(pml, name)), (what, pml.ID, name)
    if G.maxpoles and T:
        Tg = s.download_T()
        for n, name in enumerate(('Tx', 'Ty', 'Tz')):
            assert np.array_equal(Tg[n], getattr(G, name)), (what, name)


def _assert_rx_equal(buf, G, its):
    for r, rx in enumerate(G.rxs):
        for c, name in enumerate(COMPS):
            assert np.array_equal(buf[c, its, r], rx.outputs[name][its]), (r, name)


def test_config2_halfspace_full_size_default_form_from_zero():
    """300x300x200, 40 iterations from the zero state with the default form:
    it times six forms on iterations 2-25 and keeps the fastest; final
    fields, Phi and all nine receiver traces equal the oracle's."""
    from gprmax_amd.hip import HipSolver
    from gprmax_amd.workloads import heterogeneous_halfspace
    n = 40
    G = heterogeneous_halfspace(iterations=n)
    assert (G.nx, G.ny, G.nz) == (300, 300, 200) and len(G.rxs) == 9 and len(G.materials) == 258
    with HipSolver(G) as s:
        assert s.step_choice() == -1
        s.run(0, n)
        assert s.step_choice() in (0, 1, 2, 3, 4, 5)
        print('config 2: form kept', s.step_form(), flush=True)
        F = s.download_fields()
        rxb = s.download_rx()
        phis = [[s.download_phi(q, w) for w in range(4)] for q in range(len(G.pmls))]
    Gc = heterogeneous_halfspace(iterations=n)
    osolver.solve_cpu(1, 1, Gc, nthreads=_threads())
    for q, c in enumerate(COMPS):
        assert np.array_equal(F[q], getattr(Gc, c)), c
    for q, pml in enumerate(Gc.pmls):
        for w, name in enumerate(('EPhi1', 'EPhi2', 'HPhi1', 'HPhi2')):
            assert np.array_equal(phis[q][w], getattr(pml, name)), (pml.ID, name)
    _assert_rx_equal(rxb, Gc, slice(0, n))
    assert np.abs(rxb[5]).max() > 0          # the Hz dipole reached its receivers


@pytest.mark.parametrize('form', [4, 2, 6, 3, 5, 7, 0, 1], ids=['default', 'side-by-side', 'merged', 'split',
                                                             'split-beside', 'split-merged', 'two-pass',
                                                             'fused-step'])
def test_config2_halfspace_full_size_from_random_state(form):
    """300x300x200 from a random state of every field and Phi value: 30
    iterations (the default form times the forms on 24 of them), state and
    receiver samples equal to the oracle's."""
    from gprmax_amd.hip import HipSolver
    from gprmax_amd.workloads import heterogeneous_halfspace
    G = heterogeneous_halfspace(iterations=40)
    _random_state(G, 2002)
    lib = oracle.load(np.float32)
    its = range(2, 32)
    with HipSolver(G) as s:
        s.set_step(form)
        _push_state(s, G)
        s.run(its.start, len(its))
        if form == 4:
            assert s.step_choice() in (0, 1, 2, 3, 4, 5)
        for it in its:
            osolver.step(lib, G, it, _threads())
        _assert_state_equal(s, G, 'config 2 form {}'.format(form))
        _assert_rx_equal(s.download_rx(), G, slice(its.start, its.stop))


def test_config3_dispersive_soil_full_size_from_random_state():
    """400x400x300, 20-cell PML, 3-pole Debye on 30 % of the cells: 30
    iterations from a random state of fields, Phi and T with the default form
    (which times the passes alone, side by side and merged); fields, Phi, T and the
    receiver samples equal to the oracle's."""
    from gprmax_amd.hip import HipSolver
    from gprmax_amd.workloads import dispersive_soil
    G = dispersive_soil(iterations=40)
    assert (G.nx, G.ny, G.nz) == (400, 400, 300) and G.maxpoles == 3 and G.pmls[0].thickness == 20
    _random_state(G, 2003)
    lib = oracle.load(np.float32)
    its = range(2, 32)
    with HipSolver(G) as s:
        _push_state(s, G)
        s.run(its.start, len(its))
        assert s.step_choice() in (0, 3, 4)
        print('config 3: form kept', s.step_form(), flush=True)
        for it in its:
            osolver.step(lib, G, it, _threads())
        _assert_state_equal(s, G, 'config 3')
        _assert_rx_equal(s.download_rx(), G, slice(its.start, its.stop))


def test_config3_dispersive_soil_full_size_from_zero():
    """400x400x300 from the zero state, 40 iterations: the nine receiver traces
    and the final fields equal the oracle's."""
    from gprmax_amd.hip import HipSolver
    from gprmax_amd.workloads import dispersive_soil
    n = 40
    G = dispersive_soil(iterations=n)
    with HipSolver(G) as s:
        s.run(0, n)
        F = s.download_fields()
        rxb = s.download_rx()
    del G
    Gc = dispersive_soil(iterations=n)
    osolver.solve_cpu(1, 1, Gc, nthreads=_threads())
    for q, c in enumerate(COMPS):
        assert np.array_equal(F[q], getattr(Gc, c)), c
    _assert_rx_equal(rxb, Gc, slice(0, n))
    assert np.abs(F[2]).max() > 0


@pytest.mark.parametrize('tile', ['tuned', 0, 1024])
def test_config4_bscan_traces_batched_match_oracle(tile):
    """Five uneven B-scan traces (sides 128-256 cells, 1 mm, 10-cell PML, random
    soil) resident together and advanced by the batched launches in one HIP
    graph, with the shape the autotuner picks or a fixed one (the passes over
    all traces, the fused step), each from its own random state: every trace
    ends in the oracle's state with the oracle's receiver samples."""
    from gprmax_amd.hip import HipBatch, HipSolver
    from gprmax_amd.workloads import bscan_trace
    ntr, its = 5, range(3, 23)
    models = [bscan_trace(t, iterations=40) for t in range(ntr)]
    shapes = [(G.nx, G.ny, G.nz) for G in models]
    assert all(128 <= v <= 256 for sh in shapes for v in sh) and len(set(shapes)) == ntr
    solvers = [HipSolver(G) for G in models]
    lib = oracle.load(np.float32)
    try:
        with HipBatch(solvers) as b:
            if tile == 'tuned':
                t, xc, ms = b.tune(steps=2)           # leaves every model at the zero state
                print('config 4: tuned tile', t, 'xchunk', xc, flush=True)
            else:
                b.set_shape(tile, 0)
            for q, (s, G) in enumerate(zip(solvers, models)):
                _random_state(G, 2004 + q)
                _push_state(s, G)
            b.run(its.start, len(its), graph=True)
        for q, (s, G) in enumerate(zip(solvers, models)):
            for it in its:
                osolver.step(lib, G, it, _threads())
            _assert_state_equal(s, G, 'trace {}'.format(q))
            _assert_rx_equal(s.download_rx(), G, slice(its.start, its.stop))
    finally:
        for s in solvers:
            s.close()


def _adjoint_setup(shape, iterations, K, seed):
    from gprmax_amd.inversion import InversionGeometry
    rng = np.random.default_rng(seed)
    geo = InversionGeometry(shape, 2e-3, iterations, pml=10, er_pml=7.0, K=K)
    waves = geo.waveform(freq=900e6)
    return rng, geo, waves


@pytest.mark.parametrize('poles', [0, 3])
def test_config5_adjoint_40_cube_matches_float64_autograd(poles):
    """40^3 cells of 2 mm, 10-cell PML, random true model (er in [4, 10],
    sigma in [1e-3, 2e-2]) and a perturbed homogeneous start model, two
    sources and six receivers, 150 iterations with K = 64 (two whole segments
    and a partial one), with no poles or with 3 Debye poles (deltaer ~
    U[0.5, 5], tau = 10^U[-11, -9] s) on 40 % of the cells: the GPU gradient
    equals the float64 PyTorch autograd gradient within 1e-3 (relative L2
    over the cells outside the PML)."""
    import torch
    from gprmax_amd.inversion import InversionGrid, misfit_and_gradient
    from torch_fdtd import forward_traces, misfit
    shape = (40, 40, 40)
    rng, geo, waves = _adjoint_setup(shape, 150, 64, 2005)
    er_t = 4 + 6 * rng.random(shape)
    se_t = 1e-3 + 1.9e-2 * rng.random(shape)
    er0 = 7.0 + 0.3 * rng.standard_normal(shape)
    se0 = 1.05e-2 + 1e-3 * rng.standard_normal(shape)
    if poles:
        mask = rng.random(shape) < 0.4
        de = rng.uniform(0.5, 5, (poles,) + shape) * mask
        tau = 10 ** rng.uniform(-11, -9, (poles,) + shape)
        geo.set_debye(de, tau)
    src = [(15, 20, 25), (24, 18, 25)]
    rx = [(12 + 4 * n, 14 + 2 * n, 25) for n in range(6)]
    with torch.no_grad():
        obs = [forward_traces(torch.tensor(er_t), torch.tensor(se_t), geo, s, rx, waves).numpy() for s in src]
    er = torch.tensor(er0, requires_grad=True)
    se = torch.tensor(se0, requires_grad=True)
    L_ref = misfit(er, se, geo, src, rx, obs, waves)
    L_ref.backward()
    grid = InversionGrid(shape, 2e-3, geo.iterations, pml=10, er_pml=7.0, K=64)
    try:
        if poles:
            grid.set_debye(de, tau)
        L, ger, gse = misfit_and_gradient(grid, er0, se0, src, rx, obs, waves)
    finally:
        grid.close()
    m = geo.interior
    assert abs(L - L_ref.item()) <= 1e-4 * L_ref.item()
    for ours, ref in ((ger, er.grad.numpy()), (gse, se.grad.numpy())):
        err = np.linalg.norm(ours[m] - ref[m]) / np.linalg.norm(ref[m])
        print('config 5 (40^3, {} poles): relative L2 gradient error'.format(poles), err, flush=True)
        assert err < 1e-3, err


@pytest.mark.slow
def test_config5_adjoint_full_size_matches_central_differences():
    """160x160x120 cells of 2 mm, 10-cell PML, smooth true model, homogeneous
    start model (er 7, sigma 1.05e-2), one source and 16 Ez receivers 5 cells
    inside the top PML, K = 64 over 2000 iterations: dL/d(er) at 32 random
    cells outside the PML against central differences of the float32 forward
    (step 1e-2 of er): within 1e-2 relative, with an absolute floor of 1e-3 of
    the largest gradient for cells the traces hardly see."""
    from gprmax_amd.inversion import InversionGrid, misfit_and_gradient
    shape, its = (160, 160, 120), 2000
    rng = np.random.default_rng(1005)

    def smooth(lo, hi):
        f = rng.standard_normal(shape)
        for ax in range(3):
            for _ in range(3):
                f = sum(np.roll(f, s, axis=ax) for s in range(-4, 5)) / 9.0
        f = (f - f.min()) / (f.max() - f.min())
        return lo + (hi - lo) * f

    er_t, se_t = smooth(4.0, 10.0), smooth(1e-3, 2e-2)
    er0, se0 = np.full(shape, 7.0), np.full(shape, 1.05e-2)
    z = 120 - 10 - 5
    src = [(40, 55, z)]
    rx = [(30 + 33 * (n % 4), 30 + 33 * (n // 4), z) for n in range(16)]
    grid = InversionGrid(shape, 2e-3, its, pml=10, er_pml=7.0, K=64)
    try:
        waves = grid.waveform(freq=900e6)
        grid.set_model(er_t, se_t)
        obs = [grid.forward(s, rx, waves).astype(np.float64) for s in src]
        _, g, _ = misfit_and_gradient(grid, er0, se0, src, rx, obs, waves)
        idx = np.argwhere(grid.interior)
        picks = idx[rng.choice(len(idx), 32, replace=False)]
        gmax = np.abs(g).max()
        errs = []
        for c in picks:
            c = tuple(int(v) for v in c)
            h = 1e-2 * er0[c]
            L = []
            for sgn in (1, -1):
                e = er0.copy()
                e[c] += sgn * h
                grid.set_model(e, se0)
                r = grid.forward(src[0], rx, waves).astype(np.float64) - obs[0]
                L.append(0.5 * float(np.sum(r * r)))
            fd = (L[0] - L[1]) / (2 * h)
            errs.append((c, g[c], fd))
        for c, gc, fd in errs:
            print('config 5 FD', c, gc, fd, flush=True)
        for c, gc, fd in errs:
            assert abs(gc - fd) <= 1e-2 * abs(fd) + 1e-3 * gmax, (c, gc, fd)
    finally:
        grid.close()
