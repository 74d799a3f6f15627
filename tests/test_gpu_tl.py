"""Transmission lines on the GPU (k_tl_h, k_tl_e) against the oracle,
through the C ABI: fields, PML Phi, receivers, the lines' Vtotal/Itotal
and their final state, bit for bit, in every step form (models with lines
step in the two-pass form)."""

import numpy as np
import pytest

from oracle import solver as osolver
from tl_models import tl_model

pytestmark = pytest.mark.gpu

COMPS = ('Ex', 'Ey', 'Ez', 'Hx', 'Hy', 'Hz')


def _hip():
    from gprmax_amd.hip import HipSolver
    return HipSolver


@pytest.mark.parametrize('fused', [0, 1, 2, 6, 4], ids=['two-pass', 'fused-requested', 'side-by-side', 'merged',
                                                   'default'])
@pytest.mark.parametrize('case', [('HORIPML', 1), ('MRIPML', 2)], ids=['horipml1', 'mripml2'])
def test_lines_match_oracle(case, fused):
    form, order = case
    G = tl_model(iterations=100, order=order, formulation=form)
    Gc = tl_model(iterations=100, order=order, formulation=form)
    with _hip()(G) as s:
        s.set_step(fused)
        assert s.step_form() == (fused if fused in (2, 6) else 0)
        s.run(0, G.iterations)
        if fused == 4:   # timed the passes alone, side by side and merged, kept the fastest
            assert s.step_choice() in (0, 3, 4)
        F = s.download_fields()
        rxb = s.download_rx()
        phis = [[s.download_phi(n, w) for w in range(4)] for n in range(len(G.pmls))]
        lines = [s.download_transmission_line(n) for n in range(len(G.transmissionlines))]
    osolver.solve_cpu(1, 1, Gc, nthreads=8)
    for n, c in enumerate(COMPS):
        assert np.array_equal(F[n], getattr(Gc, c)), c
    for r, rx in enumerate(Gc.rxs):
        for c, name in enumerate(COMPS):
            assert np.array_equal(rxb[c, :, r], rx.outputs[name]), (r, name)
    for n, pml in enumerate(Gc.pmls):
        for w, name in enumerate(('EPhi1', 'EPhi2', 'HPhi1', 'HPhi2')):
            assert np.array_equal(phis[n][w], getattr(pml, name)), (pml.ID, name)
    for (vt, it_, v, i), tl in zip(lines, Gc.transmissionlines):
        assert np.any(vt != 0)
        assert np.array_equal(vt, tl.Vtotal) and np.array_equal(it_, tl.Itotal), tl.ID
        assert np.array_equal(v, tl.voltage[:tl.nl]) and np.array_equal(i, tl.current[:tl.nl]), tl.ID


def test_solve_hip_fills_line_outputs():
    from gprmax_amd.solver import solve_hip
    G = tl_model(iterations=80)
    Gc = tl_model(iterations=80)
    solve_hip(1, 1, G)
    osolver.solve_cpu(1, 1, Gc, nthreads=8)
    for a, b in zip(G.transmissionlines, Gc.transmissionlines):
        assert np.array_equal(a.Vtotal, b.Vtotal) and np.array_equal(a.Itotal, b.Itotal)


def test_conflicting_sources_are_refused():
    """A magnetic dipole on an H component the line reads, or a Hertzian
    dipole on its edge, cannot be ordered as the reference orders them."""
    from gprmax_amd.utilities import GeneralError
    for kind in ('magnetic', 'hertzian'):
        G = tl_model(iterations=20, extra_sources=False)
        bld_src = {'magnetic': ('x', (12, 10, 10)), 'hertzian': ('z', (12, 11, 10))}[kind]
        from gprmax_amd.sources import HertzianDipole, MagneticDipole
        src = MagneticDipole() if kind == 'magnetic' else HertzianDipole()
        src.polarisation = bld_src[0]
        src.xcoord, src.ycoord, src.zcoord = bld_src[1]
        src.ID, src.waveformID, src.start, src.stop = 'conflict', 'g', 0, G.timewindow
        if kind == 'hertzian':
            src.dl = G.dz
        src.calculate_waveform_values(G)
        (G.magneticdipoles if kind == 'magnetic' else G.hertziandipoles).append(src)
        with _hip()(G) as s:
            with pytest.raises(GeneralError, match='transmission line'):
                s.run(0, 2)


def test_batch_refuses_lines():
    from gprmax_amd.hip import HipBatch
    from gprmax_amd.utilities import GeneralError
    G = tl_model(iterations=20)
    with _hip()(G) as s:
        with pytest.raises(GeneralError, match='transmission lines'):
            HipBatch([s])
