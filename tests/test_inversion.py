"""Gradients of a trace misfit with respect to per-cell permittivity and
conductivity: the GPU adjoint (gprmax_amd.inversion) against a float64
pure-PyTorch autograd reference of the same discrete forward
(tests/torch_fdtd.py) and against central finite differences.

The reference itself is pinned first: its autograd gradient equals central
differences of its own float64 forward (CPU test)."""

import numpy as np
import pytest
import torch

from gprmax_amd.inversion import InversionGeometry
from torch_fdtd import forward_traces, misfit

SHAPE = (16, 15, 14)
PML = 4


def _setup(seed=0, iterations=70, freq=4e9):
    rng = np.random.default_rng(seed)
    geo = InversionGeometry(SHAPE, 2e-3, iterations, pml=PML, er_pml=7.0, K=16)
    waves = geo.waveform(freq=freq)
    er_true = 4 + 6 * rng.random(SHAPE)
    se_true = 1e-3 + 1.9e-2 * rng.random(SHAPE)
    sources = [(7, 7, 6), (9, 6, 8)]
    receivers = [(10, 7, 6), (5, 8, 7), (8, 10, 9), (6, 5, 8)]
    er0 = 7.0 + 0.3 * rng.standard_normal(SHAPE)
    se0 = 1.05e-2 + 1e-3 * rng.standard_normal(SHAPE)
    return geo, waves, er_true, se_true, er0, se0, sources, receivers


def _observed(geo, waves, er, se, sources, receivers):
    with torch.no_grad():
        return [forward_traces(torch.tensor(er), torch.tensor(se), geo, s, receivers, waves).numpy()
                for s in sources]


def _autograd(geo, waves, er0, se0, sources, receivers, obs):
    er = torch.tensor(er0, requires_grad=True)
    se = torch.tensor(se0, requires_grad=True)
    L = misfit(er, se, geo, sources, receivers, obs, waves)
    L.backward()
    return L.item(), er.grad.numpy(), se.grad.numpy()


def _fd(geo, waves, er0, se0, sources, receivers, obs, cell, which, rel=1e-4):
    base = (er0 if which == 'er' else se0)
    h = rel * abs(base[cell])

    def loss(delta):
        e, s = er0.copy(), se0.copy()
        (e if which == 'er' else s)[cell] += delta
        with torch.no_grad():
            return misfit(torch.tensor(e), torch.tensor(s), geo, sources, receivers, obs, waves).item()
    return (loss(h) - loss(-h)) / (2 * h)


def test_torch_reference_gradient_matches_finite_differences():
    geo, waves, er_t, se_t, er0, se0, src, rx = _setup(iterations=40)
    src, rx = src[:1], rx[:2]
    obs = _observed(geo, waves, er_t, se_t, src, rx)
    L, ger, gse = _autograd(geo, waves, er0, se0, src, rx, obs)
    assert L > 0
    for cell in [(7, 7, 7), (9, 6, 6), (6, 8, 8)]:
        assert abs(_fd(geo, waves, er0, se0, src, rx, obs, cell, 'er') - ger[cell]) < 1e-6 * abs(ger).max()
        assert abs(_fd(geo, waves, er0, se0, src, rx, obs, cell, 'se') - gse[cell]) < 1e-6 * abs(gse).max()


def _debye_params(shape, seed, P=2, frac=0.5):
    """Debye poles on a random fraction of the cells: deltaer ~ U[0.5, 5],
    tau = 10^U[-11, -9] s (deltaer 0 elsewhere)."""
    rng = np.random.default_rng(seed)
    mask = rng.random(shape) < frac
    de = rng.uniform(0.5, 5, (P,) + shape) * mask
    tau = 10 ** rng.uniform(-11, -9, (P,) + shape)
    return de, tau


def test_torch_debye_forward_matches_the_oracle():
    """The float64 PyTorch forward with Debye poles (the autograd reference)
    against the oracle's float64 restatement of the reference's dispersive
    parts A and B (pinned to the reference's compiled kernels): every cell
    its own material, painted voxel by voxel, no PML, 80 iterations; the
    receiver traces agree to float64 rounding."""
    from gprmax_amd.grid import ModelBuilder
    from gprmax_amd.materials import add_dispersion
    from oracle.solver import solve_cpu
    shape, d, its = (9, 8, 7), 2e-3, 80
    rng = np.random.default_rng(8)
    er = 4 + 6 * rng.random(shape)
    se = 1e-3 + 1.9e-2 * rng.random(shape)
    de, tau = _debye_params(shape, 9)
    geo = InversionGeometry(shape, d, its, pml=0, er_pml=7.0, K=16)
    geo.set_debye(de, tau)
    waves = geo.waveform(freq=4e9)
    src, rx = (4, 4, 3), [(2, 3, 4), (6, 5, 2), (4, 4, 3)]
    with torch.no_grad():
        ours = forward_traces(torch.tensor(er), torch.tensor(se), geo, src, rx, waves).numpy()
    b = ModelBuilder(*shape, d, d, d, its, pml_cells=0, precision='double')
    assert b.G.dt == geo.dt
    ids = np.zeros(shape, dtype=np.uint32)
    for n, c in enumerate(np.ndindex(*shape)):
        m = b.add_material(er[c], se[c], 1, 0, 'c{}'.format(n))
        if de[(0,) + c] > 0:
            add_dispersion(m, 'debye', list(de[(slice(None),) + c]), list(tau[(slice(None),) + c]))
        ids[c] = m.numID
    b.paint_voxels(np.ones(shape, dtype=bool), ids)
    b.add_waveform('ricker', 1, 4e9, 'w')
    h = b.add_hertzian_dipole('z', src, 'w')
    for r in rx:
        b.add_rx(r)
    G = b.finalise()
    h.waveformvalues_wholestep = waves.astype(np.float64)   # the inversion's float32 samples
    solve_cpu(1, 1, G, nthreads=4)
    ref = np.stack([r.outputs['Ez'] for r in G.rxs])
    assert np.abs(ref).max() > 0
    assert np.abs(ours - ref).max() <= 1e-9 * np.abs(ref).max()


def test_torch_debye_reference_gradient_matches_finite_differences():
    """The autograd gradient of the float64 reference with Debye poles on half
    the cells equals its own central differences."""
    geo, waves, er_t, se_t, er0, se0, src, rx = _setup(iterations=40)
    de, tau = _debye_params(SHAPE, 12)
    geo.set_debye(de, tau)
    src, rx = src[:1], rx[:2]
    obs = _observed(geo, waves, er_t, se_t, src, rx)
    L, ger, gse = _autograd(geo, waves, er0, se0, src, rx, obs)
    assert L > 0
    for cell in [(7, 7, 7), (9, 6, 6), (6, 8, 8)]:
        assert abs(_fd(geo, waves, er0, se0, src, rx, obs, cell, 'er') - ger[cell]) < 1e-6 * abs(ger).max()
        assert abs(_fd(geo, waves, er0, se0, src, rx, obs, cell, 'se') - gse[cell]) < 1e-6 * abs(gse).max()


@pytest.mark.gpu
def test_gpu_forward_matches_float64_reference():
    from gprmax_amd.inversion import InversionGrid
    geo, waves, er_t, se_t, er0, se0, src, rx = _setup()
    grid = InversionGrid(SHAPE, 2e-3, geo.iterations, pml=PML, er_pml=7.0, K=16)
    try:
        grid.set_model(er_t, se_t)
        for s in src:
            ours = grid.forward(s, rx, waves).astype(np.float64)
            with torch.no_grad():
                ref = forward_traces(torch.tensor(er_t), torch.tensor(se_t), geo, s, rx, waves).numpy()
            assert np.abs(ours - ref).max() <= 1e-4 * np.abs(ref).max()
    finally:
        grid.close()


@pytest.mark.gpu
def test_gpu_adjoint_matches_float64_autograd_and_finite_differences():
    """Checkpoint interval 16 over 70 iterations (the last segment partial),
    two sources, four receivers, PML on every face."""
    from gprmax_amd.inversion import InversionGrid, misfit_and_gradient
    geo, waves, er_t, se_t, er0, se0, src, rx = _setup()
    obs = _observed(geo, waves, er_t, se_t, src, rx)
    L_ref, ger_ref, gse_ref = _autograd(geo, waves, er0, se0, src, rx, obs)
    grid = InversionGrid(SHAPE, 2e-3, geo.iterations, pml=PML, er_pml=7.0, K=16)
    try:
        L, ger, gse = misfit_and_gradient(grid, er0, se0, src, rx, obs, waves)
    finally:
        grid.close()
    m = geo.interior
    assert abs(L - L_ref) <= 1e-4 * L_ref
    for ours, ref in ((ger, ger_ref), (gse, gse_ref)):
        err = np.linalg.norm(ours[m] - ref[m]) / np.linalg.norm(ref[m])
        assert err < 1e-3, err
    rng = np.random.default_rng(4)
    idx = np.argwhere(m)
    for n in rng.choice(len(idx), 6, replace=False):
        cell = tuple(idx[n])
        for which, g in (('er', ger), ('se', gse)):
            fd = _fd(geo, waves, er0, se0, src, rx, obs, cell, which)
            assert abs(g[cell] - fd) <= 1e-2 * abs(fd) + 1e-3 * np.abs(g[m]).max(), (cell, which, g[cell], fd)


@pytest.mark.gpu
def test_misfit_autograd_function():
    from gprmax_amd.inversion import InversionGrid, misfit as gpu_misfit
    geo, waves, er_t, se_t, er0, se0, src, rx = _setup(iterations=40)
    obs = _observed(geo, waves, er_t, se_t, src, rx)
    grid = InversionGrid(SHAPE, 2e-3, geo.iterations, pml=PML, er_pml=7.0, K=16)
    try:
        er = torch.tensor(er0, requires_grad=True)
        se = torch.tensor(se0, requires_grad=True)
        L = gpu_misfit(er, se, grid, src, rx, obs, waves)
        (2.0 * L).backward()
    finally:
        grid.close()
    _, ger_ref, gse_ref = _autograd(geo, waves, er0, se0, src, rx, obs)
    m = geo.interior
    assert np.linalg.norm(er.grad.numpy()[m] - 2 * ger_ref[m]) < 1e-3 * np.linalg.norm(2 * ger_ref[m])
    assert np.linalg.norm(se.grad.numpy()[m] - 2 * gse_ref[m]) < 1e-3 * np.linalg.norm(2 * gse_ref[m])


@pytest.mark.gpu
@pytest.mark.parametrize('poles', [1, 3])
def test_gpu_debye_adjoint_matches_float64_autograd(poles):
    """Debye poles (1 or 3, on half the cells) in the forward and the adjoint:
    the GPU forward within 1e-4 of the float64 reference's peak, and the GPU
    gradient within 1e-3 (relative L2 outside the PML) of its autograd
    gradient, over 70 iterations with K = 16 (a partial last segment)."""
    from gprmax_amd.inversion import InversionGrid, misfit_and_gradient
    geo, waves, er_t, se_t, er0, se0, src, rx = _setup()
    de, tau = _debye_params(SHAPE, 20 + poles, P=poles)
    geo.set_debye(de, tau)
    obs = _observed(geo, waves, er_t, se_t, src, rx)
    L_ref, ger_ref, gse_ref = _autograd(geo, waves, er0, se0, src, rx, obs)
    grid = InversionGrid(SHAPE, 2e-3, geo.iterations, pml=PML, er_pml=7.0, K=16)
    try:
        grid.set_debye(de, tau)
        grid.set_model(er_t, se_t)
        with torch.no_grad():
            ref = forward_traces(torch.tensor(er_t), torch.tensor(se_t), geo, src[0], rx, waves).numpy()
        ours = grid.forward(src[0], rx, waves).astype(np.float64)
        assert np.abs(ours - ref).max() <= 1e-4 * np.abs(ref).max()
        L, ger, gse = misfit_and_gradient(grid, er0, se0, src, rx, obs, waves)
    finally:
        grid.close()
    m = geo.interior
    assert abs(L - L_ref) <= 1e-4 * L_ref
    for ours, ref in ((ger, ger_ref), (gse, gse_ref)):
        err = np.linalg.norm(ours[m] - ref[m]) / np.linalg.norm(ref[m])
        assert err < 1e-3, err


@pytest.mark.gpu
@pytest.mark.parametrize('poles', [0, 3])
def test_gpu_kept_segments_give_the_recomputed_gradient(poles):
    """The forward keeping the stored states of none, some (2 of 5, the last
    one partial) or all of its checkpoint segments: traces, loss and both
    gradients equal those of the all-recompute backward bit for bit, also
    when the kept set shrinks between forwards."""
    from gprmax_amd.inversion import InversionGrid, misfit_and_gradient
    geo, waves, er_t, se_t, er0, se0, src, rx = _setup()
    if poles:
        de, tau = _debye_params(SHAPE, 30 + poles, P=poles)
        geo.set_debye(de, tau)
    obs = _observed(geo, waves, er_t, se_t, src, rx)
    grid = InversionGrid(SHAPE, 2e-3, geo.iterations, pml=PML, er_pml=7.0, K=16)
    try:
        if poles:
            grid.set_debye(de, tau)
        L0, ger0, gse0 = misfit_and_gradient(grid, er0, se0, src, rx, obs, waves)
        seg = grid.keep_bytes_all() / 5          # 70 iterations, K = 16: 5 segments
        for nbytes in (grid.keep_bytes_all(), 2.5 * seg, 0):
            grid.set_keep(nbytes)
            L, ger, gse = misfit_and_gradient(grid, er0, se0, src, rx, obs, waves)
            assert L == L0, nbytes
            assert np.array_equal(ger, ger0) and np.array_equal(gse, gse0), nbytes
    finally:
        grid.close()


@pytest.mark.gpu
def test_gpu_adjoint_without_fast_box_matches_float64_autograd():
    """A grid too thin for the gathers' fast box (10 cells along z with a
    4-cell PML on each face: the adjoint runs its generic path on every
    point, in one launch per half-step) with a receiver inside the PML: the
    gradient within 1e-3 (relative L2 outside the PML) of the float64
    autograd gradient, over 50 iterations with K = 16."""
    from gprmax_amd.inversion import InversionGrid, misfit_and_gradient
    shape, pml = (12, 11, 10), 4
    rng = np.random.default_rng(11)
    geo = InversionGeometry(shape, 2e-3, 50, pml=pml, er_pml=7.0, K=16)
    waves = geo.waveform(freq=4e9)
    er_t = 4 + 6 * rng.random(shape)
    se_t = 1e-3 + 1.9e-2 * rng.random(shape)
    er0 = 7.0 + 0.3 * rng.standard_normal(shape)
    se0 = 1.05e-2 + 1e-3 * rng.standard_normal(shape)
    src = [(6, 5, 5)]
    rx = [(5, 5, 4), (7, 6, 5), (2, 5, 5)]
    obs = _observed(geo, waves, er_t, se_t, src, rx)
    L_ref, ger_ref, gse_ref = _autograd(geo, waves, er0, se0, src, rx, obs)
    grid = InversionGrid(shape, 2e-3, geo.iterations, pml=pml, er_pml=7.0, K=16)
    try:
        L, ger, gse = misfit_and_gradient(grid, er0, se0, src, rx, obs, waves)
    finally:
        grid.close()
    m = geo.interior
    assert m.sum() == 3 * 2 * 1   # cells whose every edge lies outside the slabs
    assert abs(L - L_ref) <= 1e-4 * L_ref
    for ours, ref in ((ger, ger_ref), (gse, gse_ref)):
        err = np.linalg.norm(ours[m] - ref[m]) / np.linalg.norm(ref[m])
        assert err < 1e-3, err
