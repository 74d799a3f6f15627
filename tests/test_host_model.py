"""Host model layer: time step, coefficients, PML profiles, waveforms, ID build.

Each check is against a closed form or an independent brute force, not a
retyping of the code under test.
"""

import numpy as np
import pytest

from gprmax_amd.constants import c, e0, m0
from gprmax_amd.grid import ModelBuilder
from gprmax_amd.materials import Material
from gprmax_amd.pml import CFS
from gprmax_amd.utilities import courant_dt, iterations_for, round_value
from gprmax_amd.waveforms import Waveform


def test_dt_and_iterations_match_reference_files():
    # The reference's receiver files record dt and Iterations for 1 mm cells
    # and a 3 ns window (tests/models_basic/*/*.out attributes).
    ref = np.load('tests/golden/hertzian_dipole_fs_ref.npz')
    dt, mode = courant_dt(100, 100, 100, 1e-3, 1e-3, 1e-3)
    assert mode == '3D'
    assert dt == float(ref['dt'])
    assert iterations_for(3e-9, dt)[0] == int(ref['iterations'])
    # 2-D files fix the 2-D Courant limit too
    ref2 = np.load('tests/golden/2D_EzHxHy_ref.npz')
    assert courant_dt(100, 100, 1, 1e-3, 1e-3, 1e-3)[0] == float(ref2['dt'])
    # dt never exceeds the Courant limit and is within 1e-15 of it
    raw = 1 / (c * np.sqrt(3) / 1e-3)
    assert dt <= raw and raw - dt < 1e-26


def test_round_value_half_down():
    assert round_value(2.5) == 2 and round_value(3.5) == 3 and round_value(2.51) == 3
    assert round_value(1.23456789, 3) == 1.234


def test_free_space_and_pec_coefficients_closed_form():
    b = ModelBuilder(20, 20, 20, 1e-3, 1e-3, 1e-3, 10, pml_cells=0)
    G = b.finalise()
    dt = G.dt
    # free space: CA = 1, CB = dt/(e0 dx), DA = 1, DB = dt/(m0 dx)
    assert np.isclose(G.updatecoeffsE[1, 0], 1.0, rtol=0, atol=0)
    assert np.isclose(G.updatecoeffsE[1, 1], dt / (e0 * 1e-3), rtol=1e-7)
    assert np.isclose(G.updatecoeffsE[1, 4], dt / e0, rtol=1e-7)
    assert np.isclose(G.updatecoeffsH[1, 1], dt / (m0 * 1e-3), rtol=1e-7)
    assert np.all(G.updatecoeffsE[0] == 0)          # pec
    assert np.isclose(G.updatecoeffsH[0, 0], 1.0)   # pec still updates H


def test_lossy_coefficient_closed_form():
    b = ModelBuilder(20, 20, 20, 2e-3, 2e-3, 2e-3, 10, pml_cells=0)
    m = b.add_material(6.0, 0.02, 1, 0, 'soil')
    G = b.finalise()
    dt = G.dt
    EA = e0 * 6 / dt + 0.01
    EB = e0 * 6 / dt - 0.01
    assert np.isclose(G.updatecoeffsE[m.numID, 0], EB / EA, rtol=1e-7)
    assert np.isclose(G.updatecoeffsE[m.numID, 2], 1 / (EA * 2e-3), rtol=1e-7)


def test_debye_coefficients_closed_form():
    # One Debye pole: eqt = exp(-dt/tau), zt = (w/q)(1-eqt)/dt with w = de/tau, q = -1/tau
    b = ModelBuilder(20, 20, 20, 1e-3, 1e-3, 1e-3, 10, pml_cells=0)
    m = b.add_material(4.9, 0, 1, 0, 'water')
    from gprmax_amd.materials import add_dispersion
    add_dispersion(m, 'debye', [75.2], [9.231e-12])
    G = b.finalise()
    dt, tau = G.dt, 9.231e-12
    eqt = np.exp(-dt / tau)
    zt = (75.2 / tau) / (-1 / tau) * (1 - eqt) / dt
    row = G.updatecoeffsdispersive[m.numID]
    assert np.isclose(row[1].real, eqt, rtol=1e-6) and row[1].imag == 0
    assert np.isclose(row[2].real, zt, rtol=1e-5) and row[2].imag == 0
    assert np.isclose(row[0].real, e0 * np.exp(-dt / (2 * tau)), rtol=1e-6)
    # non-dispersive materials of a dispersive model have zero pole terms
    assert np.all(G.updatecoeffsdispersive[1] == 0)
    # complex permittivity at DC: er_inf + de
    assert np.isclose(m.calculate_er(1e-3).real, 4.9 + 75.2, rtol=1e-9)


def test_pml_profiles_closed_form():
    cfs = CFS()
    cfs.sigma.max = 2.0
    E, H = cfs.calculate_values(10, cfs.sigma)
    # quartic: sigma_E(p) = smax (p/t)^4, sigma_H(p) = smax ((p+1/2)/t)^4
    p = np.arange(10)
    assert np.allclose(E, 2.0 * (p / 10) ** 4, rtol=1e-12)
    assert np.allclose(H, 2.0 * ((p + 0.5) / 10) ** 4, rtol=1e-12)
    kE, kH = cfs.calculate_values(10, cfs.kappa)
    assert np.all(kE == 1) and np.all(kH == 1)


def test_pml_sigma_max_from_first_slab_only():
    # x0 slab sits on er=8 cells for half its first layer; every other slab
    # must reuse that sigma max (shared CFS object).
    b = ModelBuilder(30, 30, 30, 1e-3, 1e-3, 1e-3, 10, pml_cells=5)
    hs = b.add_material(8, 0, 1, 0, 'hs')
    b.box(0, 30, 0, 30, 0, 15, hs, averaging=True)
    G = b.finalise()
    from gprmax_amd.constants import z0
    smax = 0.8 * 5 / (z0 * 1e-3 * np.sqrt((8 + 1) / 2))
    assert np.isclose(G.cfs[0].sigma.max, smax, rtol=1e-12)
    assert [p.ID for p in G.pmls] == ['x0', 'y0', 'z0', 'xmax', 'ymax', 'zmax']
    # HORIPML order 1, default CFS: RF = 2 sigma dt / (2 e0 + dt sigma)
    sig = smax * ((np.arange(5) + 0.0) / 5) ** 4
    rf = 2 * sig * G.dt / (2 * e0 + G.dt * sig)
    for pml in G.pmls:
        assert np.allclose(pml.ERF[0], rf, rtol=1e-6)


def test_ricker_closed_form_and_integral():
    w = Waveform('r', 'ricker', 1.0, 1.5e9)
    f = 1.5e9
    chi, zeta = np.sqrt(2) / f, np.pi ** 2 * f ** 2
    t = np.linspace(0, 3e-9, 301)
    v = np.array([w.calculate_value(x, 1e-12) for x in t])
    d = t - chi
    assert np.allclose(v, (1 - 2 * zeta * d ** 2) * np.exp(-zeta * d ** 2), rtol=1e-12, atol=1e-15)
    # the Ricker pulse has zero mean
    tt = np.linspace(0, 4 * chi, 20001)
    vv = np.array([w.calculate_value(x, 1e-12) for x in tt])
    assert abs(np.trapezoid(vv, tt)) < 1e-3 * chi


def test_gaussiandot_is_derivative_of_gaussian():
    g = Waveform('g', 'gaussian', 1.0, 1e9)
    gd = Waveform('gd', 'gaussiandot', 1.0, 1e9)
    h = 1e-15
    for x in np.linspace(0.2e-9, 2e-9, 13):
        fd = (g.calculate_value(x + h, 1e-12) - g.calculate_value(x - h, 1e-12)) / (2 * h)
        assert np.isclose(gd.calculate_value(x, 1e-12), fd, rtol=1e-5, atol=1e-3)


def _brute_voxels(nx, ny, nz, mask, numids):
    """build_voxel (gprMax/geometry_primitives_ext.pyx:337-376) one voxel at a time."""
    ID = np.ones((6, nx + 1, ny + 1, nz + 1), dtype=np.uint32)
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                if not mask[i, j, k]:
                    continue
                n = numids[i, j, k]
                for c in (0, 3):
                    for (a, b) in ((0, 0), (1, 1), (1, 0), (0, 1)):
                        ID[c, i, j + a, k + b] = n
                for c in (1, 4):
                    for (a, b) in ((0, 0), (1, 1), (1, 0), (0, 1)):
                        ID[c, i + a, j, k + b] = n
                for c in (2, 5):
                    for (a, b) in ((0, 0), (1, 1), (1, 0), (0, 1)):
                        ID[c, i + a, j + b, k] = n
    return ID


def test_voxel_painting_matches_brute_force():
    rng = np.random.default_rng(3)
    nx, ny, nz = 7, 6, 5
    b = ModelBuilder(nx, ny, nz, 1e-3, 1e-3, 1e-3, 5, pml_cells=0)
    for n in range(5):
        b.add_material(2 + n, 0, 1, 0, 'm%d' % n)
    mask = rng.random((nx, ny, nz)) < 0.6
    ids = rng.integers(2, 7, (nx, ny, nz)).astype(np.uint32)
    b.paint_voxels(mask, ids)
    assert np.array_equal(b.G.ID, _brute_voxels(nx, ny, nz, mask, ids))


def test_averaged_half_space_edges():
    # hertzian_dipole_hs: edges on the z=50 interface average four cells
    # (two of each material), H faces there average two.
    b = ModelBuilder(20, 20, 20, 1e-3, 1e-3, 1e-3, 10, pml_cells=0)
    hs = b.add_material(8, 0, 1, 0, 'half_space')
    b.box(0, 20, 0, 20, 0, 10, hs, averaging=True)
    G = b.finalise()
    names = {m.numID: m for m in G.materials}
    ex = names[int(G.ID[0, 5, 5, 10])]
    assert ex.er == pytest.approx(4.5) and ex.ID.count('half_space') == 2
    assert int(G.ID[0, 5, 5, 9]) == hs.numID and int(G.ID[0, 5, 5, 11]) == 1
    assert int(G.ID[1, 5, 5, 10]) == ex.numID
    assert int(G.ID[2, 5, 5, 9]) == hs.numID
    # The reference's magnetic match accepts a material holding neither name an
    # even number of times (zero), so the first such material (pec) is taken.
    assert int(G.ID[5, 5, 5, 10]) == 0


@pytest.mark.parametrize('axis', [0, 1, 2])
def test_2d_model_has_no_pml_across_and_pec_on_its_planes(axis):
    """A grid one cell thick is a 2-D model: the Courant limit of the other two
    axes, no PML slab on the invariant axis, and the two E components across
    that axis PEC on both of its grid planes (model_build_run.py:222-241)."""
    shape = [12, 12, 12]
    shape[axis] = 1
    b = ModelBuilder(*shape, 1e-3, 2e-3, 3e-3, 10, pml_cells=3)
    G = b.finalise()
    assert G.mode == '2D TM' + 'xyz'[axis]
    d = [1e-3, 2e-3, 3e-3]
    d.pop(axis)
    raw = 1 / (c * np.sqrt(1 / d[0] ** 2 + 1 / d[1] ** 2))
    assert G.dt <= raw and raw - G.dt < 1e-26
    assert sorted(p.direction[0] for p in G.pmls) == sorted(['xyz'[a] for a in range(3) if a != axis] * 2)
    for comp in range(3):
        for plane in (0, 1):
            sl = [slice(None)] * 3
            sl[axis] = plane
            ids = G.ID[comp][tuple(sl)]
            assert (not ids.any()) == (comp != axis), (comp, plane)
    # the in-plane H and the axis E keep free space
    assert (G.ID[axis] == 1).any()


def test_axis_aligned_cylinder_matches_brute_force():
    """#cylinder: the cells whose centre lies within r of the axis, tested with
    the reference's float32 cell positions (geometry_primitives_ext.pyx:
    816-822), for each alignment; a pure-Python loop over all cells."""
    f = np.float32
    for axis, (p1, p2) in enumerate((((0.0, 0.021, 0.017), (0.030, 0.021, 0.017)),
                                     ((0.013, 0.0, 0.019), (0.013, 0.040, 0.019)),
                                     ((0.019, 0.018, 0.0), (0.019, 0.018, 0.024)))):
        b = ModelBuilder(40, 41, 42, 1e-3, 1e-3, 1e-3, 10, pml_cells=0)
        m = b.add_material(4, 0, 1, 0, 'rod')
        r = 0.0071
        b.cylinder(*p1, *p2, r, m, averaging=False)
        got = b.G.solid == m.numID
        u, v = [a for a in range(3) if a != axis]
        want = np.zeros_like(got)
        for i in range(40):
            for j in range(41):
                for k in range(42):
                    idx = (i, j, k)
                    lo = round_value(p1[axis] / 1e-3)
                    hi = round_value(p2[axis] / 1e-3)
                    if not (lo <= idx[axis] < hi):
                        continue
                    du = float(f(idx[u]) * f(1e-3)) + 0.5 * float(f(1e-3)) - float(f(p1[u]))
                    dv = float(f(idx[v]) * f(1e-3)) + 0.5 * float(f(1e-3)) - float(f(p1[v]))
                    want[idx] = np.sqrt(du * du + dv * dv) <= float(f(r))
        assert want.sum() > 100
        assert np.array_equal(got, want), axis
        # a rigid cylinder writes every edge of its cells
        assert (b.G.ID[axis][tuple(np.argwhere(want)[0])] == m.numID)


def _loop_edge_build(G):
    """The reference's edge loops (gprMax/yee_cell_build_ext.pyx:110-257),
    one edge at a time in loop order, for comparison with the vectorised
    build."""
    from gprmax_amd.grid import _average_material
    nx, ny, nz = G.nx, G.ny, G.nz
    specs = [(0, ((0, nx), (1, ny), (1, nz)), ((0, 0, 0), (0, -1, 0), (0, -1, -1), (0, 0, -1)), True),
             (1, ((1, nx), (0, ny), (1, nz)), ((0, 0, 0), (-1, 0, 0), (-1, 0, -1), (0, 0, -1)), True),
             (2, ((1, nx), (1, ny), (0, nz)), ((0, 0, 0), (-1, 0, 0), (-1, -1, 0), (0, -1, 0)), True),
             (3, ((1, nx), (0, ny), (0, nz)), ((0, 0, 0), (-1, 0, 0)), False),
             (4, ((0, nx), (1, ny), (0, nz)), ((0, 0, 0), (0, -1, 0)), False),
             (5, ((0, nx), (0, ny), (1, nz)), ((0, 0, 0), (0, 0, -1)), False)]
    for comp, ((i0, i1), (j0, j1), (k0, k1)), offs, electric in specs:
        rigid = G.rigidcellE if electric else G.rigidcellH
        cache = {}
        for i in range(i0, i1):
            for j in range(j0, j1):
                for k in range(k0, k1):
                    cells = [(i + a, j + b, k + c) for a, b, c in offs]
                    if any(rigid[c] for c in cells):
                        continue
                    ids = [int(G.solid[c]) for c in cells]
                    if all(n == ids[0] for n in ids):
                        G.ID[comp, i, j, k] = ids[0]
                    else:
                        names = [G.materials[n].ID for n in ids]
                        G.ID[comp, i, j, k] = _average_material(G, names, ids, cache, electric)


@pytest.mark.parametrize('seed', [0, 1])
def test_vectorised_edge_build_matches_edge_loop(seed):
    """Averaged random media with rigid voxels and an averaged box: the same
    edge indices and the same averaged materials, created in the same order
    with the same names and properties, as the edge-by-edge loop."""
    from gprmax_amd.grid import build_electric_components, build_magnetic_components

    def model():
        rng = np.random.default_rng(seed)
        b = ModelBuilder(14, 12, 11, 1e-3, 1e-3, 1e-3, 10, pml_cells=0)
        mats = [b.add_material(2 + 0.7 * n, 0.002 * n, 1, 0, 'm{}'.format(n)) for n in range(5)]
        b.G.solid[:, :, :7] = rng.integers(2, 7, (14, 12, 7))
        b.box(3, 9, 2, 8, 5, 9, mats[1], averaging=True)
        mask = rng.random((14, 12, 11)) < 0.08
        b.paint_voxels(mask, rng.integers(0, 7, (14, 12, 11)).astype(np.uint32))
        return b.G
    G1, G2 = model(), model()
    build_electric_components(G1)
    build_magnetic_components(G1)
    _loop_edge_build(G2)
    assert np.array_equal(G1.ID, G2.ID)
    assert [m.ID for m in G1.materials] == [m.ID for m in G2.materials]
    assert len(G1.materials) > 30
    for a, b in zip(G1.materials, G2.materials):
        assert (a.er, a.se, a.mr, a.sm) == (b.er, b.se, b.mr, b.sm)


class _RefGrid:
    """The attributes the reference's edge build reads from its grid."""
    IDlookup = {'Ex': 0, 'Ey': 1, 'Ez': 2, 'Hx': 3, 'Hy': 4, 'Hz': 5}

    def __init__(self, G):
        self.nx, self.ny, self.nz = G.nx, G.ny, G.nz
        self.materials = G.materials
        self.ID = G.ID


@pytest.mark.parametrize('seed', [0, 1, 2])
def test_edge_build_matches_the_reference_compiled_build(seed, ref_edge_build):
    """The reference's own yee_cell_build_ext (compiled from its .pyx into
    oracle/_ref) and the vectorised host build of gprmax_amd.grid on the same grid:
    identical edge indices, and identical averaged materials created in the
    same order with the same names and constitutive values.  The rigid flags
    go to the reference as its 12 (E) and 6 (H) edge slots per cell, all set
    for a rigid cell (set_rigid_E / set_rigid_H write every slot of a cell).
    The GPU edge build is compared with the host build in test_gpu_edges.py."""
    from gprmax_amd.grid import build_electric_components, build_magnetic_components
    from models import edge_model
    G1, G2 = edge_model(seed), edge_model(seed)
    build_electric_components(G1)
    build_magnetic_components(G1)
    rigidE = np.ascontiguousarray(np.repeat(G2.rigidcellE[None], 12, axis=0).astype(np.int8))
    rigidH = np.ascontiguousarray(np.repeat(G2.rigidcellH[None], 6, axis=0).astype(np.int8))
    R = _RefGrid(G2)
    ref_edge_build.build_electric_components(G2.solid, rigidE, G2.ID, R)
    ref_edge_build.build_magnetic_components(G2.solid, rigidH, G2.ID, R)
    assert np.array_equal(G1.ID, G2.ID)
    assert len(G1.materials) > 50
    assert [m.ID for m in G1.materials] == [m.ID for m in G2.materials]
    assert [m.numID for m in G2.materials] == list(range(len(G2.materials)))
    for a, b in zip(G1.materials, G2.materials):
        assert (a.er, a.se, a.mr, a.sm) == (b.er, b.se, b.mr, b.sm), a.ID
        assert a.type == b.type, a.ID


def test_device_tables_from_reference_like_objects():
    """solve_hip reads attributes only: the reference's own source, line,
    snapshot and grid objects (which lack this package's helper methods)
    pack into the same device tables."""
    from types import SimpleNamespace
    from gprmax_amd.hip import grid_maxpoles
    from gprmax_amd.snapshots import store_components
    from gprmax_amd.sources import device_source_table, device_tl_args
    from tl_models import tl_model
    G = tl_model(iterations=40)
    plain = lambda o: SimpleNamespace(**{k: v for k, v in vars(o).items()})
    for kind, srcs in (('hertzian', G.hertziandipoles), ('magnetic', G.magneticdipoles),
                       ('voltage', G.voltagesources)):
        ours = device_source_table(srcs, G)
        theirs = device_source_table([plain(s) for s in srcs], G, kind)
        for a, b in zip(ours, theirs):
            assert np.array_equal(a, b), kind
    for tl in G.transmissionlines:
        for a, b in zip(device_tl_args(tl, G), device_tl_args(plain(tl), G)):
            assert np.array_equal(a, b)
        assert np.array_equal(device_tl_args(tl, G)[1], np.array(tl.coefficients(G)))
    g = SimpleNamespace(updatecoeffsdispersive=np.zeros((7, 9), np.complex64))
    assert grid_maxpoles(g) == 3 and grid_maxpoles(SimpleNamespace()) == 0 and grid_maxpoles(G) == G.maxpoles
    snap = SimpleNamespace()
    comps = [np.full((2, 3, 4), v, np.float32) for v in range(6)]
    store_components(snap, *comps)
    assert np.array_equal(snap.electric, np.stack(comps[:3]).reshape(-1, order='F'))
    assert np.array_equal(snap.magnetic, np.stack(comps[3:]).reshape(-1, order='F'))
