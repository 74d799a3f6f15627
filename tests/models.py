"""The reference's basic test models, built programmatically.

Each function mirrors one ``tests/models_basic/<model>/<model>.in`` of the
reference (100 x 100 x 100 cells of 1 mm, 3 ns, default 10-cell PML,
gaussiandot 1 GHz source at (50, 50, 50), rx at (70, 70, 70)).
"""

from gprmax_amd.grid import ModelBuilder
from gprmax_amd.materials import add_dispersion


def _base(precision='single', n=100, time_window=3e-9, pml_cells=10, pml_formulation='HORIPML', cfs=None):
    b = ModelBuilder(n, n, n, 0.001, 0.001, 0.001, time_window, pml_cells=pml_cells,
                     pml_formulation=pml_formulation, cfs=cfs, precision=precision)
    # source at the centre, receiver 20% of the side further along each axis
    b.src = (n // 2,) * 3
    b.rx = (n // 2 + n // 5,) * 3
    return b


def hertzian_dipole_fs(precision='single', **kw):
    """tests/models_basic/hertzian_dipole_fs/hertzian_dipole_fs.in"""
    b = _base(precision, **kw)
    b.add_waveform('gaussiandot', 1, 1e9, 'myWave')
    b.add_hertzian_dipole('z', b.src, 'myWave')
    b.add_rx(b.rx)
    return b.finalise()


def magnetic_dipole_fs(precision='single', **kw):
    """tests/models_basic/magnetic_dipole_fs/magnetic_dipole_fs.in"""
    b = _base(precision, **kw)
    b.add_waveform('gaussiandot', 1, 1e9, 'myWave')
    b.add_magnetic_dipole('z', b.src, 'myWave')
    b.add_rx(b.rx)
    return b.finalise()


def hertzian_dipole_hs(precision='single', n=100, **kw):
    """tests/models_basic/hertzian_dipole_hs/hertzian_dipole_hs.in: a box of
    er=8 over z in [0, 50) cells, averaged at its surface."""
    b = _base(precision, n=n, **kw)
    hs = b.add_material(8, 0, 1, 0, 'half_space')
    b.add_waveform('gaussiandot', 1, 1e9, 'myWave')
    b.add_hertzian_dipole('z', b.src, 'myWave')
    b.add_rx(b.rx)
    b.box(0, n, 0, n, 0, n // 2, hs, averaging=True)
    return b.finalise()


def hertzian_dipole_dispersive(precision='single', n=100, **kw):
    """tests/models_basic/hertzian_dipole_dispersive/hertzian_dipole_dispersive.in:
    water (er_inf 4.9, one Debye pole of 75.2 and 9.231 ps) filling the domain."""
    b = _base(precision, n=n, **kw)
    water = b.add_material(4.9, 0, 1, 0, 'myWater')
    add_dispersion(water, 'debye', [75.2], [9.231e-12])
    b.add_waveform('gaussiandot', 1, 1e9, 'myWave')
    b.add_hertzian_dipole('z', b.src, 'myWave')
    b.add_rx(b.rx)
    b.box(0, n, 0, n, 0, n, water, averaging=True)
    return b.finalise()


def _plane(axis, precision='single', n=100, **kw):
    """tests/models_basic/2D_{ExHyHz,EyHxHz,EzHxHy}/*.in: a 100 x 100 cell
    plane, one cell thick along `axis`, default PML (none on that axis),
    gaussiandot 1 GHz dipole along `axis` at the centre, rx 20 cells further
    along both in-plane axes."""
    size = [n, n, n]
    size[axis] = 1
    b = ModelBuilder(*size, 0.001, 0.001, 0.001, 3e-9, precision=precision, **kw)
    src = [n // 2] * 3
    rx = [n // 2 + n // 5] * 3
    src[axis] = rx[axis] = 0
    b.add_waveform('gaussiandot', 1, 1e9, 'myWave')
    b.add_hertzian_dipole('xyz'[axis], tuple(src), 'myWave')
    b.add_rx(tuple(rx))
    return b.finalise()


def plane_ExHyHz(precision='single', **kw):
    return _plane(0, precision, **kw)


def plane_EyHxHz(precision='single', **kw):
    return _plane(1, precision, **kw)


def plane_EzHxHy(precision='single', **kw):
    return _plane(2, precision, **kw)


def cylinder_Ascan_2D(precision='single', **kw):
    """tests/models_basic/cylinder_Ascan_2D/cylinder_Ascan_2D.in: 240 x 210 mm
    TMz plane of 2 mm cells, er=6 half-space below y = 170 mm holding a PEC
    cylinder of 10 mm radius at (120, 80) mm, ricker 1.5 GHz dipole at
    (100, 170) mm, rx at (140, 170) mm."""
    b = ModelBuilder(120, 105, 1, 0.002, 0.002, 0.002, 3e-9, precision=precision, **kw)
    hs = b.add_material(6, 0, 1, 0, 'half_space')
    b.add_waveform('ricker', 1, 1.5e9, 'my_ricker')
    b.add_hertzian_dipole('z', (50, 85, 0), 'my_ricker')
    b.add_rx((70, 85, 0))
    b.box(0, 120, 0, 85, 0, 1, hs, averaging=True)
    b.cylinder(0.120, 0.080, 0, 0.120, 0.080, 0.002, 0.010, b.G.materials[0])
    return b.finalise()


PLANE_MODELS = {
    '2D_ExHyHz': plane_ExHyHz,
    '2D_EyHxHz': plane_EyHxHz,
    '2D_EzHxHy': plane_EzHxHy,
    'cylinder_Ascan_2D': cylinder_Ascan_2D,
}


BASIC_MODELS = {
    'hertzian_dipole_fs': hertzian_dipole_fs,
    'magnetic_dipole_fs': magnetic_dipole_fs,
    'hertzian_dipole_hs': hertzian_dipole_hs,
    'hertzian_dipole_dispersive': hertzian_dipole_dispersive,
}
BASIC_MODELS.update(PLANE_MODELS)


def edge_model(seed, shape=(30, 26, 21), nmat=6, frac=0.05):
    """A grid whose edges need averaging everywhere: random cell materials in
    the lower half, an averaged box and cylinder, and rigidly painted random
    voxels (some of them PEC and free space) on top."""
    import numpy as np
    rng = np.random.default_rng(seed)
    b = ModelBuilder(*shape, 1e-3, 1e-3, 1e-3, 10, pml_cells=0)
    mats = [b.add_material(2 + 0.7 * n, 0.002 * n, 1 + 0.1 * n, 0, 'm{}'.format(n)) for n in range(nmat)]
    nx, ny, nz = shape
    b.G.solid[:, :, :nz // 2] = rng.integers(2, 2 + nmat, (nx, ny, nz // 2))
    b.box(nx // 4, nx // 2, 2, ny - 3, nz // 3, nz - 2, mats[1], averaging=True)
    b.cylinder(0.5e-3 * nx, 0.5e-3 * ny, 0, 0.5e-3 * nx, 0.5e-3 * ny, 1e-3 * nz, 4e-3, mats[2], averaging=True)
    mask = rng.random(shape) < frac
    b.paint_voxels(mask, rng.integers(0, 2 + nmat, shape).astype(np.uint32))
    return b.G
