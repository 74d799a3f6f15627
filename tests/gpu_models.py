"""Small seeded models that exercise every feature of the solve path at once:
random materials on every edge, PML on all six faces, all source kinds,
receivers, snapshots and optional dispersion."""

import numpy as np

from gprmax_amd.grid import ModelBuilder
from gprmax_amd.materials import add_dispersion
from gprmax_amd.pml import CFS


def randomise_tables(G, seed, scaled=False):
    """Replace the coefficient tables (every column of updatecoeffsE/H drawn
    on its own) and every slab's R tables with random values, so that a
    swapped coefficient column (CBx/CBy/CBz, DBx/DBy/DBz, srce/srcm) or a
    profile read at the wrong index changes the results.  The model keeps
    its geometry, sources and dispersive tables.

    scaled=False draws every entry from U[-1, 1] (R from U[0.5, 1.5]), for
    single sub-steps.  scaled=True multiplies every physical entry by its own
    factor from U[0.5, 1.5] (R entries by U[0.9, 1.1]), which keeps the
    magnitudes (the PML divides by cell sizes of a millimetre) so that runs
    of several iterations stay finite."""
    rng = np.random.default_rng(seed)
    n = len(G.updatecoeffsE)
    if scaled:
        G.updatecoeffsE = (G.updatecoeffsE * rng.uniform(0.5, 1.5, (n, 5))).astype(np.float32)
        G.updatecoeffsH = (G.updatecoeffsH * rng.uniform(0.5, 1.5, (n, 5))).astype(np.float32)
    else:
        G.updatecoeffsE = rng.uniform(-1, 1, (n, 5)).astype(np.float32)
        G.updatecoeffsH = rng.uniform(-1, 1, (n, 5)).astype(np.float32)
    for p in G.pmls:
        for name in ('ERA', 'ERB', 'ERE', 'ERF', 'HRA', 'HRB', 'HRE', 'HRF'):
            a = getattr(p, name)
            if scaled:
                setattr(p, name, (a * rng.uniform(0.9, 1.1, a.shape)).astype(a.dtype))
            else:
                setattr(p, name, rng.uniform(0.5, 1.5, a.shape).astype(a.dtype))
    return G


def rich_model(seed=0, shape=(26, 22, 19), iterations=40, pml=4, order=1, formulation='HORIPML',
               poles=0, kind='debye', nmat=9, snapshots=True, cell=(1e-3, 1e-3, 1e-3)):
    """`pml` is the thickness of every face, or a dict of per-face
    thicknesses (x0, y0, z0, xmax, ymax, zmax; a missing or zero face is
    open), as #pml_cells gives them."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    faces = pml if isinstance(pml, dict) else None
    if faces is not None:
        pml = max(faces.values(), default=0)
    cfs = None
    if order == 2:
        a, b = CFS(), CFS()
        b.alpha.max = 0.05
        b.kappa.scalingprofile, b.kappa.min, b.kappa.max = 'quadratic', 1, 3
        cfs = [a, b]
    bld = ModelBuilder(nx, ny, nz, *cell, iterations, pml_cells=faces if faces is not None else pml,
                       pml_formulation=formulation, cfs=cfs)
    mats = []
    for n in range(nmat):
        m = bld.add_material(rng.uniform(1, 9), rng.uniform(0, 0.05), 1, 0, 'm{}'.format(n))
        if poles and n % 2 == 0:
            if kind == 'debye':
                add_dispersion(m, 'debye', list(rng.uniform(0.5, 5, poles)), list(10 ** rng.uniform(-11, -9, poles)))
            elif kind == 'lorentz':
                add_dispersion(m, 'lorentz', list(rng.uniform(0.5, 5, poles)), list(rng.uniform(1e9, 5e9, poles)),
                               list(rng.uniform(1e8, 5e8, poles)))
            else:
                add_dispersion(m, 'drude', None, list(rng.uniform(1e9, 3e9, poles)), list(rng.uniform(1e9, 5e9, poles)))
        mats.append(m)
    mask = rng.random((nx, ny, nz)) < 0.7
    ids = rng.integers(2, 2 + nmat, (nx, ny, nz)).astype(np.uint32)
    bld.paint_voxels(mask, ids)
    # a PEC block
    bld.box(nx // 2 - 2, nx // 2 + 1, 3, 6, nz // 2, nz // 2 + 3, bld.G.materials[0], averaging=False)
    bld.add_waveform('ricker', 1, 20e9, 'w1')
    bld.add_waveform('gaussiandot', 2.0, 15e9, 'w2')
    c = (nx // 2, ny // 2, nz // 2)
    bld.add_hertzian_dipole('z', c, 'w1')
    bld.add_hertzian_dipole('x', (c[0] - 3, c[1] + 2, c[2] - 1), 'w2', start=5e-12, stop=40e-12)
    bld.add_hertzian_dipole('z', c, 'w2')          # two sources on one edge
    bld.add_magnetic_dipole('y', (c[0] + 2, c[1] - 3, c[2]), 'w1')
    bld.add_voltage_source('y', (c[0] - 4, c[1] - 2, c[2] + 2), 50.0, 'w2')
    bld.add_voltage_source('x', (c[0] + 4, c[1] + 3, c[2] - 2), 0.0, 'w1')
    for r in range(5):
        bld.add_rx((int(rng.integers(pml, nx - pml)), int(rng.integers(pml, ny - pml)),
                    int(rng.integers(pml, nz - pml))))
    bld.add_rx(c)
    if snapshots:
        bld.add_snapshot(0, 0, 0, nx, ny, nz, 1, 1, 1, iterations // 2)
        bld.add_snapshot(2, 3, 1, 2 + 2 * 5, 3 + 3 * 4, 1 + 2 * 6, 2, 3, 2, iterations - 1)
    return bld.finalise()


def plane_model(seed=0, axis=2, size=(30, 24), iterations=40, pml=4, order=1, formulation='HORIPML',
                poles=0, kind='debye', nmat=7, snapshots=True):
    """A 2-D (one cell along `axis`) companion of rich_model: random materials,
    a PEC block, PML on the four in-plane faces, Hertzian, voltage and
    magnetic sources in the plane, receivers and snapshots."""
    rng = np.random.default_rng(seed)
    shape = [0, 0, 0]
    inplane = [a for a in range(3) if a != axis]
    shape[axis] = 1
    shape[inplane[0]], shape[inplane[1]] = size
    nx, ny, nz = shape
    cfs = None
    if order == 2:
        a, b = CFS(), CFS()
        b.alpha.max = 0.05
        b.kappa.scalingprofile, b.kappa.min, b.kappa.max = 'quadratic', 1, 3
        cfs = [a, b]
    bld = ModelBuilder(nx, ny, nz, 1e-3, 1e-3, 1e-3, iterations, pml_cells=pml,
                       pml_formulation=formulation, cfs=cfs)
    assert bld.G.mode == '2D TM' + 'xyz'[axis]
    for n in range(nmat):
        m = bld.add_material(rng.uniform(1, 9), rng.uniform(0, 0.05), rng.uniform(1, 2), rng.uniform(0, 1e3),
                             'm{}'.format(n))
        if poles and n % 2 == 0:
            if kind == 'debye':
                add_dispersion(m, 'debye', list(rng.uniform(0.5, 5, poles)), list(10 ** rng.uniform(-11, -9, poles)))
            elif kind == 'lorentz':
                add_dispersion(m, 'lorentz', list(rng.uniform(0.5, 5, poles)), list(rng.uniform(1e9, 5e9, poles)),
                               list(rng.uniform(1e8, 5e8, poles)))
            else:
                add_dispersion(m, 'drude', None, list(rng.uniform(1e9, 3e9, poles)), list(rng.uniform(1e9, 5e9, poles)))
    mask = rng.random((nx, ny, nz)) < 0.6
    ids = rng.integers(2, 2 + nmat, (nx, ny, nz)).astype(np.uint32)
    bld.paint_voxels(mask, ids)

    def cell(u, v):
        c = [0, 0, 0]
        c[inplane[0]], c[inplane[1]] = u, v
        return tuple(c)

    lo, hi = [0, 0, 0], [nx, ny, nz]
    lo[inplane[0]], hi[inplane[0]] = size[0] // 2 - 3, size[0] // 2
    lo[inplane[1]], hi[inplane[1]] = 2, 5
    bld.box(lo[0], hi[0], lo[1], hi[1], lo[2], hi[2], bld.G.materials[0], averaging=False)
    bld.add_waveform('ricker', 1, 20e9, 'w1')
    bld.add_waveform('gaussiandot', 2.0, 15e9, 'w2')
    pol = 'xyz'[axis]
    u0, v0 = size[0] // 2, size[1] // 2
    bld.add_hertzian_dipole(pol, cell(u0, v0), 'w1')
    bld.add_hertzian_dipole(pol, cell(u0 - 3, v0 + 2), 'w2', start=5e-12, stop=40e-12)
    bld.add_magnetic_dipole('xyz'[inplane[0]], cell(u0 + 2, v0 - 3), 'w1')
    bld.add_voltage_source(pol, cell(u0 - 4, v0 - 2), 50.0, 'w2')
    bld.add_voltage_source(pol, cell(u0 + 4, v0 + 3), 0.0, 'w1')
    for r in range(4):
        bld.add_rx(cell(int(rng.integers(pml, size[0] - pml)), int(rng.integers(pml, size[1] - pml))))
    bld.add_rx(cell(u0, v0))
    if snapshots:
        bld.add_snapshot(0, 0, 0, nx, ny, nz, 1, 1, 1, iterations // 2)
    return bld.finalise()
