"""Synthetic workloads of the solve path, seeded and sized as the benchmark
configurations describe them.

* ``freespace_ricker``: 100^3 free space, 1 mm, 10-cell PML, z-Hertzian dipole
  at the centre driven by a 1.5 GHz Ricker, one receiver 20 cells along x,
  500 iterations.
* ``heterogeneous_halfspace``: 300x300x200 cells of 2 mm, 10-cell PML; the
  lower 100 z-layers are soil with a seeded per-cell material out of 256
  (er ~ U[4, 12], sigma ~ U[1e-3, 5e-2] S/m); free space above; a PEC
  cylinder of radius 15 cells along y, axis at x = 150, z = 60 cells;
  Hz-polarised magnetic dipole (Ricker 900 MHz) at (150, 150, 110); nine
  receivers at x = 110..190 step 10, y = 150, z = 110; 2000 iterations.
* ``dispersive_soil``: 400x400x300 cells of 2 mm, 20-cell PML; each cell is,
  with probability 0.3, one of 64 three-pole Debye materials (delta er ~
  U[0.5, 5], tau = 10^U[-11, -8] s, er_inf ~ U[4, 12]) and otherwise one of 192
  non-dispersive materials; z-Hertzian at the centre (Ricker 900 MHz); nine
  receivers; 3000 iterations.

Cell materials are painted rigidly (no averaging), as gprMax paints
#geometry_objects_read voxels (gprMax/geometry_primitives_ext.pyx:314-376), the
cylinder after the soil.  Seeds: numpy default_rng(1000 + k) for workload k.
"""

import numpy as np

from .grid import ModelBuilder
from .materials import add_dispersion


def freespace_ricker(iterations=500, n=100):
    b = ModelBuilder(n, n, n, 1e-3, 1e-3, 1e-3, iterations, pml_cells=10)
    b.add_waveform('ricker', 1, 1.5e9, 'ricker15')
    c = n // 2
    b.add_hertzian_dipole('z', (c, c, c), 'ricker15')
    b.add_rx((c + 20, c, c))
    G = b.finalise()
    G.title = 'freespace_ricker_{}'.format(n)
    return G


def _soil_materials(b, rng, count, name):
    mats = []
    er = rng.uniform(4, 12, count)
    se = rng.uniform(1e-3, 5e-2, count)
    for n in range(count):
        mats.append(b.add_material(er[n], se[n], 1, 0, '{}{}'.format(name, n)))
    return mats


def heterogeneous_halfspace(iterations=2000, shape=(300, 300, 200), soil_layers=100, seed=1002):
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    b = ModelBuilder(nx, ny, nz, 2e-3, 2e-3, 2e-3, iterations, pml_cells=10)
    mats = _soil_materials(b, rng, 256, 'soil')
    first = mats[0].numID
    ids = np.ones((nx, ny, nz), dtype=np.uint32)
    soil = np.zeros((nx, ny, nz), dtype=bool)
    soil[:, :, :soil_layers] = True
    ids[soil] = rng.integers(first, first + 256, int(soil.sum()), dtype=np.uint32)
    b.paint_voxels(soil, ids)
    # PEC cylinder along y: ((i + 0.5) dx - x1)^2 + ((k + 0.5) dz - z1)^2 <= r^2
    # (gprMax/geometry_primitives_ext.pyx:808-813)
    d = 2e-3
    x1, z1, r = (nx // 2) * d, (soil_layers - 40) * d, 15 * d
    i = (np.arange(nx) + 0.5) * d
    k = (np.arange(nz) + 0.5) * d
    disc = (i[:, None] - x1) ** 2 + (k[None, :] - z1) ** 2 <= r ** 2
    cyl = np.broadcast_to(disc[:, None, :], (nx, ny, nz))
    b.paint_voxels(np.ascontiguousarray(cyl), 0)
    b.add_waveform('ricker', 1, 900e6, 'ricker900')
    cx, cy, cz = nx // 2, ny // 2, soil_layers + 10
    b.add_magnetic_dipole('z', (cx, cy, cz), 'ricker900')
    for x in range(cx - 40, cx + 41, 10):
        b.add_rx((x, cy, cz))
    G = b.finalise()
    G.title = 'heterogeneous_halfspace_{}x{}x{}'.format(nx, ny, nz)
    return G


def dispersive_soil(iterations=3000, shape=(400, 400, 300), seed=1003, pml=20):
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    b = ModelBuilder(nx, ny, nz, 2e-3, 2e-3, 2e-3, iterations, pml_cells=pml)
    debye = _soil_materials(b, rng, 64, 'debye')
    for m in debye:
        add_dispersion(m, 'debye', list(rng.uniform(0.5, 5, 3)), list(10 ** rng.uniform(-11, -8, 3)))
    plain = _soil_materials(b, rng, 192, 'soil')
    ids = np.empty((nx, ny, nz), dtype=np.uint32)
    disp = rng.random((nx, ny, nz)) < 0.3
    nd = int(disp.sum())
    ids[disp] = rng.integers(debye[0].numID, debye[0].numID + 64, nd, dtype=np.uint32)
    ids[~disp] = rng.integers(plain[0].numID, plain[0].numID + 192, disp.size - nd, dtype=np.uint32)
    b.paint_voxels(np.ones((nx, ny, nz), dtype=bool), ids)
    b.add_waveform('ricker', 1, 900e6, 'ricker900')
    cx, cy, cz = nx // 2, ny // 2, nz // 2
    b.add_hertzian_dipole('z', (cx, cy, cz), 'ricker900')
    for x in range(cx - 40, cx + 41, 10):
        b.add_rx((x, cy, cz + 10))
    G = b.finalise()
    G.title = 'dispersive_soil_{}x{}x{}'.format(nx, ny, nz)
    return G


def bscan_trace(t, seed=1004, iterations=1500, sides=(128, 256)):
    """Trace t of the uneven B-scan: its own domain, each side drawn from
    `sides` cells at 1 mm, 10-cell PML, lower half soil with a per-cell random
    choice of 64 lossless materials (er ~ U[3, 15]), a z-directed Hertzian
    dipole and an Ez receiver 10 cells apart in x, 5 cells above the soil,
    stepped 2 cells per trace (wrapping inside the domain)."""
    rng = np.random.default_rng([seed, t])
    nx, ny, nz = (int(v) for v in rng.integers(sides[0], sides[1] + 1, 3))
    b = ModelBuilder(nx, ny, nz, 1e-3, 1e-3, 1e-3, iterations, pml_cells=10)
    er = rng.uniform(3, 15, 64)
    mats = [b.add_material(float(er[q]), 0, 1, 0, 'soil{}'.format(q)) for q in range(64)]
    soil = np.zeros((nx, ny, nz), dtype=bool)
    soil[:, :, :nz // 2] = True
    ids = np.ones((nx, ny, nz), dtype=np.uint32)
    ids[soil] = rng.integers(mats[0].numID, mats[0].numID + 64, int(soil.sum()), dtype=np.uint32)
    b.paint_voxels(soil, ids)
    b.add_waveform('ricker', 1, 1.5e9, 'ricker15')
    xs = 15 + (2 * t) % (nx - 40)
    zs = nz // 2 + 5
    b.add_hertzian_dipole('z', (xs, ny // 2, zs), 'ricker15')
    b.add_rx((xs + 10, ny // 2, zs))
    G = b.finalise()
    G.title = 'bscan_trace{}_{}x{}x{}'.format(t, nx, ny, nz)
    return G


WORKLOADS = {
    'freespace_ricker': freespace_ricker,
    'heterogeneous_halfspace': heterogeneous_halfspace,
    'dispersive_soil': dispersive_soil,
}


def algorithmic_bytes_per_step(G):
    """Canonical bytes per iteration: 96 Nc + 32 order N_pml + 24 P Nc.

    96 B per cell and step = two passes x (3 own-field reads + 3 writes + 3
    other-field reads + 3 uint32 IDs) x 4 B; 32 B x order per PML cell for the
    Phi read+write of both passes; 24 B x poles per cell for real T read+write
    of three components.  N_pml counts every slab cell (overlaps per slab).
    """
    nc = G.nx * G.ny * G.nz
    npml = sum(p.nx * p.ny * p.nz for p in G.pmls)
    order = len(G.cfs) if G.pmls else 0
    return 96 * nc + 32 * order * npml + 24 * G.maxpoles * nc


def pass_bytes(G, field):
    """Algorithmic bytes of one H (field=1) or E (field=0) pass."""
    nc = G.nx * G.ny * G.nz
    npml = sum(p.nx * p.ny * p.nz for p in G.pmls)
    order = len(G.cfs) if G.pmls else 0
    b = 48 * nc + 16 * order * npml
    if field == 0:
        b += 24 * G.maxpoles * nc
    return b


def kernel_bytes(G, field, box):
    """Algorithmic bytes of the (interior, boundary) kernels of one pass.

    The interior kernel updates the cells of `box` ([x0, x1, y0, y1, z0, z1)
    in field-array indices) at 48 B each (plus 24 B x poles for E) plus the
    Phi traffic of the PML cells inside the box; the boundary kernel does the
    rest, so the two add up to ``pass_bytes``.
    """
    x0, x1, y0, y1, z0, z1 = box
    lo = (x0, y0, z0)
    hi = (min(x1, G.nx), min(y1, G.ny), min(z1, G.nz))   # cells of the grid
    cells = max(0, hi[0] - lo[0]) * max(0, hi[1] - lo[1]) * max(0, hi[2] - lo[2])
    order = len(G.cfs) if G.pmls else 0
    pml_in = 0
    for p in G.pmls:
        a = (max(lo[0], p.xs), max(lo[1], p.ys), max(lo[2], p.zs))
        b = (min(hi[0], p.xf), min(hi[1], p.yf), min(hi[2], p.zf))
        pml_in += max(0, b[0] - a[0]) * max(0, b[1] - a[1]) * max(0, b[2] - a[2])
    per_cell = 48 + (24 * G.maxpoles if field == 0 else 0)
    inner = per_cell * cells + 16 * order * pml_in
    return inner, pass_bytes(G, field) - inner
