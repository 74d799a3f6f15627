"""Float64 pure-PyTorch forward of the differentiable model -- TEST
INFRASTRUCTURE, the autograd reference for the GPU adjoint.

The same discrete scheme as gprmax_amd/csrc/adjoint.hip, written with
whole-array slices (the update ranges of gprMax/fields_updates_ext.pyx:81-107
and 404-412, the HORIPML order-1 slab corrections of
gprMax/pml_updates/pml_updates_*_HORIPML_ext.pyx), every update out of place
so torch.autograd differentiates the whole time loop.  Meant for small grids.
"""

import numpy as np
import torch

from gprmax_amd.constants import e0, m0

_E_PAIRS = {0: ((1, 2, -1), (2, 1, +1)), 1: ((0, 2, +1), (2, 0, -1)), 2: ((0, 1, -1), (1, 0, +1))}
_H_PAIRS = {0: ((1, 2, +1), (2, 1, -1)), 1: ((0, 2, -1), (2, 0, +1)), 2: ((0, 1, +1), (1, 0, -1))}


def coefficients(er, se, dt, d, Z=None):
    """CA, CB, IEA of gprMax/materials.py:120-143; Z = (e0/dt) sum Re(zt2)
    per cell for cells with Debye poles."""
    a = e0 / dt
    EA = a * er + 0.5 * se
    EB = a * er - 0.5 * se
    if Z is not None:
        EA = EA - Z
        EB = EB - Z
    IEA = 1.0 / EA
    return EB / EA, (1.0 / d) * IEA, IEA


# Update ranges of Ex, Ey, Ez (gprMax/fields_updates_ext.pyx:81-107); an
# edge takes the coefficients of the cell of the same index.
def _e_ranges(nx, ny, nz):
    return [(slice(0, nx), slice(1, ny), slice(1, nz)),
            (slice(1, nx), slice(0, ny), slice(1, nz)),
            (slice(1, nx), slice(1, ny), slice(0, nz))]


def _slab_geometry(pml, field):
    """Index arrays of the corrected cells, of their neighbours along the
    normal and the profile slices, as gprmax_amd.pml/oracle do it."""
    axis = {'x': 0, 'y': 1, 'z': 2}[pml.direction[0]]
    minus = pml.direction.endswith('minus')
    starts = [pml.xs, pml.ys, pml.zs]
    ends = [pml.xf, pml.yf, pml.zf]
    t = ends[axis] - starts[axis]
    p = np.arange(t)
    if minus:
        along = ends[axis] - p if field == 0 else ends[axis] - (p + 1)
    else:
        along = starts[axis] + p
    idx = [np.arange(starts[n], ends[n]) for n in range(3)]
    idx[axis] = along
    nbr = [a.copy() for a in idx]
    nbr[axis] = along + (-1 if field == 0 else 1)
    shape = [1, 1, 1]
    shape[axis] = t
    return axis, np.ix_(*idx), np.ix_(*nbr), shape, [ends[n] - starts[n] for n in range(3)]


def forward_traces(er, se, grid, source, receivers, waves):
    """Ez traces (torch, float64) for one z-directed Hertzian dipole.

    grid: a gprmax_amd.inversion.InversionGrid (geometry, dt and PML tables);
    er, se: float64 tensors (nx, ny, nz)."""
    nx, ny, nz, d, dt = grid.nx, grid.ny, grid.nz, grid.d, grid.dt
    f64 = torch.float64
    deb = getattr(grid, 'debye', None)
    if deb is not None:
        Zt = torch.tensor(deb['Z'])
        dcoef = {k: torch.tensor(deb[k]) for k in ('ee', 'eqt', 'zt')}
        P = deb['ee'].shape[0]
    CA, CB, IEA = coefficients(er, se, dt, d, Zt if deb is not None else None)
    rng = _e_ranges(nx, ny, nz)
    if deb is not None:   # T of each E component on its update range, per pole
        shapes = [(nx, ny - 1, nz - 1), (nx - 1, ny, nz - 1), (nx - 1, ny - 1, nz)]
        T = [[torch.zeros(sh, dtype=f64) for _ in range(P)] for sh in shapes]
    DB = (1.0 / d) * (dt / m0)
    cH = dt / m0
    shp = (nx + 1, ny + 1, nz + 1)
    E = [torch.zeros(shp, dtype=f64) for _ in range(3)]
    H = [torch.zeros(shp, dtype=f64) for _ in range(3)]
    geo = []
    for p in grid.slabs:
        ge = _slab_geometry(p, 0)
        gh = _slab_geometry(p, 1)
        R = {k: torch.tensor(getattr(p, k)[0].astype(np.float64)) for k in ('ERA', 'ERB', 'ERE', 'ERF',
                                                                           'HRA', 'HRB', 'HRE', 'HRF')}
        n = ge[4]
        phiE = [torch.zeros(tuple(n), dtype=f64), torch.zeros(tuple(n), dtype=f64)]
        phiH = [torch.zeros(tuple(n), dtype=f64), torch.zeros(tuple(n), dtype=f64)]
        geo.append((ge, gh, R, phiE, phiH))
    si, sj, sk = source
    sfac = d * (1.0 / (d * d * d))
    rx = np.asarray(receivers, dtype=np.int64).reshape(-1, 3)
    traces = []
    for n_it in range(grid.iterations):
        traces.append(E[2][rx[:, 0], rx[:, 1], rx[:, 2]])
        Ex, Ey, Ez = E
        Hx, Hy, Hz = H
        nHx, nHy, nHz = Hx.clone(), Hy.clone(), Hz.clone()
        nHx[1:nx + 1, 0:ny, 0:nz] = (1.0 * Hx[1:nx + 1, 0:ny, 0:nz]
                                     - DB * (Ez[1:nx + 1, 1:ny + 1, 0:nz] - Ez[1:nx + 1, 0:ny, 0:nz])
                                     + DB * (Ey[1:nx + 1, 0:ny, 1:nz + 1] - Ey[1:nx + 1, 0:ny, 0:nz]))
        nHy[0:nx, 1:ny + 1, 0:nz] = (1.0 * Hy[0:nx, 1:ny + 1, 0:nz]
                                     - DB * (Ex[0:nx, 1:ny + 1, 1:nz + 1] - Ex[0:nx, 1:ny + 1, 0:nz])
                                     + DB * (Ez[1:nx + 1, 1:ny + 1, 0:nz] - Ez[0:nx, 1:ny + 1, 0:nz]))
        nHz[0:nx, 0:ny, 1:nz + 1] = (1.0 * Hz[0:nx, 0:ny, 1:nz + 1]
                                     - DB * (Ey[1:nx + 1, 0:ny, 1:nz + 1] - Ey[0:nx, 0:ny, 1:nz + 1])
                                     + DB * (Ex[0:nx, 1:ny + 1, 1:nz + 1] - Ex[0:nx, 0:ny, 1:nz + 1]))
        nH = [nHx, nHy, nHz]
        for (ge, gh, R, phiE, phiH) in geo:
            axis, cells, nbrs, shape, n = gh
            for w, (comp, src, sign) in enumerate(_H_PAIRS[axis]):
                dd = (E[src][nbrs] - E[src][cells]) / d
                ph = phiH[w]
                corr = cH * ((R['HRA'].reshape(shape) - 1.0) * dd + R['HRB'].reshape(shape) * ph)
                phiH[w] = R['HRE'].reshape(shape) * ph - R['HRF'].reshape(shape) * dd
                upd = nH[comp].clone()
                upd[cells] = nH[comp][cells] + corr if sign > 0 else nH[comp][cells] - corr
                nH[comp] = upd
        Hx, Hy, Hz = nH
        nEx, nEy, nEz = Ex.clone(), Ey.clone(), Ez.clone()
        nEx[0:nx, 1:ny, 1:nz] = (CA[0:nx, 1:ny, 1:nz] * Ex[0:nx, 1:ny, 1:nz]
                                 + CB[0:nx, 1:ny, 1:nz] * (Hz[0:nx, 1:ny, 1:nz] - Hz[0:nx, 0:ny - 1, 1:nz])
                                 - CB[0:nx, 1:ny, 1:nz] * (Hy[0:nx, 1:ny, 1:nz] - Hy[0:nx, 1:ny, 0:nz - 1]))
        nEy[1:nx, 0:ny, 1:nz] = (CA[1:nx, 0:ny, 1:nz] * Ey[1:nx, 0:ny, 1:nz]
                                 + CB[1:nx, 0:ny, 1:nz] * (Hx[1:nx, 0:ny, 1:nz] - Hx[1:nx, 0:ny, 0:nz - 1])
                                 - CB[1:nx, 0:ny, 1:nz] * (Hz[1:nx, 0:ny, 1:nz] - Hz[0:nx - 1, 0:ny, 1:nz]))
        nEz[1:nx, 1:ny, 0:nz] = (CA[1:nx, 1:ny, 0:nz] * Ez[1:nx, 1:ny, 0:nz]
                                 + CB[1:nx, 1:ny, 0:nz] * (Hy[1:nx, 1:ny, 0:nz] - Hy[0:nx - 1, 1:ny, 0:nz])
                                 - CB[1:nx, 1:ny, 0:nz] * (Hx[1:nx, 1:ny, 0:nz] - Hx[1:nx, 0:ny - 1, 0:nz]))
        nE = [nEx, nEy, nEz]
        if deb is not None:
            # dispersive part A (gprMax/fields_updates_ext.pyx:113-179):
            # phi from T^n, T advanced with E^n, E reduced by IEA phi
            Tn = []
            for c, rc in enumerate(rng):
                phi = 0.0
                Tc = []
                for p in range(P):
                    phi = phi + dcoef['ee'][p][rc] * T[c][p]
                    Tc.append(dcoef['eqt'][p][rc] * T[c][p] + dcoef['zt'][p][rc] * E[c][rc])
                Tn.append(Tc)
                upd = nE[c].clone()
                upd[rc] = nE[c][rc] - IEA[rc] * phi
                nE[c] = upd
        for (ge, gh, R, phiE, phiH) in geo:
            axis, cells, nbrs, shape, n = ge
            c = IEA[cells]
            for w, (comp, src, sign) in enumerate(_E_PAIRS[axis]):
                dd = (nH[src][cells] - nH[src][nbrs]) / d
                ph = phiE[w]
                corr = c * ((R['ERA'].reshape(shape) - 1.0) * dd + R['ERB'].reshape(shape) * ph)
                phiE[w] = R['ERE'].reshape(shape) * ph - R['ERF'].reshape(shape) * dd
                upd = nE[comp].clone()
                upd[cells] = nE[comp][cells] + corr if sign > 0 else nE[comp][cells] - corr
                nE[comp] = upd
        ez = nE[2].clone()
        ez[si, sj, sk] = nE[2][si, sj, sk] - IEA[si, sj, sk] * float(waves[n_it]) * sfac
        nE[2] = ez
        if deb is not None:   # dispersive part B (fields_updates_ext.pyx:182-235)
            T = [[Tn[c][p] - dcoef['zt'][p][rc] * nE[c][rc] for p in range(P)] for c, rc in enumerate(rng)]
        E = nE
        H = [Hx, Hy, Hz]
    return torch.stack(traces, dim=1)


def misfit(er, se, grid, sources, receivers, observed, waves):
    loss = 0.0
    for s, src in enumerate(sources):
        tr = forward_traces(er, se, grid, src, receivers, waves)
        r = tr - torch.as_tensor(observed[s], dtype=torch.float64)
        loss = loss + 0.5 * torch.sum(r * r)
    return loss
