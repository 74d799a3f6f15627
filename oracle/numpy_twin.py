"""NumPy slicing twin of the oracle kernels -- TEST INFRASTRUCTURE.

An independent formulation of gprMax/fields_updates_ext.pyx and
gprMax/pml_updates/*_ext.pyx written with whole-array slices instead of loops.
NumPy evaluates float32 products and sums without fusing them, so on the same
inputs it must agree with the C oracle (built without contraction) bit for bit.
Meant for small grids.
"""

import numpy as np


def update_electric(nx, ny, nz, coeffs, ID, Ex, Ey, Ez, Hx, Hy, Hz):
    """3-D E update over Ex i in [0,nx), Ey j in [0,ny), Ez k in [0,nz)
    (gprMax/fields_updates_ext.pyx:81-107)."""
    s = np.s_
    m = ID[0, 0:nx, 1:ny, 1:nz]
    c = coeffs[m]
    Ex[0:nx, 1:ny, 1:nz] = (c[..., 0] * Ex[0:nx, 1:ny, 1:nz]
                            + c[..., 2] * (Hz[0:nx, 1:ny, 1:nz] - Hz[0:nx, 0:ny - 1, 1:nz])
                            - c[..., 3] * (Hy[0:nx, 1:ny, 1:nz] - Hy[0:nx, 1:ny, 0:nz - 1]))
    m = ID[1, 1:nx, 0:ny, 1:nz]
    c = coeffs[m]
    Ey[1:nx, 0:ny, 1:nz] = (c[..., 0] * Ey[1:nx, 0:ny, 1:nz]
                            + c[..., 3] * (Hx[1:nx, 0:ny, 1:nz] - Hx[1:nx, 0:ny, 0:nz - 1])
                            - c[..., 1] * (Hz[1:nx, 0:ny, 1:nz] - Hz[0:nx - 1, 0:ny, 1:nz]))
    m = ID[2, 1:nx, 1:ny, 0:nz]
    c = coeffs[m]
    Ez[1:nx, 1:ny, 0:nz] = (c[..., 0] * Ez[1:nx, 1:ny, 0:nz]
                            + c[..., 1] * (Hy[1:nx, 1:ny, 0:nz] - Hy[0:nx - 1, 1:ny, 0:nz])
                            - c[..., 2] * (Hx[1:nx, 1:ny, 0:nz] - Hx[1:nx, 0:ny - 1, 0:nz]))
    del s


def update_magnetic(nx, ny, nz, coeffs, ID, Ex, Ey, Ez, Hx, Hy, Hz):
    """3-D H update (gprMax/fields_updates_ext.pyx:404-412)."""
    c = coeffs[ID[3, 1:nx + 1, 0:ny, 0:nz]]
    Hx[1:nx + 1, 0:ny, 0:nz] = (c[..., 0] * Hx[1:nx + 1, 0:ny, 0:nz]
                                - c[..., 2] * (Ez[1:nx + 1, 1:ny + 1, 0:nz] - Ez[1:nx + 1, 0:ny, 0:nz])
                                + c[..., 3] * (Ey[1:nx + 1, 0:ny, 1:nz + 1] - Ey[1:nx + 1, 0:ny, 0:nz]))
    c = coeffs[ID[4, 0:nx, 1:ny + 1, 0:nz]]
    Hy[0:nx, 1:ny + 1, 0:nz] = (c[..., 0] * Hy[0:nx, 1:ny + 1, 0:nz]
                                - c[..., 3] * (Ex[0:nx, 1:ny + 1, 1:nz + 1] - Ex[0:nx, 1:ny + 1, 0:nz])
                                + c[..., 1] * (Ez[1:nx + 1, 1:ny + 1, 0:nz] - Ez[0:nx, 1:ny + 1, 0:nz]))
    c = coeffs[ID[5, 0:nx, 0:ny, 1:nz + 1]]
    Hz[0:nx, 0:ny, 1:nz + 1] = (c[..., 0] * Hz[0:nx, 0:ny, 1:nz + 1]
                                - c[..., 1] * (Ey[1:nx + 1, 0:ny, 1:nz + 1] - Ey[0:nx, 0:ny, 1:nz + 1])
                                + c[..., 2] * (Ex[0:nx, 1:ny + 1, 1:nz + 1] - Ex[0:nx, 0:ny, 1:nz + 1]))


# PML: per axis, the two corrected components, their ID rows, the source
# component and the sign (+1 adds the correction, -1 subtracts it).
_E_PAIRS = {0: (('Ey', 1, 'Hz', -1), ('Ez', 2, 'Hy', +1)),
            1: (('Ex', 0, 'Hz', +1), ('Ez', 2, 'Hx', -1)),
            2: (('Ex', 0, 'Hy', -1), ('Ey', 1, 'Hx', +1))}
_H_PAIRS = {0: (('Hy', 4, 'Ez', +1), ('Hz', 5, 'Ey', -1)),
            1: (('Hx', 3, 'Ez', -1), ('Hz', 5, 'Ex', +1)),
            2: (('Hx', 3, 'Ey', +1), ('Hy', 4, 'Ex', -1))}


def pml_update(field, formulation, direction, box, coeffs, ID, F, Phi1, Phi2, RA, RB, RE, RF, d):
    """One slab, all cells at once.

    field 0 = E, 1 = H; formulation 0 = HORIPML, 1 = MRIPML; direction 0..5 as
    xminus, yminus, zminus, xplus, yplus, zplus; box = (xs, xf, ys, yf, zs, zf);
    F = dict of the six field arrays (modified in place).
    """
    axis, minus = direction % 3, direction < 3
    xs, xf, ys, yf, zs, zf = box
    starts, ends = [xs, ys, zs], [xf, yf, zf]
    t = ends[axis] - starts[axis]
    order = RA.shape[0]
    dt = F['Ex'].dtype.type
    p = np.arange(t)
    # cell coordinate along the normal for each profile index p
    if minus:
        along = ends[axis] - p if field == 0 else ends[axis] - (p + 1)
    else:
        along = starts[axis] + p
    idx = [np.arange(starts[n], ends[n]) for n in range(3)]
    idx[axis] = along
    cells = np.ix_(*idx)
    shift = [0, 0, 0]
    shift[axis] = -1 if field == 0 else 1
    nbr_idx = [a.copy() for a in idx]
    nbr_idx[axis] = along + shift[axis]
    nbrs = np.ix_(*nbr_idx)
    # per-profile coefficients broadcast along the normal axis
    shape = [1, 1, 1]
    shape[axis] = t

    def prof(a, o):
        return a[o].reshape(shape)

    pairs = (_E_PAIRS if field == 0 else _H_PAIRS)[axis]
    n = [ends[0] - starts[0], ends[1] - starts[1], ends[2] - starts[2]]
    for (comp, idrow, src, sign), Phi in zip(pairs, (Phi1, Phi2)):
        c = coeffs[ID[idrow][cells], 4]
        S = F[src]
        if field == 0:
            dd = (S[cells] - S[nbrs]) / dt(d)
        else:
            dd = (S[nbrs] - S[cells]) / dt(d)
        ph0 = Phi[0, :n[0], :n[1], :n[2]]
        if formulation == 0:
            if order == 1:
                RA01 = prof(RA, 0) - dt(1)
                corr = c * (RA01 * dd + prof(RB, 0) * ph0)
                new0 = prof(RE, 0) * ph0 - prof(RF, 0) * dd
                new1 = None
            else:
                ph1 = Phi[1, :n[0], :n[1], :n[2]]
                RA01 = prof(RA, 0) * prof(RA, 1) - dt(1)
                corr = c * (RA01 * dd + prof(RA, 1) * prof(RB, 0) * ph0 + prof(RB, 1) * ph1)
                new1 = prof(RE, 1) * ph1 - prof(RF, 1) * (prof(RA, 0) * dd + prof(RB, 0) * ph0)
                new0 = prof(RE, 0) * ph0 - prof(RF, 0) * dd
        else:
            if order == 1:
                IRA = dt(1) / prof(RA, 0)
                IRA1 = IRA - dt(1)
                RC0 = IRA * prof(RB, 0) * prof(RF, 0)
                corr = c * (IRA1 * dd - IRA * ph0)
                new0 = prof(RE, 0) * ph0 + RC0 * dd - RC0 * ph0
                new1 = None
            else:
                ph1 = Phi[1, :n[0], :n[1], :n[2]]
                IRA = dt(1) / (prof(RA, 0) + prof(RA, 1))
                IRA1 = IRA - dt(1)
                RC0 = IRA * prof(RF, 0)
                RC1 = IRA * prof(RF, 1)
                psi = prof(RB, 0) * ph0 + prof(RB, 1) * ph1
                corr = c * (IRA1 * dd - IRA * psi)
                new1 = prof(RE, 1) * ph1 + RC1 * (dd - psi)
                new0 = prof(RE, 0) * ph0 + RC0 * (dd - psi)
        F[comp][cells] = F[comp][cells] + corr if sign > 0 else F[comp][cells] - corr
        if new1 is not None:
            Phi[1, :n[0], :n[1], :n[2]] = new1
        Phi[0, :n[0], :n[1], :n[2]] = new0
