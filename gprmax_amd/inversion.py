"""Gradients of a trace misfit with respect to per-cell permittivity and
conductivity, through the adjoint of the FDTD solve on the GPU.

The forward model is the reference's solver loop (gprMax/model_build_run.py:
424-470) with every cell its own material, as a model painted voxel by voxel
without averaging: E edge (c, i, j, k) takes the coefficients of cell
(i, j, k) (gprMax/materials.py:178-201), mu_r = 1, no magnetic loss, a
10-cell (by default) HORIPML with one CFS term on every face, z-directed
Hertzian dipoles and Ez receivers, optionally with fixed Debye poles per
cell (``set_debye``).  ``gpx_inv_backward`` returns dL/dCA,
dL/dCB and dL/dIEA per cell; this module applies the chain rule to
dL/d(eps_r) and dL/d(sigma).

    grid = InversionGrid((160, 160, 120), 2e-3, 2000, K=64)
    loss = misfit(er, se, grid, sources, receivers, observed)   # torch tensors
    loss.backward()

Cells inside a PML slab get zero gradient: their coefficient also enters the
PML correction, which the adjoint propagates through but does not
differentiate.
"""

import ctypes

import numpy as np

from .constants import e0
from .hip import library
from .pml import CFS, PML, DIRECTION_CODE
from .utilities import GeneralError, courant_dt
from .waveforms import Waveform


class GpxInvDesc(ctypes.Structure):
    _fields_ = [('nx', ctypes.c_int), ('ny', ctypes.c_int), ('nz', ctypes.c_int),
                ('iterations', ctypes.c_int), ('K', ctypes.c_int), ('device', ctypes.c_int),
                ('d', ctypes.c_float), ('dt', ctypes.c_float)]


_P, _I, _F = ctypes.c_void_p, ctypes.c_int, ctypes.c_float
SIGNATURES = {
    'gpx_inv_create': ([ctypes.POINTER(GpxInvDesc), ctypes.POINTER(_P)], _I),
    'gpx_inv_destroy': ([_P], _I),
    'gpx_inv_last_error': ([_P], ctypes.c_char_p),
    'gpx_inv_add_pml': ([_P] + [_I] * 7 + [_P, _P], _I),
    'gpx_inv_set_model': ([_P, _P, _P, _P], _I),
    'gpx_inv_set_dispersive': ([_P, _I, _P, _P, _P], _I),
    'gpx_inv_forward': ([_P, _I, _I, _I, _F, _P, _I, _P, _P], _I),
    'gpx_inv_backward': ([_P, _P, _P, _P, _P], _I),
    'gpx_inv_set_keep': ([_P, ctypes.c_double], _I),
    'gpx_inv_timings': ([_P, ctypes.POINTER(_F), ctypes.POINTER(_F)], _I),
}


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class _PmlGrid:
    """The attributes gprmax_amd.pml.PML reads."""

    def __init__(self, nx, ny, nz, d, dt):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.dx = self.dy = self.dz = d
        self.dt = dt
        self.cfs = [CFS()]
        self.pmlformulation = 'HORIPML'
        self.floattype = np.float32


class InversionGeometry:
    """Geometry, time step and PML coefficients of the differentiable model.

    er_pml is the relative permittivity used for the PML's sigma max (the
    reference averages the slab's first layer, gprMax/pml.py:412-425); with a
    homogeneous start model it is that model's value.
    """

    def __init__(self, shape, d, iterations, pml=10, er_pml=1.0, K=64):
        self.nx, self.ny, self.nz = (int(v) for v in shape)
        self.d = float(d)
        self.dt = courant_dt(self.nx, self.ny, self.nz, d, d, d)[0]
        self.iterations = int(iterations)
        self.pml = int(pml)
        self.K = int(K)
        G = _PmlGrid(self.nx, self.ny, self.nz, self.d, self.dt)
        self.slabs = []
        t, nx, ny, nz = self.pml, self.nx, self.ny, self.nz
        boxes = [('xminus', 0, t, 0, ny, 0, nz), ('yminus', 0, nx, 0, t, 0, nz), ('zminus', 0, nx, 0, ny, 0, t),
                 ('xplus', nx - t, nx, 0, ny, 0, nz), ('yplus', 0, nx, ny - t, ny, 0, nz),
                 ('zplus', 0, nx, 0, ny, nz - t, nz)]
        if t > 0:
            for direction, xs, xf, ys, yf, zs, zf in boxes:
                p = PML(G, ID=direction, direction=direction, xs=xs, xf=xf, ys=ys, yf=yf, zs=zs, zf=zf)
                p.calculate_update_coeffs(er_pml, 1.0, G)
                self.slabs.append(p)
        m = np.zeros((nx, ny, nz), dtype=bool)
        if t > 0:
            m[t + 1:nx - t, t + 1:ny - t, t + 1:nz - t] = True
        else:
            m[...] = True
        self.interior = m   # cells whose every edge lies outside the PML slabs
        self.debye = None

    def set_debye(self, deltaer, tau):
        """Debye poles per cell, fixed during the inversion: deltaer and tau
        [P][nx][ny][nz] (deltaer 0 where a cell has fewer poles).  Per pole
        and cell, float64, the coefficients of gprMax/materials.py:92-143:
        w = deltaer / tau, q = -1 / tau, eqt = exp(q dt), eqt2 = exp(q dt / 2),
        zt = (w / q)(1 - eqt) / dt, zt2 = (w / q)(1 - eqt2); the update uses
        ee = e0 eqt2, eqt and zt, and EA gains -(e0 / dt) sum zt2 (Z).
        None removes the poles."""
        if deltaer is None:
            self.debye = None
            return
        de = np.asarray(deltaer, dtype=np.float64)
        tau = np.asarray(tau, dtype=np.float64)
        if de.ndim != 4 or de.shape != tau.shape or de.shape[1:] != (self.nx, self.ny, self.nz):
            raise GeneralError('Debye parameters must be [P][nx][ny][nz]')
        if de.shape[0] > 4 or np.any(tau <= 0):
            raise GeneralError('at most 4 Debye poles, with positive relaxation times')
        dt = self.dt
        w, q = de / tau, -1.0 / tau
        eqt, eqt2 = np.exp(q * dt), np.exp(q * (dt / 2))
        zt = (w / q) * (1 - eqt) / dt
        zt2 = (w / q) * (1 - eqt2)
        self.debye = {'ee': e0 * eqt2, 'eqt': eqt, 'zt': zt, 'Z': (e0 / dt) * zt2.sum(axis=0)}

    def waveform(self, freq=900e6, kind='ricker', amp=1.0):
        w = Waveform()
        w.type, w.amp, w.freq = kind, amp, freq
        w.calculate_coefficients()
        return np.array([w.calculate_value(n * self.dt, self.dt) for n in range(self.iterations)], dtype=np.float32)

    def coefficients(self, er, se):
        """CA, CB, IEA (float32) and their derivatives (float64) per cell."""
        er = np.asarray(er, dtype=np.float64)
        se = np.asarray(se, dtype=np.float64)
        a = e0 / self.dt
        EA = a * er + 0.5 * se
        EB = a * er - 0.5 * se
        if self.debye is not None:
            EA = EA - self.debye['Z']
            EB = EB - self.debye['Z']
        CA, IEA = EB / EA, 1.0 / EA
        CB = (1.0 / self.d) * IEA
        EA2 = EA * EA
        der = {'CA': a * (EA - EB) / EA2, 'IEA': -a / EA2}
        dse = {'CA': -0.5 * (EA + EB) / EA2, 'IEA': -0.5 / EA2}
        der['CB'] = der['IEA'] / self.d
        dse['CB'] = dse['IEA'] / self.d
        f32 = [np.ascontiguousarray(x, dtype=np.float32) for x in (CA, CB, IEA)]
        return f32, der, dse


class InversionGrid(InversionGeometry):
    """InversionGeometry plus the device context that runs the forward and
    adjoint solves."""

    def __init__(self, shape, d, iterations, pml=10, er_pml=1.0, K=64, device=0):
        super().__init__(shape, d, iterations, pml, er_pml, K)
        nx, ny, nz = self.nx, self.ny, self.nz
        self.lib = library()
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(self.lib, name)
            fn.argtypes = args
            fn.restype = res
        desc = GpxInvDesc(nx, ny, nz, self.iterations, self.K, device, self.d, self.dt)
        v = ctypes.c_void_p()
        rc = self.lib.gpx_inv_create(ctypes.byref(desc), ctypes.byref(v))
        if rc:
            raise GeneralError('gpx_inv_create failed ({}): {}'.format(rc, self.lib.gpx_inv_last_error(None).decode()))
        self.v = v
        for p in self.slabs:
            eR = np.ascontiguousarray(np.concatenate([p.ERA[0], p.ERB[0], p.ERE[0], p.ERF[0]]), dtype=np.float32)
            hR = np.ascontiguousarray(np.concatenate([p.HRA[0], p.HRB[0], p.HRE[0], p.HRF[0]]), dtype=np.float32)
            self._check(self.lib.gpx_inv_add_pml(self.v, DIRECTION_CODE[p.direction], p.xs, p.xf, p.ys, p.yf,
                                                 p.zs, p.zf, _ptr(eR), _ptr(hR)))

    def _check(self, rc):
        if rc:
            raise GeneralError('inversion error {}: {}'.format(rc, self.lib.gpx_inv_last_error(self.v).decode()))

    def close(self):
        if getattr(self, 'v', None):
            self.lib.gpx_inv_destroy(self.v)
            self.v = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- solves --------------------------------------------------------------
    def set_debye(self, deltaer, tau):
        super().set_debye(deltaer, tau)
        if self.debye is None:
            self._check(self.lib.gpx_inv_set_dispersive(self.v, 0, None, None, None))
            return
        f32 = [np.ascontiguousarray(self.debye[k], dtype=np.float32) for k in ('ee', 'eqt', 'zt')]
        self._check(self.lib.gpx_inv_set_dispersive(self.v, f32[0].shape[0], *(_ptr(x) for x in f32)))

    def set_model(self, er, se):
        (CA, CB, IEA), der, dse = self.coefficients(er, se)
        self._check(self.lib.gpx_inv_set_model(self.v, _ptr(CA), _ptr(CB), _ptr(IEA)))
        self._derivs = (der, dse)

    def forward(self, source, receivers, waves):
        """Ez traces [nrx][iterations] for a z-directed Hertzian dipole on
        Ez edge `source` (cell-length dipole)."""
        rx = np.ascontiguousarray(receivers, dtype=np.int32).reshape(-1, 3)
        waves = np.ascontiguousarray(waves, dtype=np.float32)
        traces = np.zeros((len(rx), self.iterations), dtype=np.float32)
        sfac = np.float32(self.d * (1.0 / (self.d * self.d * self.d)))
        i, j, k = (int(v) for v in source)
        self._check(self.lib.gpx_inv_forward(self.v, i, j, k, float(sfac), _ptr(waves), len(rx), _ptr(rx),
                                             _ptr(traces)))
        return traces

    def backward(self, resid):
        """dL/d(eps_r), dL/d(sigma) per cell for dL/d(trace) = resid, after
        the matching forward(); zero inside the PML."""
        resid = np.ascontiguousarray(resid, dtype=np.float32)
        g = [np.zeros((self.nx, self.ny, self.nz), dtype=np.float32) for _ in range(3)]
        self._check(self.lib.gpx_inv_backward(self.v, _ptr(resid), *(_ptr(x) for x in g)))
        der, dse = self._derivs
        gCA, gCB, gIEA = (x.astype(np.float64) for x in g)
        ger = gCA * der['CA'] + gCB * der['CB'] + gIEA * der['IEA']
        gse = gCA * dse['CA'] + gCB * dse['CB'] + gIEA * dse['IEA']
        ger[~self.interior] = 0
        gse[~self.interior] = 0
        return ger, gse

    def set_keep(self, nbytes):
        """Let the forward keep the stored states of its last checkpoint
        segments in up to `nbytes` of device memory, so that the backward
        skips their recompute (0, the default: recompute every segment).
        The gradient does not change."""
        self._check(self.lib.gpx_inv_set_keep(self.v, float(nbytes)))

    def keep_bytes_all(self):
        """Device bytes that keep every segment: per iteration 6 (9 with
        poles) floats per grid point, rounded up to a multiple of 4 floats."""
        per = 9 if self.debye is not None else 6
        floats = (per * (self.nx + 1) * (self.ny + 1) * (self.nz + 1) + 3) // 4 * 4
        return 4.0 * floats * self.K * (-(-self.iterations // self.K))

    def timings(self):
        f, b = ctypes.c_float(), ctypes.c_float()
        self._check(self.lib.gpx_inv_timings(self.v, ctypes.byref(f), ctypes.byref(b)))
        return f.value, b.value


def misfit_and_gradient(grid, er, se, sources, receivers, observed, waves):
    """L = 0.5 sum over sources, receivers and iterations of (trace - observed)^2
    and its gradient with respect to er and se (numpy, float64)."""
    grid.set_model(er, se)
    loss = 0.0
    ger = np.zeros((grid.nx, grid.ny, grid.nz))
    gse = np.zeros_like(ger)
    for s, src in enumerate(sources):
        tr = grid.forward(src, receivers, waves)
        r = tr.astype(np.float64) - observed[s]
        loss += 0.5 * float(np.sum(r * r))
        a, b = grid.backward(r.astype(np.float32))
        ger += a
        gse += b
    return loss, ger, gse


def make_misfit_function():
    """torch.autograd.Function for L(er, se); imported lazily so the module
    loads without torch."""
    import torch

    class Misfit(torch.autograd.Function):
        @staticmethod
        def forward(ctx, er, se, grid, sources, receivers, observed, waves):
            loss, ger, gse = misfit_and_gradient(grid, er.detach().cpu().numpy(), se.detach().cpu().numpy(),
                                                 sources, receivers, observed, waves)
            ctx.save_for_backward(torch.from_numpy(ger).to(er), torch.from_numpy(gse).to(se))
            return torch.tensor(loss, dtype=er.dtype, device=er.device)

        @staticmethod
        def backward(ctx, g):
            ger, gse = ctx.saved_tensors
            return g * ger, g * gse, None, None, None, None, None

    return Misfit


def misfit(er, se, grid, sources, receivers, observed, waves):
    """Differentiable trace misfit (torch tensors er, se of shape (nx, ny, nz))."""
    return make_misfit_function().apply(er, se, grid, sources, receivers, observed, waves)
