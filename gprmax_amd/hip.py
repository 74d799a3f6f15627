"""ctypes binding of libgprmax_hip.so (C ABI in include/gpx.h).

The library must have been built (``gprmax_amd.build.build()``); loading fails
loudly otherwise -- there is no CPU fallback on this path.
"""

import ctypes
import os

import numpy as np

from .pml import DIRECTION_CODE
from .sources import device_source_table, device_tl_args
from .receivers import rx_coords
from .utilities import GeneralError

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, 'libgprmax_hip.so')

PART = {'store_rx': 0, 'snapshot': 1, 'h_main': 2, 'h_pml': 3, 'h_src': 4,
        'e_main': 5, 'e_pml': 6, 'e_src': 7, 'e_disp_b': 8}
SRC_KIND = {'hertzian': 0, 'magnetic': 1, 'voltage': 2}


class GpxDesc(ctypes.Structure):
    _fields_ = [('nx', ctypes.c_int), ('ny', ctypes.c_int), ('nz', ctypes.c_int),
                ('iterations', ctypes.c_int), ('nmat', ctypes.c_int), ('maxpoles', ctypes.c_int),
                ('pml_formulation', ctypes.c_int), ('device', ctypes.c_int)]


# Exported symbols and their argument types (include/gpx.h).
_P, _I, _F = ctypes.c_void_p, ctypes.c_int, ctypes.c_float
SIGNATURES = {
    'gpx_create': ([ctypes.POINTER(GpxDesc), ctypes.POINTER(_P)], _I),
    'gpx_destroy': ([_P], _I),
    'gpx_last_error': ([_P], ctypes.c_char_p),
    'gpx_upload_id': ([_P, _P], _I),
    'gpx_upload_coeffs': ([_P, _P, _P], _I),
    'gpx_upload_dispersive': ([_P, _P], _I),
    'gpx_add_pml': ([_P, _I, _I, _I, _I, _I, _I, _I, _I, _F, _P, _P, ctypes.POINTER(_I)], _I),
    'gpx_set_sources': ([_P, _I, _I, _P, _P, _P], _I),
    'gpx_set_receivers': ([_P, _I, _P], _I),
    'gpx_add_transmission_line': ([_P] * 9 + [ctypes.POINTER(_I)], _I),
    'gpx_download_transmission_line': ([_P, _I, _P, _P, _P, _P], _I),
    'gpx_add_snapshot': ([_P] + [_I] * 10 + [ctypes.POINTER(_I)], _I),
    'gpx_run': ([_P, _I, _I], _I),
    'gpx_run_timed': ([_P, _I, _I, _I, ctypes.POINTER(_F), ctypes.POINTER(_F), ctypes.POINTER(_F)], _I),
    'gpx_run_profiled': ([_P, _I, _I, ctypes.POINTER(_F)], _I),
    'gpx_interior_cells': ([_P], ctypes.c_longlong),
    'gpx_interior_box': ([_P, ctypes.POINTER(_I)], _I),
    'gpx_part': ([_P, _I, _I, _I], _I),
    'gpx_set_launch': ([_P, _I, _I], _I),
    'gpx_set_step': ([_P, _I, _I], _I),
    'gpx_step_kind': ([_P], _I),
    'gpx_step_form': ([_P], _I),
    'gpx_step_choice': ([_P], _I),
    'gpx_synchronize': ([_P], _I),
    'gpx_download_rx': ([_P, _P], _I),
    'gpx_download_fields': ([_P, _P], _I),
    'gpx_upload_fields': ([_P, _P], _I),
    'gpx_download_pml_phi': ([_P, _I, _I, _P], _I),
    'gpx_upload_pml_phi': ([_P, _I, _I, _P], _I),
    'gpx_download_dispersive_T': ([_P, _P], _I),
    'gpx_upload_dispersive_T': ([_P, _P], _I),
    'gpx_download_snapshot': ([_P, _I, _P], _I),
    'gpx_device_info': ([_I, ctypes.c_char_p, _I, ctypes.POINTER(ctypes.c_size_t),
                         ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(_I)], _I),
    'gpx_memory_bytes': ([_P], ctypes.c_size_t),
    'gpx_batch_create': ([ctypes.POINTER(_P), _I, ctypes.POINTER(_P)], _I),
    'gpx_batch_destroy': ([_P], _I),
    'gpx_batch_last_error': ([_P], ctypes.c_char_p),
    'gpx_batch_set_shape': ([_P, _I, _I], _I),
    'gpx_batch_set_group_bytes': ([_P, ctypes.c_double], _I),
    'gpx_batch_tune': ([_P, _I, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_F)], _I),
    'gpx_batch_run': ([_P, _I, _I, _I, ctypes.POINTER(_F)], _I),
}

_lib = None


def library():
    """Load libgprmax_hip.so once and declare its signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBRARY):
        raise GeneralError('libgprmax_hip.so is not built; run gprmax_amd.build.build() '
                           '(the HIP path has no CPU fallback)')
    lib = ctypes.CDLL(LIBRARY)
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def grid_maxpoles(G):
    """Dispersive poles of grid G: ``G.maxpoles`` on this package's grid; the
    reference keeps it on its Material class and sizes
    ``updatecoeffsdispersive`` [nmat][3*maxpoles] by it."""
    if hasattr(G, 'maxpoles'):
        return int(G.maxpoles)
    d = getattr(G, 'updatecoeffsdispersive', None)
    return int(d.shape[1]) // 3 if d is not None and d.ndim == 2 else 0


def _arr(a, dtype):
    a = np.ascontiguousarray(a, dtype=dtype)
    return a, a.ctypes.data_as(ctypes.c_void_p)


class HipSolver:
    """One model resident on one GPU, built from a grid ``G``.

    Uploads ``G.ID``, the coefficient tables, the PML slabs (in ``G.pmls``
    order), the sources, receivers and snapshots; then steps on the device.
    """

    def __init__(self, G, device=0):
        self.lib = library()
        self.G = G
        if getattr(G, 'floattype', np.float32) != np.float32:
            raise GeneralError('the HIP solver computes in float32')
        self.maxpoles = grid_maxpoles(G)
        desc = GpxDesc(G.nx, G.ny, G.nz, G.iterations, len(G.updatecoeffsE), self.maxpoles,
                       0 if G.pmlformulation == 'HORIPML' else 1, device)
        ctx = ctypes.c_void_p()
        rc = self.lib.gpx_create(ctypes.byref(desc), ctypes.byref(ctx))
        if rc:
            raise GeneralError('gpx_create failed ({}): {}'.format(rc, self.lib.gpx_last_error(None).decode()))
        self.ctx = ctx
        self.device = device
        self.ntot = (G.nx + 1) * (G.ny + 1) * (G.nz + 1)
        self._check(self.lib.gpx_upload_id(ctx, _arr(G.ID, np.uint32)[1]))
        e, pe = _arr(G.updatecoeffsE, np.float32)
        h, ph = _arr(G.updatecoeffsH, np.float32)
        self._check(self.lib.gpx_upload_coeffs(ctx, pe, ph))
        if self.maxpoles > 0:
            d = np.ascontiguousarray(G.updatecoeffsdispersive, dtype=np.complex64)
            self._check(self.lib.gpx_upload_dispersive(ctx, d.ctypes.data_as(ctypes.c_void_p)))
        for pml in G.pmls:
            eR = np.ascontiguousarray(np.stack([pml.ERA, pml.ERB, pml.ERE, pml.ERF]), dtype=np.float32)
            hR = np.ascontiguousarray(np.stack([pml.HRA, pml.HRB, pml.HRE, pml.HRF]), dtype=np.float32)
            slab = ctypes.c_int()
            self._check(self.lib.gpx_add_pml(ctx, DIRECTION_CODE[pml.direction], pml.xs, pml.xf, pml.ys,
                                             pml.yf, pml.zs, pml.zf, len(pml.CFS), float(pml.d),
                                             eR.ctypes.data_as(ctypes.c_void_p),
                                             hR.ctypes.data_as(ctypes.c_void_p), ctypes.byref(slab)))
        for kind, sources in (('hertzian', G.hertziandipoles), ('magnetic', G.magneticdipoles),
                              ('voltage', G.voltagesources)):
            if sources:
                info, fac, waves = device_source_table(sources, G, kind)
                self._check(self.lib.gpx_set_sources(ctx, SRC_KIND[kind], len(sources),
                                                     _arr(info, np.int32)[1], _arr(fac, np.float32)[1],
                                                     _arr(waves, np.float32)[1]))
        if G.rxs:
            self._check(self.lib.gpx_set_receivers(ctx, len(G.rxs), _arr(rx_coords(G), np.int32)[1]))
        for tl in getattr(G, 'transmissionlines', []):
            args = device_tl_args(tl, G)
            idx = ctypes.c_int()
            self._check(self.lib.gpx_add_transmission_line(ctx, *(a.ctypes.data_as(ctypes.c_void_p) for a in args),
                                                           ctypes.byref(idx)))
        for snap in G.snapshots:
            idx = ctypes.c_int()
            self._check(self.lib.gpx_add_snapshot(ctx, snap.xs, snap.ys, snap.zs, snap.dx, snap.dy, snap.dz,
                                                  snap.nx, snap.ny, snap.nz, snap.time, ctypes.byref(idx)))

    # -- plumbing ----------------------------------------------------------
    def _check(self, rc):
        if rc:
            raise GeneralError('HIP solver error {}: {}'.format(rc, self.lib.gpx_last_error(self.ctx).decode()))

    def close(self):
        if getattr(self, 'ctx', None):
            self.lib.gpx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- stepping ----------------------------------------------------------
    def set_launch(self, xchunk=0, block=0):
        self._check(self.lib.gpx_set_launch(self.ctx, xchunk, block))

    def set_step(self, fused=4, tile_cells=0):
        """Step form of gpx_run: 0 the H and E passes, 1 the fused step, 2 the
        passes with interior and boundary kernels side by side, 3 the split
        step, 5 the split step with the H boundary pass beside the interior
        kernel, 6 the passes as one launch each over every cell, 7 the split
        step with the H boundary pass inside the interior kernel's launch, 4
        (the library's default) the fastest of 7, 6, 2, 5, 3 and 0 (of 6, 2
        and 0 where the split step does not apply), timed on the first run of
        at least 26 iterations.  All give the same bits."""
        self._check(self.lib.gpx_set_step(self.ctx, int(fused), int(tile_cells)))

    def step_kind(self):
        """Cells per fused block, or 0 when another form runs."""
        return int(self.lib.gpx_step_kind(self.ctx))

    def step_form(self):
        """The form gpx_run uses: 0 passes, 1 fused, 2 side by side, 3 split,
        5 split with the H boundary pass beside the interior kernel, 6 merged
        passes, 7 split with the H boundary pass in the interior launch."""
        return int(self.lib.gpx_step_form(self.ctx))

    def step_choice(self):
        """The default form's choice: -1 not timed yet, 0 passes, 1 split,
        2 split with the H boundary pass beside, 3 passes side by side, 4
        merged passes, 5 split with the H boundary pass in the interior
        launch."""
        return int(self.lib.gpx_step_choice(self.ctx))

    def run(self, it0=0, n=None):
        n = self.G.iterations - it0 if n is None else n
        self._check(self.lib.gpx_run(self.ctx, it0, n))

    def run_timed(self, it0=0, n=None, per_kernel=False):
        """Run and return (ms_total, ms_h_mean, ms_e_mean) from HIP events."""
        n = self.G.iterations - it0 if n is None else n
        t, h, e = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
        self._check(self.lib.gpx_run_timed(self.ctx, it0, n, int(per_kernel), ctypes.byref(t),
                                           ctypes.byref(h), ctypes.byref(e)))
        return t.value, h.value, e.value

    def run_profiled(self, it0, n):
        """Mean ms per iteration of each kernel: (H interior, H boundary, E
        interior, E boundary) for the passes; (interior, H boundary, 0, E
        boundary) for the split step; (step, 0, 0, 0) for the fused step."""
        ms = (ctypes.c_float * 4)()
        self._check(self.lib.gpx_run_profiled(self.ctx, it0, n, ms))
        return tuple(ms)

    def interior_cells(self):
        return int(self.lib.gpx_interior_cells(self.ctx))

    def interior_box(self):
        b = (ctypes.c_int * 6)()
        self._check(self.lib.gpx_interior_box(self.ctx, b))
        return tuple(b)

    def part(self, name, arg=0, it=0):
        self._check(self.lib.gpx_part(self.ctx, PART[name], arg, it))

    def synchronize(self):
        self._check(self.lib.gpx_synchronize(self.ctx))

    # -- state ---------------------------------------------------------------
    def download_fields(self):
        out = np.empty((6, self.G.nx + 1, self.G.ny + 1, self.G.nz + 1), dtype=np.float32)
        self._check(self.lib.gpx_download_fields(self.ctx, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def upload_fields(self, fields6):
        a, p = _arr(fields6, np.float32)
        if a.size != 6 * self.ntot:
            raise GeneralError('upload_fields expects 6 x (nx+1)(ny+1)(nz+1) values')
        self._check(self.lib.gpx_upload_fields(self.ctx, p))

    def download_transmission_line(self, n):
        """(Vtotal, Itotal) per iteration and the line's (voltage, current) now."""
        tl = self.G.transmissionlines[n]
        its = self.G.iterations
        vt, it_ = np.zeros(its, np.float32), np.zeros(its, np.float32)
        v, i = np.zeros(tl.nl, np.float32), np.zeros(tl.nl, np.float32)
        self._check(self.lib.gpx_download_transmission_line(self.ctx, n, *(a.ctypes.data_as(ctypes.c_void_p)
                                                                            for a in (vt, it_, v, i))))
        return vt, it_, v, i

    def download_rx(self):
        out = np.zeros((6, self.G.iterations, max(len(self.G.rxs), 1)), dtype=np.float32)
        if self.G.rxs:
            self._check(self.lib.gpx_download_rx(self.ctx, out.ctypes.data_as(ctypes.c_void_p)))
        return out[:, :, :len(self.G.rxs)]

    def download_phi(self, slab, which):
        pml = self.G.pmls[slab]
        out = np.empty(pml.phi_shapes()[which], dtype=np.float32)
        self._check(self.lib.gpx_download_pml_phi(self.ctx, slab, which, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def upload_phi(self, slab, which, values):
        pml = self.G.pmls[slab]
        a, p = _arr(values, np.float32)
        if a.shape != tuple(pml.phi_shapes()[which]):
            raise GeneralError('Phi shape mismatch')
        self._check(self.lib.gpx_upload_pml_phi(self.ctx, slab, which, p))

    def download_T(self):
        G = self.G
        out = np.empty((3, self.maxpoles, G.nx + 1, G.ny + 1, G.nz + 1), dtype=np.complex64)
        self._check(self.lib.gpx_download_dispersive_T(self.ctx, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def upload_T(self, T):
        a = np.ascontiguousarray(T, dtype=np.complex64)
        self._check(self.lib.gpx_upload_dispersive_T(self.ctx, a.ctypes.data_as(ctypes.c_void_p)))

    def download_snapshot(self, index):
        s = self.G.snapshots[index]
        out = np.empty((6, s.nx, s.ny, s.nz), dtype=np.float32)
        self._check(self.lib.gpx_download_snapshot(self.ctx, index, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def memory_bytes(self):
        return int(self.lib.gpx_memory_bytes(self.ctx))


def device_info(device=0):
    lib = library()
    name = ctypes.create_string_buffer(256)
    total, free, cus = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_int()
    rc = lib.gpx_device_info(device, name, 256, ctypes.byref(total), ctypes.byref(free), ctypes.byref(cus))
    if rc:
        raise GeneralError('gpx_device_info failed: {}'.format(lib.gpx_last_error(None).decode()))
    return {'name': name.value.decode(), 'totalmem': total.value, 'freemem': free.value, 'cus': cus.value}


class HipBatch:
    """Several HipSolver models advanced together: per iteration, four launches
    (the H and E passes, interior and boundary) each over all models, or one
    launch of the fused step over all models; optionally all iterations in one
    HIP graph."""

    def __init__(self, solvers):
        self.lib = library()
        self.solvers = list(solvers)
        arr = (ctypes.c_void_p * len(self.solvers))(*[s.ctx.value for s in self.solvers])
        b = ctypes.c_void_p()
        rc = self.lib.gpx_batch_create(arr, len(self.solvers), ctypes.byref(b))
        if rc:
            raise GeneralError('gpx_batch_create failed ({}): {}'.format(rc, self.lib.gpx_batch_last_error(None).decode()))
        self.b = b

    def _check(self, rc):
        if rc:
            raise GeneralError('HIP batch error {}: {}'.format(rc, self.lib.gpx_batch_last_error(self.b).decode()))

    def close(self):
        if getattr(self, 'b', None):
            self.lib.gpx_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_shape(self, tile_cells=0, xchunk=0):
        """tile_cells 0: the H and E passes; 256, 512, 1024: the fused step."""
        self._check(self.lib.gpx_batch_set_shape(self.b, int(tile_cells), int(xchunk)))

    def set_group_bytes(self, nbytes=0):
        """State bytes per group of models the passes advance together
        (0: 192 MiB)."""
        self._check(self.lib.gpx_batch_set_group_bytes(self.b, float(nbytes)))

    def tune(self, steps=3):
        """Pick the fastest launch shape; returns (tile_cells, xchunk, ms per
        iteration), tile_cells 0 for the passes.  Leaves every model at the
        zero state."""
        t, x, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_float()
        self._check(self.lib.gpx_batch_tune(self.b, int(steps), ctypes.byref(t), ctypes.byref(x), ctypes.byref(ms)))
        return t.value, x.value, ms.value

    def run(self, it0, n, graph=True):
        """Advance every model by n iterations from it0; returns device ms."""
        ms = ctypes.c_float()
        self._check(self.lib.gpx_batch_run(self.b, int(it0), int(n), int(bool(graph)), ctypes.byref(ms)))
        return ms.value


def _state_bytes(G):
    """Device bytes of a model's fields and material indices (six float32
    components and six uint16 indices per grid point), the measure the batch
    groups models by."""
    return 36.0 * (G.nx + 1) * (G.ny + 1) * (G.nz + 1)


def bscan_groups(grids, group_bytes=0):
    """Consecutive grids in groups of at most ``group_bytes`` of state (0:
    192 MiB), as lists of indices; a grid larger than the budget is a group
    of its own."""
    budget = float(group_bytes) if group_bytes else 192.0 * 2 ** 20
    groups, acc = [], 0.0
    for n, G in enumerate(grids):
        b = _state_bytes(G)
        if not groups or acc + b > budget:
            groups.append([])
            acc = 0.0
        groups[-1].append(n)
        acc += b
    return groups


def solve_bscan(grids, device=0, graph=True, group_bytes=0, tune_steps=3):
    """Solve every model of a B-scan (one grid per trace, any sizes, the same
    iteration count) from the zero state; returns (the receiver buffer of
    each grid, as HipSolver.download_rx gives it; device ms of the runs).

    Consecutive models form groups as the batch forms them (state bytes up to
    ``group_bytes``, 0: 192 MiB): a group of several runs as one HipBatch
    (tuned, shared launches, ``graph``), a group of one as that model's plain
    solve with its own step form.  While a group runs, a second host thread
    builds and uploads the next one, so each group's set-up hides behind the
    device work of the group before it; a group's models are freed once their
    receivers are read."""
    from concurrent.futures import ThreadPoolExecutor
    grids = list(grids)
    if not grids:
        return [], 0.0
    groups = bscan_groups(grids, group_bytes)

    def create(idx):
        made = []
        try:
            for n in idx:
                made.append(HipSolver(grids[n], device))
        except Exception:
            for s in made:
                s.close()
            raise
        return made

    out = [None] * len(grids)
    device_ms = 0.0
    with ThreadPoolExecutor(max_workers=1) as pool:
        nxt = pool.submit(create, groups[0])
        try:
            for gi, idx in enumerate(groups):
                solvers = nxt.result()
                nxt = pool.submit(create, groups[gi + 1]) if gi + 1 < len(groups) else None
                try:
                    if len(solvers) == 1:
                        ms, _, _ = solvers[0].run_timed(0, grids[idx[0]].iterations)
                    else:
                        with HipBatch(solvers) as b:
                            if group_bytes:
                                b.set_group_bytes(group_bytes)
                            b.tune(steps=tune_steps)
                            ms = b.run(0, grids[idx[0]].iterations, graph=graph)
                    device_ms += ms
                    for n, s in zip(idx, solvers):
                        out[n] = s.download_rx()
                finally:
                    for s in solvers:
                        s.close()
        finally:
            if nxt is not None:  # a group built ahead of a failure
                try:
                    for s in nxt.result():
                        s.close()
                except Exception:
                    pass
    return out, device_ms
