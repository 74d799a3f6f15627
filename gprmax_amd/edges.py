"""Edge material indices built on the GPU (gpx_edges_*): the edge loops of
build_electric_components / build_magnetic_components
(gprMax/yee_cell_build_ext.pyx:110-257).

The device writes every edge whose cells agree and collects the distinct sets
of cell materials of the others; this module finds or creates each set's
averaged material once, in the order of the set's first edge, with the same
search as the host build (``grid._average_material``), and the device writes
the results.  The outcome (edge indices, materials, their order) is that of
``grid.build_electric_components`` / ``build_magnetic_components``.
"""

import ctypes

import numpy as np

from .grid import _average_material
from .hip import library
from .utilities import GeneralError

_P, _I = ctypes.c_void_p, ctypes.c_int
_U64P = ctypes.POINTER(ctypes.c_ulonglong)
SIGNATURES = {
    'gpx_edges_create': ([_I, _I, _I, _I, _P, _P, _P, _P, _I, ctypes.POINTER(_P)], _I),
    'gpx_edges_destroy': ([_P], _I),
    'gpx_edges_last_error': ([_P], ctypes.c_char_p),
    'gpx_edges_classify': ([_P, _I, _P, _P, _I, ctypes.POINTER(_I)], _I),
    'gpx_edges_resolve': ([_P, _I, _P, _I], _I),
    'gpx_edges_download': ([_P, _P], _I),
    'gpx_edges_static': ([_I, _I, _I, _I, _P, _P, _P, _I, _I, _P, _P, _P, _P, _I, ctypes.POINTER(_I)], _I),
}

# Sets per call of gpx_edges_static (given or returned).
STATIC_MAX_SETS = 1 << 15

# The cells around an edge in the reference's numID1.. order, per component.
_OFFSETS = [((0, 0, 0), (0, -1, 0), (0, -1, -1), (0, 0, -1)),
            ((0, 0, 0), (-1, 0, 0), (-1, 0, -1), (0, 0, -1)),
            ((0, 0, 0), (-1, 0, 0), (-1, -1, 0), (0, -1, 0)),
            ((0, 0, 0), (-1, 0, 0)),
            ((0, 0, 0), (0, -1, 0)),
            ((0, 0, 0), (0, 0, -1))]


def _lib():
    lib = library()
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _resolve_sets(G, solid, field, first, ntot):
    """The material of each mixed set, found or created in the order of the
    sets' first edges (the reference's loop order), as the host build does."""
    first = np.asarray(first).astype(np.int64)
    numids = np.zeros(len(first), dtype=np.uint32)
    cache = {}
    for s in np.argsort(first, kind='stable'):
        comp, point = divmod(int(first[s]), ntot)
        i, rest = divmod(point, (G.ny + 1) * (G.nz + 1))
        j, k = divmod(rest, G.nz + 1)
        row = [int(solid[i + a, j + b, k + c]) for a, b, c in _OFFSETS[comp]]
        names = [G.materials[m].ID for m in row]
        if any('+' in nm for nm in names):
            raise GeneralError('the device edge build needs material names without "+"')
        numids[s] = _average_material(G, names, row, cache, field == 0)
    return numids


def build_components_on_device(G, device=0, capacity=1 << 16):
    """Build G.ID's edge indices on the GPU; creates averaged materials in
    G.materials as the host build does."""
    if len(G.materials) >= 0xFFFF:
        raise GeneralError('the device edge build needs fewer than 65535 cell materials')
    lib = _lib()
    solid = np.ascontiguousarray(G.solid, dtype=np.uint32)
    rigid = [np.ascontiguousarray(r, dtype=np.uint8) for r in (G.rigidcellE, G.rigidcellH)]
    ID = np.ascontiguousarray(G.ID, dtype=np.uint32)
    h = ctypes.c_void_p()
    rc = lib.gpx_edges_create(device, G.nx, G.ny, G.nz, _ptr(solid), _ptr(rigid[0]), _ptr(rigid[1]), _ptr(ID),
                              int(capacity), ctypes.byref(h))
    if rc:
        raise GeneralError('gpx_edges_create failed ({}): {}'.format(rc, lib.gpx_edges_last_error(None).decode()))
    ntot = (G.nx + 1) * (G.ny + 1) * (G.nz + 1)
    try:
        for field in (0, 1):
            count = ctypes.c_int()
            keys = np.zeros(capacity, dtype=np.uint64)
            first = np.zeros(capacity, dtype=np.uint64)
            rc = lib.gpx_edges_classify(h, field, _ptr(keys), _ptr(first), capacity, ctypes.byref(count))
            if rc:
                raise GeneralError('gpx_edges_classify failed ({}): {}'.format(rc, lib.gpx_edges_last_error(h).decode()))
            n = count.value
            numids = _resolve_sets(G, solid, field, first[:n], ntot)
            rc = lib.gpx_edges_resolve(h, field, _ptr(numids), n)
            if rc:
                raise GeneralError('gpx_edges_resolve failed ({}): {}'.format(rc, lib.gpx_edges_last_error(h).decode()))
        rc = lib.gpx_edges_download(h, _ptr(ID))
        if rc:
            raise GeneralError('gpx_edges_download failed ({}): {}'.format(rc, lib.gpx_edges_last_error(h).decode()))
    finally:
        lib.gpx_edges_destroy(h)
    G.ID[...] = ID


def build_components_static(G, solid_dev, rigidE_dev, rigidH_dev, ID_dev, device=0):
    """The same build through gpx_edges_static, which allocates nothing: the
    caller's device arrays (addresses as integers; solid uint32 [nx][ny][nz],
    rigid uint8 [nx][ny][nz], ID uint32 [6][nx+1][ny+1][nz+1] holding the
    geometry's indices) are read and ID is written in place on `device`.
    Averaged materials are created in G.materials as the host build does; the
    host copy of the cell materials, G.solid, names each set's cells."""
    if len(G.materials) >= 0xFFFF:
        raise GeneralError('the device edge build needs fewer than 65535 cell materials')
    lib = _lib()
    solid = np.ascontiguousarray(G.solid, dtype=np.uint32)
    ntot = (G.nx + 1) * (G.ny + 1) * (G.nz + 1)
    keys = np.zeros(STATIC_MAX_SETS, dtype=np.uint64)
    first = np.zeros(STATIC_MAX_SETS, dtype=np.uint64)
    count = ctypes.c_int()

    def call(field, rigid, known_keys, known_ids):
        rc = lib.gpx_edges_static(device, G.nx, G.ny, G.nz, solid_dev, rigid, ID_dev, field, len(known_keys),
                                  _ptr(known_keys), _ptr(known_ids), _ptr(keys), _ptr(first), len(keys),
                                  ctypes.byref(count))
        if rc:
            raise GeneralError('gpx_edges_static failed ({}): {}'.format(rc, lib.gpx_edges_last_error(None).decode()))
        return count.value

    for field, rigid in ((0, rigidE_dev), (1, rigidH_dev)):
        none_k, none_i = np.zeros(1, dtype=np.uint64), np.zeros(1, dtype=np.uint32)
        n = call(field, rigid, none_k[:0], none_i[:0])
        if n:
            set_keys = keys[:n].copy()
            numids = _resolve_sets(G, solid, field, first[:n], ntot)
            left = call(field, rigid, set_keys, numids)
            if left:
                raise GeneralError('gpx_edges_static: {} sets left unresolved'.format(left))
