"""The edge build that allocates nothing (gpx_edges_static) against the edge
build that allocates its tables (gpx_edges_create / classify / resolve /
download), both through the C ABI: the same edge indices bit for bit and the
same averaged materials in the same order.  Two calls in a row on one GPU,
the second on a smaller grid, so the second finds the first's contents in
the library's fixed tables."""

import ctypes

import numpy as np
import pytest

from models import edge_model

pytestmark = pytest.mark.gpu


class _DeviceArrays:
    """Device copies of the grid's cell materials, rigid flags and edge
    indices, made with the HIP runtime the library itself links against
    (libamdhip64.so.7, loaded with the library), so that the test holds one
    HIP runtime per process."""

    H2D, D2H = 1, 2

    def __init__(self, G):
        from gprmax_amd.hip import library
        library()
        self.rt = ctypes.CDLL('libamdhip64.so.7')
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.rt.hipSetDevice(0)
        self.ptrs = []
        self.solid = self._up(np.ascontiguousarray(G.solid, dtype=np.uint32))
        self.rigidE = self._up(np.ascontiguousarray(G.rigidcellE, dtype=np.uint8))
        self.rigidH = self._up(np.ascontiguousarray(G.rigidcellH, dtype=np.uint8))
        self.ID = self._up(np.ascontiguousarray(G.ID, dtype=np.uint32))

    def _up(self, a):
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        self.ptrs.append(p)
        assert self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, self.H2D) == 0
        return p.value

    def download_ID(self, out):
        assert self.rt.hipMemcpy(out.ctypes.data, self.ID, out.nbytes, self.D2H) == 0

    def free(self):
        for p in self.ptrs:
            self.rt.hipFree(p)
        self.ptrs = []


def _static_build(G):
    from gprmax_amd.edges import build_components_static
    d = _DeviceArrays(G)
    try:
        build_components_static(G, d.solid, d.rigidE, d.rigidH, d.ID, device=0)
        ID = np.empty(G.ID.shape, dtype=np.uint32)
        d.download_ID(ID)
    finally:
        d.free()
    G.ID[...] = ID


@pytest.mark.parametrize('shapes', [((34, 29, 23), (17, 12, 9)), ((40, 33, 1), (12, 9, 1))],
                         ids=['3d', '2d'])
def test_static_edge_build_matches_allocating_build_twice(shapes):
    from gprmax_amd.edges import build_components_on_device
    averaged = []
    for n, shape in enumerate(shapes):
        seed = 40 + n
        ref, ours = edge_model(seed, shape), edge_model(seed, shape)
        base = len(ref.materials)
        build_components_on_device(ref)
        _static_build(ours)
        assert np.array_equal(ref.ID, ours.ID), (shape, int(np.sum(ref.ID != ours.ID)))
        assert [m.ID for m in ref.materials] == [m.ID for m in ours.materials], shape
        for a, b in zip(ref.materials, ours.materials):
            assert (a.er, a.se, a.mr, a.sm) == (b.er, b.se, b.mr, b.sm)
        averaged.append(len(ref.materials) - base)
    # the first (larger) grid of each pair needs averaged materials
    assert averaged[0] > 0, averaged

