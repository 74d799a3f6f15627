"""The C-ABI library loads and exports every symbol include/gpx.h declares.
No compute call is made (this runs without a GPU)."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, 'include', 'gpx.h')).read()
    return sorted(set(re.findall(r'\b(gpx_[A-Za-z_]+)\s*\(', text)))


def test_header_and_binding_agree():
    from gprmax_amd.hip import SIGNATURES
    from gprmax_amd.inversion import SIGNATURES as INV
    from gprmax_amd.edges import SIGNATURES as EDGES
    assert declared_functions() == sorted(set(SIGNATURES) | set(INV) | set(EDGES))


def test_library_exports_every_declared_symbol():
    from gprmax_amd.hip import LIBRARY, library
    if not os.path.exists(LIBRARY):
        pytest.fail('libgprmax_hip.so is not built (python -m gprmax_amd.build)')
    lib = library()
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_create_rejects_bad_descriptors_before_touching_a_device():
    from gprmax_amd.hip import GpxDesc, library
    lib = library()
    ctx = ctypes.c_void_p()
    bad = [GpxDesc(1, 1, 10, 5, 2, 0, 0, 0),      # one cell on two axes
           GpxDesc(0, 10, 10, 5, 2, 0, 0, 0),     # no cells
           GpxDesc(10, 10, 10, 0, 2, 0, 0, 0),    # no iterations
           GpxDesc(10, 10, 10, 5, 2, 9, 0, 0),    # too many poles
           GpxDesc(10, 10, 10, 5, 2, 0, 7, 0),    # unknown PML formulation
           GpxDesc(10, 10, 10, 5, 5000, 0, 0, 0)]  # table larger than LDS
    for d in bad:
        rc = lib.gpx_create(ctypes.byref(d), ctypes.byref(ctx))
        assert rc < 0, (d.nx, d.iterations, d.maxpoles)
        assert lib.gpx_last_error(None)
    assert lib.gpx_create(None, ctypes.byref(ctx)) < 0


def test_create_refuses_grids_past_the_32_bit_offset_limit():
    """A field component of 2^30 points or more would need byte offsets past
    4 GiB; gpx_create refuses it (GPX_ELIMIT) before touching a device, rather
    than letting the kernels' 32-bit offsets wrap.  The reference accepts any
    grid its memory check passes (gprMax/grid.py:227-245), so the refusal
    names the limit."""
    from gprmax_amd.hip import GpxDesc, library
    lib = library()
    ctx = ctypes.c_void_p()
    for shape in [(1024, 1024, 1024), (1023, 1023, 1024), (4095, 4095, 63), (2, 2**16, 2**14)]:
        pts = (shape[0] + 1) * (shape[1] + 1) * (shape[2] + 1)
        assert pts * 4 > 0xFFFFFFFF
        d = GpxDesc(*shape, 5, 2, 0, 0, 0)
        assert lib.gpx_create(ctypes.byref(d), ctypes.byref(ctx)) == -3, shape
        msg = lib.gpx_last_error(None).decode()
        assert 'points per field component' in msg and str(pts) in msg, msg
        assert not ctx.value


def test_oracle_and_product_are_separate():
    """The product package never imports the oracle."""
    pkg = os.path.join(ROOT, 'gprmax_amd')
    for name in os.listdir(pkg):
        if name.endswith('.py'):
            src = open(os.path.join(pkg, name)).read()
            assert 'import oracle' not in src and 'from oracle' not in src, name


def test_static_edge_build_refuses_too_many_sets_before_touching_a_device():
    """gpx_edges_static holds its tables in a fixed device array: more
    resolved sets than it holds are refused with a clear error (GPX_ELIMIT)
    before any device work."""
    import numpy as np
    from gprmax_amd.edges import STATIC_MAX_SETS, _lib
    lib = _lib()
    n = STATIC_MAX_SETS + 1
    keys = np.arange(n, dtype=np.uint64)
    ids = np.zeros(n, dtype=np.uint32)
    count = ctypes.c_int(-1)
    rc = lib.gpx_edges_static(0, 4, 4, 4, 1, 1, 1, 0, n, keys.ctypes.data, ids.ctypes.data, None, None, 0,
                              ctypes.byref(count))
    assert rc == -3
    assert 'at most 32768' in lib.gpx_edges_last_error(None).decode()
    assert count.value == -1


def test_bscan_groups_follow_the_state_budget():
    """solve_bscan's groups: consecutive grids while their state (36 B per
    grid point) fits the budget, a grid past it alone, order kept."""
    from types import SimpleNamespace
    from gprmax_amd.hip import bscan_groups
    small = SimpleNamespace(nx=99, ny=99, nz=99)      # 1e6 points, 36e6 B
    large = SimpleNamespace(nx=255, ny=255, nz=255)   # 16.8e6 points, 604e6 B
    grids = [small, small, small, large, small, small, large, large]
    assert bscan_groups(grids) == [[0, 1, 2], [3], [4, 5], [6], [7]]
    assert bscan_groups(grids, group_bytes=80e6) == [[0, 1], [2], [3], [4, 5], [6], [7]]
    assert bscan_groups(grids, group_bytes=1) == [[n] for n in range(8)]
    assert bscan_groups([]) == []
