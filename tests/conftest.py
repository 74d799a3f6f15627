import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

REF_TREE = os.path.join(ROOT, 'oracle', '_ref', 'exact')


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an AMD GPU (MI355X) and the built HIP library')
    config.addinivalue_line('markers', 'slow: runs for more than a few seconds on the CPU')


def gpu_present():
    return os.path.exists('/dev/kfd')


def pytest_collection_modifyitems(config, items):
    if gpu_present():
        return
    skip = pytest.mark.skip(reason='no GPU device in this environment')
    for item in items:
        if 'gpu' in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope='session')
def oracle_lib():
    import numpy as np
    import oracle
    return oracle.load(np.float32)


@pytest.fixture(scope='session')
def ref_modules():
    """The reference's own Cython kernels compiled into oracle/_ref (only in a
    container where the reference is mounted)."""
    if not os.path.isdir(REF_TREE):
        pytest.skip('reference kernels not built (oracle/build_ref.py)')
    import importlib
    sys.path.insert(0, REF_TREE)
    try:
        mods = {
            'fields': importlib.import_module('gprMax.fields_updates_ext'),
            'snap': importlib.import_module('gprMax.snapshots_ext'),
        }
        for field in ('electric', 'magnetic'):
            for form in ('HORIPML', 'MRIPML'):
                mods[(field, form)] = importlib.import_module(
                    'gprMax.pml_updates.pml_updates_{}_{}_ext'.format(field, form))
    finally:
        sys.path.remove(REF_TREE)
    return mods


@pytest.fixture(scope='session')
def ref_edge_build():
    """The reference's compiled edge build (gprMax/yee_cell_build_ext.pyx),
    from oracle/_ref.  The module takes its Material class from
    gprMax.materials at import; it only constructs materials and sets their
    er, se, mr, sm and type, so it is handed gprmax_amd's Material class,
    whose attributes are the reference's."""
    if not os.path.isdir(REF_TREE):
        pytest.skip('reference kernels not built (oracle/build_ref.py)')
    import importlib
    import types
    from gprmax_amd.materials import Material
    sys.path.insert(0, REF_TREE)
    shim = types.ModuleType('gprMax.materials')
    shim.Material = Material
    saved = sys.modules.get('gprMax.materials')
    sys.modules['gprMax.materials'] = shim
    try:
        mod = importlib.import_module('gprMax.yee_cell_build_ext')
    finally:
        sys.path.remove(REF_TREE)
        if saved is None:
            del sys.modules['gprMax.materials']
        else:
            sys.modules['gprMax.materials'] = saved
    return mod
