"""CPU oracle of the FDTD solve path -- TEST INFRASTRUCTURE.

This package restates the reference's CPU solver (the Cython/OpenMP kernels of
gprMax/fields_updates_ext.pyx and gprMax/pml_updates/*_ext.pyx, driven in the
step order of gprMax/model_build_run.py:408-474).  It is the checker for the
HIP path and the timed CPU baseline.  Only ``tests/``, the smoke check and the
CPU-baseline leg of ``bench.py`` may import it; the product path in
``gprmax_amd`` never does.

Pinning (what checks each function, so that a dropped term, a wrong sign or a
wrong index fails something):

* update_electric / update_magnetic: golden receiver traces of the reference
  (tests/models_basic/hertzian_dipole_fs, magnetic_dipole_fs,
  hertzian_dipole_hs), the closed-form Hertzian dipole field, an independent
  NumPy slicing twin (``oracle.numpy_twin``) bit for bit, and -- where
  /root/reference is present -- the reference's own Cython kernels compiled
  into oracle/_ref (``oracle.build_ref``), bit for bit.
* dispersive A/B: golden traces of hertzian_dipole_dispersive (1 Debye pole);
  the compiled reference kernels bit for bit with random complex coefficients
  (1 and 3 poles, covering Lorentz/Drude); multipole with extra zero-weight
  poles equals one pole bit for bit; real Debye coefficients keep the
  imaginary part of T at exactly zero.
* PML: HORIPML order 1 is exercised by every golden trace (default 10-cell
  PML).  All 24 slab kernels (HORIPML/MRIPML x order 1/2 x 6 directions,
  electric and magnetic) match the compiled reference kernels bit for bit and
  the NumPy twin; reflection from the PML is checked against a large-domain
  run.  No reference trace exists for HORIPML order 2 or MRIPML, so outside
  this container (no oracle/_ref) those rest on the twin and the reflection
  bound.
* snapshots: the compiled reference kernel bit for bit and the NumPy twin of
  the slicing definition (no reference vector).
"""

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}


def build():
    """Compile the oracle libraries in place (gcc, OpenMP)."""
    subprocess.run(['make', '-s', '-C', HERE], check=True)


def load(dtype=np.float32):
    """Return the ctypes library for float32 or float64."""
    key = np.dtype(dtype).name
    if key in _LIBS:
        return _LIBS[key]
    tag = {'float32': 'f32', 'float64': 'f64'}[key]
    path = os.path.join(HERE, 'liboracle_{}.so'.format(tag))
    if not os.path.exists(path):
        build()
    lib = ctypes.CDLL(path)
    real = ctypes.c_float if tag == 'f32' else ctypes.c_double
    P = ctypes.c_void_p
    i = ctypes.c_int
    spec = {
        'update_electric': [i, i, i, i, P, P, P, P, P, P, P, P],
        'update_magnetic': [i, i, i, i, P, P, P, P, P, P, P, P],
        'update_electric_dispersive_A': [i, i, i, i, i, P, P, P, P, P, P, P, P, P, P, P, P],
        'update_electric_dispersive_B': [i, i, i, i, i, P, P, P, P, P, P, P, P],
        'pml_update': [i] * 4 + [i] * 6 + [i] + [i, i, i] + [P, P] + [P] * 6 + [P, P] + [P] * 4 + [real],
        'snapshot_fields': [i] * 12 + [P] * 12,
    }
    funcs = {}
    for name, argtypes in spec.items():
        fn = getattr(lib, 'orc_{}_{}'.format(tag, name))
        fn.argtypes = argtypes
        fn.restype = None
        funcs[name] = fn
    lib.funcs = funcs
    lib.real = real
    _LIBS[key] = lib
    return lib


def ptr(a):
    """Data pointer of a C-contiguous array."""
    assert a.flags['C_CONTIGUOUS'], 'oracle arrays must be C-contiguous'
    return a.ctypes.data_as(ctypes.c_void_p)
