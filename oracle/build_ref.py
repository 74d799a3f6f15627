"""Compile the reference's own Cython kernels of the solve path into oracle/_ref.

TEST INFRASTRUCTURE.  Used only in a container where the reference tree is
mounted, to pin the C restatement in ``oracle/fdtd_oracle.c`` and the host
edge build bit for bit, and to time the reference's CPU kernels beside the C
restatement.  Nothing here is copied: the .pyx sources are compiled where
they lie; the generated C and the built modules go to ``oracle/_ref/``
(git-ignored and never copied to another host: the GPU tests and the
benchmark use only the C restatement).

Two builds are made:
  exact/   -O3 -fopenmp -ffp-contract=off   (every product rounded, no FMA)
  native/  -O3 -w -fopenmp -march=native    (the reference's own flags,
                                             setup.py:176-180)
Cython directives follow setup.py:206-213.

Usage: python oracle/build_ref.py [/root/reference]
"""

import os
import subprocess
import sys
import sysconfig

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, '_ref')
MODULES = [
    'gprMax/fields_updates_ext',
    'gprMax/snapshots_ext',
    'gprMax/pml_updates/pml_updates_electric_HORIPML_ext',
    'gprMax/pml_updates/pml_updates_magnetic_HORIPML_ext',
    'gprMax/pml_updates/pml_updates_electric_MRIPML_ext',
    'gprMax/pml_updates/pml_updates_magnetic_MRIPML_ext',
    # the edge build with dielectric averaging (yee_cell_build_ext cimports
    # the rigid-edge accessors, so both are built)
    'gprMax/yee_cell_setget_rigid_ext',
    'gprMax/yee_cell_build_ext',
]
FLAVOURS = {
    'exact': ['-O3', '-fopenmp', '-ffp-contract=off'],
    'native': ['-O3', '-w', '-fopenmp', '-march=native'],
}
DIRECTIVES = ['boundscheck=False', 'wraparound=False', 'initializedcheck=False',
              'embedsignature=True', 'language_level=3']


def build(reference='/root/reference'):
    csrc = os.path.join(OUT, 'csrc')
    os.makedirs(csrc, exist_ok=True)
    suffix = sysconfig.get_config_var('EXT_SUFFIX')
    include = [np.get_include(), sysconfig.get_paths()['include']]
    for mod in MODULES:
        pyx = os.path.join(reference, mod + '.pyx')
        cfile = os.path.join(csrc, mod.replace('/', '.') + '.c')
        if not os.path.exists(cfile) or os.path.getmtime(cfile) < os.path.getmtime(pyx):
            cmd = [sys.executable, '-m', 'cython', '-3', '-I', reference,
                   '--module-name', mod.replace('/', '.'), '-o', cfile, pyx]
            for d in DIRECTIVES:
                cmd += ['-X', d]
            subprocess.run(cmd, check=True)
        for flavour, flags in FLAVOURS.items():
            target = os.path.join(OUT, flavour, mod + suffix)
            os.makedirs(os.path.dirname(target), exist_ok=True)
            for pkg in (os.path.join(OUT, flavour, 'gprMax'), os.path.join(OUT, flavour, 'gprMax', 'pml_updates')):
                init = os.path.join(pkg, '__init__.py')
                if os.path.isdir(pkg) and not os.path.exists(init):
                    open(init, 'w').close()
            if os.path.exists(target) and os.path.getmtime(target) > os.path.getmtime(cfile):
                continue
            cmd = ['gcc', '-shared', '-fPIC'] + flags + ['-I' + p for p in include] + [cfile, '-o', target]
            subprocess.run(cmd, check=True)
    return OUT


if __name__ == '__main__':
    build(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
    print('built into', OUT)
