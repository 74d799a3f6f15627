"""Build libgprmax_hip.so in place with hipcc for gfx950.

Flags:
* ``-ffp-contract=off`` and ``-fhip-fp32-correctly-rounded-divide-sqrt``: every
  float expression rounds as written, so the kernels reproduce the CPU path's
  float32 arithmetic exactly;
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 math.
"""

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = [os.path.join(HERE, 'csrc', 'gpx.hip')]
LIBRARY = os.path.join(HERE, 'libgprmax_hip.so')
ARCH = os.environ.get('GPX_OFFLOAD_ARCH', 'gfx950')

HIPCC_FLAGS = [
    '-O3', '-std=c++17', '-fPIC', '-shared',
    '--offload-arch=' + ARCH,
    '-ffp-contract=off',
    '-fhip-fp32-correctly-rounded-divide-sqrt',
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wall', '-Wno-unused-function',
]


def hipcc():
    path = shutil.which('hipcc') or '/opt/rocm/bin/hipcc'
    if not os.path.exists(path):
        raise RuntimeError('hipcc not found')
    return path


def build(verbose=False):
    """Compile the HIP library; returns its path.  Always recompiles."""
    tmp = LIBRARY + '.tmp'
    cmd = [hipcc()] + HIPCC_FLAGS + ['-I', os.path.join(ROOT, 'include')] + SOURCES + ['-o', tmp]
    if verbose:
        print(' '.join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError('hipcc failed for libgprmax_hip.so')
    os.replace(tmp, LIBRARY)
    return LIBRARY


if __name__ == '__main__':
    print(build(verbose=True))
