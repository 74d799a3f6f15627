"""Build libgprmax_hip.so in place with hipcc for gfx950.

Flags:
* ``-ffp-contract=off`` and ``-fhip-fp32-correctly-rounded-divide-sqrt``: every
  float expression rounds as written, so the kernels reproduce the CPU path's
  float32 arithmetic exactly;
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 math.
"""

import glob
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
# gpx.hip holds the host code and the two-pass kernels; each step_*.hip holds
# the fused kernels of one PML form.  They compile in parallel.
SOURCES = sorted(glob.glob(os.path.join(HERE, 'csrc', '*.hip')))
LIBRARY = os.path.join(HERE, 'libgprmax_hip.so')
ARCH = os.environ.get('GPX_OFFLOAD_ARCH', 'gfx950')

HIPCC_FLAGS = [
    '-O3', '-std=c++17', '-fPIC',
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


def build(verbose=False, jobs=None):
    """Compile the HIP library; returns its path.  Always recompiles."""
    jobs = jobs or min(len(SOURCES), int(os.environ.get('MAX_JOBS', os.cpu_count() or 4)))
    inc = ['-I', os.path.join(ROOT, 'include')]
    with tempfile.TemporaryDirectory(prefix='gpx_build_') as tmpdir:
        objs, pending = [], []
        for src in SOURCES:
            obj = os.path.join(tmpdir, os.path.basename(src) + '.o')
            objs.append(obj)
            pending.append([hipcc()] + HIPCC_FLAGS + inc + ['-c', src, '-o', obj])
        running = []
        while pending or running:
            while pending and len(running) < jobs:
                cmd = pending.pop(0)
                if verbose:
                    print(' '.join(cmd), flush=True)
                running.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                                      text=True)))
            cmd, proc = running.pop(0)
            out, _ = proc.communicate()
            if proc.returncode != 0:
                for _, p in running:
                    p.kill()
                sys.stderr.write(out)
                raise RuntimeError('hipcc failed for ' + cmd[cmd.index('-c') + 1])
        tmp = LIBRARY + '.tmp'
        cmd = [hipcc(), '-shared', '-fPIC', '--offload-arch=' + ARCH] + objs + ['-o', tmp]
        if verbose:
            print(' '.join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError('linking libgprmax_hip.so failed')
    os.replace(tmp, LIBRARY)
    return LIBRARY


if __name__ == '__main__':
    print(build(verbose=True))
