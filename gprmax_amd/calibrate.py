"""Calibration run for HBM byte counters.

Runs a plain device-to-device copy of a known size (a torch tensor copy, one
elementwise copy kernel) a few times.  Profiling this script with
``rocprofv3 --pmc FETCH_SIZE`` and ``--pmc WRITE_SIZE`` gives what those
counters report for exactly ``bytes`` read and ``bytes`` written, the scale
used to turn the solver kernels' counter values into HBM bytes.

    python -m gprmax_amd.calibrate [--mib 1024] [--reps 3]
"""

import argparse
import json


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--mib', type=int, default=1024)
    ap.add_argument('--reps', type=int, default=3)
    args = ap.parse_args()
    import torch
    n = args.mib * (1 << 20) // 4
    src = torch.empty(n, dtype=torch.float32, device='cuda').uniform_()
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(args.reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({'copy_bytes_read': n * 4, 'copy_bytes_written': n * 4, 'reps': args.reps}))


if __name__ == '__main__':
    main()
