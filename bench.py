"""Benchmark of the FDTD solve path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

A step is one iteration of the time loop (fused H pass + fused E pass) over
the whole grid.  W untimed iterations run first, then exactly K iterations are
timed with HIP events on the solver's stream, bracketed by synchronisation.
The default workload is the 300x300x200 heterogeneous half-space (256 soil
materials, PEC cylinder, magnetic dipole, nine receivers).

Prints one JSON line: Mcells/s (NX*NY*NZ*K / T), the roofline of the dominant
pass (algorithmic bytes / mean pass duration against 8 TB/s HBM), and the CPU
oracle timed on the host cores on a bounded sample of the same workload.
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = 'Mcells/s + % HBM roofline; batched B-scan and adjoint-grad time vs plain (1 GPU)'


def cpu_baseline(workload, seconds=12.0):
    """CPU oracle (C restatement of the reference's OpenMP kernels) on the host
    cores, on the first iterations of the same workload, until ~`seconds` of
    CPU work have run."""
    import numpy as np  # noqa: F401
    from gprmax_amd.workloads import WORKLOADS
    from oracle import load
    from oracle.solver import default_threads, initialise_state, step

    G = WORKLOADS[workload]()
    lib = load(G.floattype)
    threads = default_threads()
    initialise_state(G)
    step(lib, G, 0, threads)  # first touch of every array
    n, t0 = 0, time.perf_counter()
    while True:
        step(lib, G, n + 1, threads)
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds or n + 2 >= G.iterations:
            break
    cells = G.nx * G.ny * G.nz
    return {'value': round(cells * n / el / 1e6, 2), 'unit': 'Mcells/s', 'cores': threads,
            'kind': 'port',
            'sample': '{} iterations (1..{}) of {} on {} OpenMP threads, {:.1f} s'.format(
                n, n, G.title, threads, el)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=500)
    ap.add_argument('--warmup', type=int, default=50)
    ap.add_argument('--workload', default='heterogeneous_halfspace')
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--xchunk', type=int, default=0)
    ap.add_argument('--block', type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py: this solver runs on one GPU; --gpus {} is not supported '
              '(one-GPU boxes only; multi-GPU sharding is a design, not a build)'.format(args.gpus),
              file=sys.stderr)
        return 2

    from gprmax_amd.hip import HipSolver, device_info
    from gprmax_amd.workloads import WORKLOADS, pass_bytes

    t_build = time.perf_counter()
    G = WORKLOADS[args.workload]()
    t_build = time.perf_counter() - t_build
    K, W = args.steps, args.warmup
    if W + K > G.iterations:
        raise SystemExit('warmup + steps ({}) exceed the workload iterations ({})'.format(W + K, G.iterations))
    info = device_info(0)
    with HipSolver(G) as s:
        s.set_launch(args.xchunk, args.block)
        s.run(0, W)
        s.synchronize()
        ms, _, _ = s.run_timed(W, K, per_kernel=False)
        s.synchronize()
        # per-pass durations, separate from the timed region
        n_prof = min(50, G.iterations - W - K)
        _, ms_h, ms_e = s.run_timed(W + K, n_prof, per_kernel=True) if n_prof > 0 else (0, 0, 0)
        mem = s.memory_bytes()
    cells = G.nx * G.ny * G.nz
    mcells = cells * K / (ms * 1e-3) / 1e6
    bh, be = pass_bytes(G, 1), pass_bytes(G, 0)
    dom = ('e_pass', be, ms_e) if ms_e >= ms_h else ('h_pass', bh, ms_h)
    achieved = dom[1] / (dom[2] * 1e-3) / 1e9 if dom[2] > 0 else 0.0
    line = {
        'metric': METRIC,
        'value': round(mcells, 2),
        'unit': 'Mcells/s',
        'n_gpus': 1,
        'steps': K,
        'warmup': W,
        'ms_per_step': round(ms / K, 5),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32',
        'data': 'synthetic',
        'config': {'workload': G.title, 'cells': cells, 'nmat': len(G.materials),
                   'pml_cells': 10, 'iterations_timed': K, 'device': info['name']},
        'roofline': {'bound': 'hbm', 'kernel': dom[0], 'achieved': round(achieved, 1),
                     'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': round(achieved / HBM_PEAK_GBS, 4),
                     'traffic': None, 'bytes_per_launch': dom[1], 'ms_per_launch': round(dom[2], 5),
                     'ms_h_pass': round(ms_h, 5), 'ms_e_pass': round(ms_e, 5)},
        'device_bytes': mem,
        'host_build_s': round(t_build, 1),
    }
    if not args.no_cpu_baseline:
        line['cpu_baseline'] = cpu_baseline(args.workload)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == '__main__':
    sys.exit(main())
