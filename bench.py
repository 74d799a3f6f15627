"""Benchmark of the FDTD solve path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

A step is one iteration of the time loop (fused H pass + fused E pass) over
the whole grid.  W untimed iterations run first, then exactly K iterations are
timed with HIP events on the solver's stream, bracketed by synchronisation.
The default workload is the 300x300x200 heterogeneous half-space (256 soil
materials, PEC cylinder, magnetic dipole, nine receivers).

Prints one JSON line: Mcells/s (NX*NY*NZ*K / T), the roofline of the dominant
kernel (its algorithmic bytes per launch / its mean launch duration from HIP
events, against 8 TB/s HBM; HBM traffic per launch from the committed counter
summary), and the CPU oracle timed on the host cores on a bounded sample of the
same workload.
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


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by what the same
    counters read for a copy of known size), or None when there is none."""
    path = os.path.join(ROOT, 'profiles', 'hbm_traffic.json')
    try:
        with open(path) as f:
            rec = json.load(f)[workload][kernel]
    except (OSError, KeyError, ValueError):
        return None
    return int(rec['fetch_kib'] * 1024 * rec['fetch_scale'] + rec['write_kib'] * 1024 * rec['write_scale'])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=500)
    ap.add_argument('--warmup', type=int, default=50)
    ap.add_argument('--workload', default='heterogeneous_halfspace')
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--xchunk', type=int, default=0)
    ap.add_argument('--block', type=int, default=0)
    ap.add_argument('--tile', type=int, default=0, help='cells per fused block (0 = auto)')
    ap.add_argument('--fused', action='store_true', help='one fused kernel per step instead of the H and E passes')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py: this solver runs on one GPU; --gpus {} is not supported '
              '(one-GPU boxes only; multi-GPU sharding is a design, not a build)'.format(args.gpus),
              file=sys.stderr)
        return 2

    from gprmax_amd.hip import HipSolver, device_info
    from gprmax_amd.workloads import WORKLOADS, algorithmic_bytes_per_step, kernel_bytes

    t_build = time.perf_counter()
    G = WORKLOADS[args.workload]()
    t_build = time.perf_counter() - t_build
    K, W = args.steps, args.warmup
    if W + K > G.iterations:
        raise SystemExit('warmup + steps ({}) exceed the workload iterations ({})'.format(W + K, G.iterations))
    info = device_info(0)
    with HipSolver(G) as s:
        s.set_launch(args.xchunk, args.block)
        s.set_step(1 if args.fused else 0, args.tile)
        fused = s.step_kind()
        s.run(0, W)
        s.synchronize()
        ms, _, _ = s.run_timed(W, K, per_kernel=False)
        s.synchronize()
        # Per-kernel durations (HIP events around every launch on the solver's
        # stream), in a separate run after the timed region.
        n_prof = min(50, G.iterations - W - K)
        ms4 = s.run_profiled(W + K, n_prof) if n_prof > 0 else (0.0,) * 4
        interior = s.interior_cells()
        mem = s.memory_bytes()
    cells = G.nx * G.ny * G.nz
    mcells = cells * K / (ms * 1e-3) / 1e6
    if fused:
        kernels = {'step': (algorithmic_bytes_per_step(G), ms4[0])}
    else:
        bh, be = kernel_bytes(G, 1, interior), kernel_bytes(G, 0, interior)
        kernels = {'h_interior': (bh[0], ms4[0]), 'h_boundary': (bh[1], ms4[1]),
                   'e_interior': (be[0], ms4[2]), 'e_boundary': (be[1], ms4[3])}
    dom = max(kernels, key=lambda k: kernels[k][1])
    dbytes, dms = kernels[dom]
    achieved = dbytes / (dms * 1e-3) / 1e9 if dms > 0 else 0.0
    step_gbs = (sum(b for b, _ in kernels.values())) / (ms / K * 1e-3) / 1e9
    traffic = measured_traffic(args.workload, dom)
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
                   'pml_cells': G.pmls[0].thickness if G.pmls else 0, 'poles': G.maxpoles,
                   'iterations_timed': K, 'device': info['name']},
        'roofline': {'bound': 'hbm', 'kernel': dom, 'achieved': round(achieved, 1),
                     'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': round(achieved / HBM_PEAK_GBS, 4),
                     'traffic': traffic, 'bytes_per_launch': dbytes, 'ms_per_launch': round(dms, 5),
                     'kernels_ms': {k: round(v[1], 5) for k, v in kernels.items()},
                     'kernels_bytes': {k: v[0] for k, v in kernels.items()},
                     'step_achieved': round(step_gbs, 1),
                     'step_form': 'fused, {} cells per block'.format(fused) if fused else 'two-pass',
                     'interior_cells': interior},
        'device_bytes': mem,
        'host_build_s': round(t_build, 1),
    }
    if not args.no_cpu_baseline:
        line['cpu_baseline'] = cpu_baseline(args.workload)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == '__main__':
    sys.exit(main())
