"""Benchmark of the FDTD solve path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

A step is one iteration of the time loop (the H and E updates with their PML
corrections, sources and receivers) over the whole grid, in the step form the
library picks.  W untimed iterations run first, then exactly K iterations are
timed with HIP events on the solver's stream, bracketed by synchronisation.
The default workload is the 300x300x200 heterogeneous half-space (256 soil
materials, PEC cylinder, magnetic dipole, nine receivers).

Prints one JSON line: Mcells/s (NX*NY*NZ*K / T), the roofline of the dominant
kernel (its algorithmic bytes per launch / its mean launch duration from HIP
events, against 8 TB/s HBM; HBM traffic per launch from the committed counter
summary), and the CPU oracle timed on the host cores on a bounded sample of the
same workload.  On the default workload the line also carries the metric's
other two parts under `forms`: the batched B-scan against sequential plain
solves (16 of the 96 traces) and the adjoint gradient against one forward
(`--forms 0` leaves them out; `--workload bscan` / `adjoint` runs them alone).
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


def host_cpu():
    """Model name and physical core count of the host, from lscpu."""
    import subprocess
    try:
        out = subprocess.run(['lscpu'], capture_output=True, text=True, timeout=20).stdout
    except (OSError, subprocess.SubprocessError):
        return 'unknown CPU', None
    info = {}
    for ln in out.splitlines():
        if ':' in ln:
            k, v = ln.split(':', 1)
            info[k.strip()] = v.strip()
    try:
        phys = int(info['Core(s) per socket']) * int(info['Socket(s)'])
    except (KeyError, ValueError):
        phys = None
    return info.get('Model name', 'unknown CPU'), phys


def cpu_baseline(workload, seconds=12.0):
    """The reference's CPU path, as the oracle's C restatement of its Cython
    kernels (same loops, same OpenMP static schedule over i, float32, no FMA
    contraction; kind "port"), on the first iterations of the same workload
    until ~`seconds` of CPU work have run.  Threads: OMP_NUM_THREADS (the
    host's CPU share for one GPU), reported as `cores`; lscpu's model and
    physical core count go into the sample text."""
    from gprmax_amd.workloads import WORKLOADS
    from oracle import load
    from oracle.solver import default_threads, initialise_state, step

    G = WORKLOADS[workload]()
    threads = default_threads()
    lib = load(G.floattype)
    model, phys = host_cpu()
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
            'sample': '{} iterations (1..{}) of {}, the C restatement of the reference kernels '
                      '(-O3 -fopenmp -ffp-contract=off), {} OpenMP threads on {} ({} physical cores), '
                      '{:.1f} s'.format(n, n, G.title, threads, model, phys if phys else 'unknown', el)}


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


def bench_bscan(args):
    line = bscan_line(args, args.traces)
    print(json.dumps(line), flush=True)
    return 0 if line['receivers_identical'] else 1


def bscan_line(args, traces):
    """Uneven B-scan: `--traces` models (sides 128-256 cells, 1 mm, 10-cell
    PML, random soil), each `--bscan-its` iterations, built on the host first
    (outside both timings).  Both legs are timed the same way, from the first
    upload to the last download:
    * sequential plain solves, as the reference's trace loop re-runs the
      solve: per trace create + upload, run, download, destroy;
    * the B-scan solve (`solve_bscan`, the value): the traces in groups as
      the batch forms them, each group's set-up (build and upload) on a
      second host thread while the group before it runs, small models
      advanced together by shared launches, large ones as plain solves;
    * the resident batch (reported beside it): every trace created and
      uploaded, the autotuner, the whole run in one HipBatch, every download.
    Also reported: the device time of the runs alone (HIP events), which
    shows how much of the gain is set-up and launch overhead rather than
    kernel time.  The batched receivers must equal the sequential ones bit
    for bit (the GPU tests compare batched traces with the oracle)."""
    import numpy as np
    from gprmax_amd.hip import HipBatch, HipSolver, solve_bscan
    from gprmax_amd.workloads import bscan_trace

    models = []
    for t in range(traces):
        models.append(bscan_trace(t, iterations=args.bscan_its))
        print('built trace {} {}'.format(t, models[-1].title), file=sys.stderr, flush=True)
    cells = sum(G.nx * G.ny * G.nz for G in models)
    plain_dev = 0.0
    plain_rx = []
    t0 = time.perf_counter()
    for t, G in enumerate(models):
        with HipSolver(G) as s:
            if args.fused:
                s.set_step(1)
            elif args.passes:
                s.set_step(0)
            ms, _, _ = s.run_timed(0, G.iterations)
            plain_rx.append(s.download_rx())
        plain_dev += ms * 1e-3
        print('trace {} plain {:.3f} s on the device'.format(t, ms * 1e-3), file=sys.stderr, flush=True)
    plain_wall = time.perf_counter() - t0
    t0 = time.perf_counter()
    rx_p, ms_p = solve_bscan(models, graph=bool(args.bscan_graph))
    pipe_wall = time.perf_counter() - t0
    print('B-scan solve {:.3f} s'.format(pipe_wall), file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    solvers = [HipSolver(G) for G in models]
    t_up = time.perf_counter() - t0
    with HipBatch(solvers) as b:
        tile, xchunk, ms_tune = b.tune(steps=3)
        t_tune = time.perf_counter() - t0 - t_up
        ms = b.run(0, args.bscan_its, graph=bool(args.bscan_graph))
        rx = [s.download_rx() for s in solvers]
    for s in solvers:
        s.close()
    batch_wall = time.perf_counter() - t0
    same = (all(np.array_equal(a, b_) for a, b_ in zip(rx, plain_rx)) and
            all(np.array_equal(a, b_) for a, b_ in zip(rx_p, plain_rx)))
    line = {
        'metric': 'batched B-scan time vs sequential plain solves (1 GPU)',
        'value': round(plain_wall / pipe_wall, 3),
        'unit': 'x (sequential wall / B-scan solve wall, uploads to downloads on both sides)',
        'n_gpus': 1, 'steps': args.bscan_its, 'warmup': 0,
        'higher_is_better': True, 'scaling': 'strong', 'vs_baseline': None,
        'dtype': 'f32', 'data': 'synthetic',
        'config': {'workload': 'uneven B-scan, {} traces, sides 128-256 cells'.format(traces),
                   'cells': cells, 'iterations': args.bscan_its},
        'sequential_wall_s': round(plain_wall, 3), 'sequential_device_s': round(plain_dev, 3),
        'bscan_solve_wall_s': round(pipe_wall, 3), 'bscan_solve_device_s': round(ms_p * 1e-3, 3),
        'resident_batch_ratio': round(plain_wall / batch_wall, 3),
        'batched_wall_s': round(batch_wall, 3), 'batched_upload_s': round(t_up, 3),
        'batched_tune_s': round(t_tune, 3), 'batched_device_s': round(ms * 1e-3, 3),
        'device_ratio': round(plain_dev / (ms * 1e-3), 3),
        'batched_mcells_per_s': round(cells * args.bscan_its / (ms * 1e-3) / 1e6, 1),
        'tuned_shape': {'tile_cells': tile, 'xchunk': xchunk, 'ms_per_iteration': round(ms_tune, 4)},
        'plain_step_form': 'fused' if args.fused else ('two-pass' if args.passes else 'default (timed choice)'),
        'receivers_identical': same,
    }
    return line


def bench_edges(args):
    """Edge material indices of a 300x300x200 grid whose lower half is an
    averaged random mixture of 6 soils (every soil edge needs an averaged
    material): the GPU build against the host build, same result required."""
    import numpy as np
    from gprmax_amd.edges import build_components_on_device
    from gprmax_amd.grid import ModelBuilder, build_electric_components, build_magnetic_components

    def model():
        rng = np.random.default_rng(1006)
        nx, ny, nz = 300, 300, 200
        b = ModelBuilder(nx, ny, nz, 2e-3, 2e-3, 2e-3, 10, pml_cells=0)
        for n in range(6):
            b.add_material(2 + 0.5 * n, 1e-3 * n, 1, 0, 'soil{}'.format(n))
        b.G.solid[:, :, :nz // 2] = rng.integers(2, 8, (nx, ny, nz // 2))
        return b.G
    G = model()
    t0 = time.perf_counter()
    build_components_on_device(G)
    t_dev = time.perf_counter() - t0
    G2 = model()
    t0 = time.perf_counter()
    build_electric_components(G2)
    build_magnetic_components(G2)
    t_host = time.perf_counter() - t0
    same = bool(np.array_equal(G.ID, G2.ID)) and [m.ID for m in G.materials] == [m.ID for m in G2.materials]
    line = {'metric': 'edge material build, host time / GPU time', 'value': round(t_host / t_dev, 2),
            'unit': 'x', 'n_gpus': 1, 'steps': 1, 'warmup': 0, 'higher_is_better': True, 'scaling': 'strong',
            'vs_baseline': None, 'dtype': 'u32', 'data': 'synthetic',
            'config': {'workload': 'averaged random 6-soil half-space 300x300x200', 'materials': len(G.materials)},
            'gpu_s': round(t_dev, 3), 'host_s': round(t_host, 3), 'identical': same}
    print(json.dumps(line), flush=True)
    return 0 if same else 1


def bench_adjoint(args):
    line = adjoint_line(args)
    print(json.dumps(line), flush=True)
    return 0


def adjoint_line(args):
    """Config of the differentiable solve: 160x160x120 cells at 2 mm, 10-cell
    PML, smooth true model (er in [4, 10], sigma in [1e-3, 2e-2]),
    homogeneous start model, 8 z-directed Hertzian sources x 16 Ez receivers,
    L2 trace misfit, checkpoint interval K = 64 over 2000 iterations.
    Reports the cost of the gradient (forward + recompute + adjoint) against
    one forward; then (--adj-keep 1) the same gradient with the forward
    keeping every segment's stored states (about 154 GB of device memory),
    which leaves forward + adjoint; and, with --fd-cells, central differences
    of the float32 forward at that many random interior cells (first source
    only)."""
    import numpy as np
    from gprmax_amd.inversion import InversionGrid, misfit_and_gradient

    shape, d, its = (160, 160, 120), 2e-3, args.adj_its
    rng = np.random.default_rng(1005)

    def smooth(lo, hi):
        f = rng.standard_normal(shape)
        for ax in range(3):               # separable box blur, 9 cells wide, 3 passes
            for _ in range(3):
                f = sum(np.roll(f, s, axis=ax) for s in range(-4, 5)) / 9.0
        f = (f - f.min()) / (f.max() - f.min())
        return lo + (hi - lo) * f

    er_true, se_true = smooth(4.0, 10.0), smooth(1e-3, 2e-2)
    er0, se0 = np.full(shape, 7.0), np.full(shape, 1.05e-2)
    grid = InversionGrid(shape, d, its, pml=10, er_pml=7.0, K=64)
    z = 120 - 10 - 5                      # 5 cells inside the top PML
    sources = [(40 + 26 * (n % 4), 55 + 50 * (n // 4), z) for n in range(8)]
    receivers = [(30 + 33 * (n % 4), 30 + 33 * (n // 4), z) for n in range(16)]
    waves = grid.waveform(freq=900e6)
    grid.set_model(er_true, se_true)
    observed = [grid.forward(s, receivers, waves) for s in sources]
    t0 = time.perf_counter()
    loss, ger, gse = misfit_and_gradient(grid, er0, se0, sources, receivers, observed, waves)
    wall = time.perf_counter() - t0
    ms_f, ms_b = grid.timings()          # of the last source
    line = {
        'metric': 'adjoint-gradient time vs one forward solve (1 GPU)',
        'value': round((ms_f + ms_b) / ms_f, 3),
        'unit': 'x (forward + recompute + adjoint) / forward',
        'n_gpus': 1, 'steps': its, 'warmup': 0, 'higher_is_better': False, 'scaling': 'strong',
        'vs_baseline': None, 'dtype': 'f32', 'data': 'synthetic',
        'config': {'workload': 'adjoint 160x160x120, 8 sources x 16 receivers, K=64', 'iterations': its},
        'forward_ms': round(ms_f, 2), 'backward_ms': round(ms_b, 2),
        'forward_mcells_per_s': round(float(np.prod(shape)) * its / (ms_f * 1e-3) / 1e6, 1),
        'gradient_wall_s_all_sources': round(wall, 2), 'loss': loss,
        'grad_norm_er': float(np.linalg.norm(ger)), 'grad_norm_se': float(np.linalg.norm(gse)),
    }
    if args.adj_keep:
        # The same gradient with the forward keeping every segment's stored
        # states in device memory, so the backward recomputes nothing.  A
        # device without that much free memory reports the failure here.
        nbytes = grid.keep_bytes_all()
        try:
            grid.set_keep(nbytes)
            t0 = time.perf_counter()
            loss_k, ger_k, gse_k = misfit_and_gradient(grid, er0, se0, sources, receivers, observed, waves)
            wall_k = time.perf_counter() - t0
            ms_fk, ms_bk = grid.timings()
            line['kept_states'] = {
                'device_gb': round(nbytes / 1e9, 1),
                'forward_ms': round(ms_fk, 2), 'backward_ms': round(ms_bk, 2),
                'ratio_to_plain_forward': round((ms_fk + ms_bk) / ms_f, 3),
                'gradient_wall_s_all_sources': round(wall_k, 2),
                'gradient_identical': bool(loss_k == loss and np.array_equal(ger_k, ger) and
                                           np.array_equal(gse_k, gse)),
            }
        except Exception as e:  # noqa: BLE001 - reported, not hidden
            line['kept_states'] = {'device_gb': round(nbytes / 1e9, 1), 'error': '{}: {}'.format(type(e).__name__, e)}
        grid.set_keep(0)
    if args.fd_cells:
        idx = np.argwhere(grid.interior)
        picks = idx[rng.choice(len(idx), args.fd_cells, replace=False)]
        src, obs = sources[:1], observed[:1]
        _, g1, _ = misfit_and_gradient(grid, er0, se0, src, receivers, obs, waves)
        errs = []
        for c in picks:
            c = tuple(c)
            h = 1e-2 * er0[c]
            Ls = []
            for sgn in (1, -1):
                e = er0.copy()
                e[c] += sgn * h
                Ls.append(misfit_and_gradient(grid, e, se0, src, receivers, obs, waves)[0])
            fd = (Ls[0] - Ls[1]) / (2 * h)
            errs.append(abs(fd - g1[c]) / max(abs(fd), 1e-30))
        line['fd_cells'] = args.fd_cells
        line['fd_rel_err_median'] = float(np.median(errs))
        line['fd_rel_err_max'] = float(np.max(errs))
    grid.close()
    return line


def step_form_label(fused, form, choice, default_form):
    """The step form that ran, and how it was chosen."""
    name = ('fused, {} cells per block'.format(fused) if fused else
            {3: 'split', 5: 'split, H boundary beside the interior kernel',
             2: 'two-pass side by side', 6: 'two-pass merged, one launch per pass',
             7: 'split, H boundary in the interior launch'}.get(form, 'two-pass'))
    if not default_form:
        return name + ' (set on the command line)'
    if choice == -1:
        return name + ' (default form, not timed: the run was too short to time the forms)'
    return name + ' (default form: timed against the others and kept)'


def calibrate(args):
    """A device copy of known size, for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
    runs: what the counters read for exactly `bytes` read and written gives
    the scale that turns the solver kernels' counts into HBM bytes."""
    import torch
    n = args.calibrate_mib * (1 << 20) // 4
    src = torch.empty(n, dtype=torch.float32, device='cuda').uniform_()
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({'copy_bytes_read': n * 4, 'copy_bytes_written': n * 4, 'reps': 3}), flush=True)
    return 0


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
    ap.add_argument('--overlap', action='store_true', help='H and E passes with interior and boundary kernels side by side')
    ap.add_argument('--passes', action='store_true', help='the H and E passes (step form 0)')
    ap.add_argument('--merged', action='store_true',
                    help='the passes as one launch each over every cell (form 6)')
    ap.add_argument('--split-merged', action='store_true',
                    help='the split step with the H boundary pass inside the interior launch (form 7)')
    ap.add_argument('--split', action='store_true',
                    help='the split step (form 3): one kernel doing H and E for the interior, boundary passes '
                         'around it; by default the library times forms 7, 6, 2, 5, the split step and the passes '
                         '(forms 6, 2 and the passes where the split step does not apply) and keeps the fastest')
    ap.add_argument('--split-beside', action='store_true',
                    help='the split step with the H boundary pass beside the interior kernel (form 5)')
    ap.add_argument('--traces', type=int, default=8, help='B-scan traces (workload bscan)')
    ap.add_argument('--forms', type=int, default=1,
                    help='headline workload: also run the B-scan and adjoint workloads and report them (1) or not (0)')
    ap.add_argument('--forms-traces', type=int, default=16, help='B-scan traces of that report')
    ap.add_argument('--bscan-its', type=int, default=1500, help='iterations per B-scan trace')
    ap.add_argument('--bscan-graph', type=int, default=1, help='B-scan batch: iterations in HIP graphs (1) or launched directly (0)')
    ap.add_argument('--adj-its', type=int, default=2000, help='iterations of the adjoint workload')
    ap.add_argument('--fd-cells', type=int, default=0, help='finite-difference check cells (adjoint workload)')
    ap.add_argument('--adj-keep', type=int, default=1,
                    help='adjoint workload: also time the gradient with every segment kept (1) or not (0)')
    ap.add_argument('--calibrate', action='store_true',
                    help='only run a device copy of --calibrate-mib MiB (counter calibration under rocprofv3)')
    ap.add_argument('--calibrate-mib', type=int, default=1024)
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py: this solver runs on one GPU; --gpus {} is not supported '
              '(sharding over several GPUs is designed, not built)'.format(args.gpus),
              file=sys.stderr)
        return 2

    if args.calibrate:
        return calibrate(args)
    if args.workload == 'bscan':
        return bench_bscan(args)
    if args.workload == 'adjoint':
        return bench_adjoint(args)
    if args.workload == 'edges':
        return bench_edges(args)
    from gprmax_amd.hip import HipSolver, device_info
    from gprmax_amd.workloads import WORKLOADS, algorithmic_bytes_per_step, kernel_bytes

    t_build = time.perf_counter()
    G = WORKLOADS[args.workload]()
    t_build = time.perf_counter() - t_build
    K, W = args.steps, args.warmup
    if W + K > G.iterations:
        raise SystemExit('warmup + steps ({}) exceed the workload iterations ({})'.format(W + K, G.iterations))
    info = device_info(0)
    default_form = not (args.fused or args.overlap or args.passes or args.split or args.split_beside or args.merged
                        or args.split_merged)
    t_setup = time.perf_counter()
    with HipSolver(G) as s:
        s.synchronize()
        t_setup = time.perf_counter() - t_setup     # create + upload of the model (PCIe)
        s.set_launch(args.xchunk, args.block)
        s.set_step(1 if args.fused else (2 if args.overlap else (0 if args.passes else (
                   3 if args.split else (5 if args.split_beside else (6 if args.merged else (
                       7 if args.split_merged else 4)))))),
                   args.tile)
        fused = s.step_kind()
        s.run(0, W)
        s.synchronize()
        ms, _, _ = s.run_timed(W, K, per_kernel=False)
        s.synchronize()
        form = s.step_form()       # after the runs: the default form times six forms on its first long run
        choice = s.step_choice()   # -1: the default form has not timed them (a run shorter than 26 iterations)
        # Per-kernel durations (HIP events around every launch on the solver's
        # stream), in a separate run after the timed region.
        n_prof = min(50, G.iterations - W - K)
        ms4 = s.run_profiled(W + K, n_prof) if n_prof > 0 else (0.0,) * 4
        interior = s.interior_cells()
        box = s.interior_box()
        mem = s.memory_bytes()
    cells = G.nx * G.ny * G.nz
    mcells = cells * K / (ms * 1e-3) / 1e6
    if fused:
        kernels = {'step': (algorithmic_bytes_per_step(G), ms4[0])}
    elif form in (3, 5, 7):
        # the fused box is the interior box less its first plane, row and column
        box3 = (box[0] + 1, box[1], box[2] + 1, box[3], box[4] + 1, box[5])
        bh, be = kernel_bytes(G, 1, box3), kernel_bytes(G, 0, box3)
        kernels = {'interior': (bh[0] + be[0], ms4[0]), 'h_boundary': (bh[1], ms4[1]),
                   'e_boundary': (be[1], ms4[3])}
        if form == 7:
            # the H boundary pass runs inside the interior kernel's launch
            kernels = {'interior_and_h_boundary': (bh[0] + be[0] + bh[1], ms4[0]), 'e_boundary': (be[1], ms4[3])}
    elif form == 6:
        # one launch per pass over every cell
        bh, be = kernel_bytes(G, 1, box), kernel_bytes(G, 0, box)
        kernels = {'h_pass': (bh[0] + bh[1], ms4[0]), 'e_pass': (be[0] + be[1], ms4[2])}
    else:
        bh, be = kernel_bytes(G, 1, box), kernel_bytes(G, 0, box)
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
                     'step_form': step_form_label(fused, form, choice, default_form),
                     'interior_cells': interior},
        'device_bytes': mem,
        'host_build_s': round(t_build, 1),
        # the model's create + upload (host arrays over PCIe), and the rate of
        # a whole solve of the workload's iterations that includes it (not `value`)
        'setup_s': round(t_setup, 3),
        'pcie_inclusive_mcells_per_s': round(cells * G.iterations / (ms / K * 1e-3 * G.iterations + t_setup) / 1e6, 1),
    }
    if args.forms and args.workload == 'heterogeneous_halfspace':
        # The metric's other two parts, from their own workloads in the same
        # process: the batched B-scan (a 16-trace sample of the 96-trace
        # workload) and the adjoint gradient.
        # A failure here is reported in the line; the headline stands.
        forms = {}
        try:
            b = bscan_line(args, args.forms_traces)
            forms['bscan'] = {'value': b['value'], 'unit': b['unit'], 'traces': args.forms_traces,
                              'device_ratio': b['device_ratio'], 'sequential_wall_s': b['sequential_wall_s'],
                              'bscan_solve_wall_s': b['bscan_solve_wall_s'],
                              'resident_batch_ratio': b['resident_batch_ratio'],
                              'receivers_identical': b['receivers_identical']}
        except Exception as e:  # noqa: BLE001 - reported, not hidden
            forms['bscan'] = {'error': '{}: {}'.format(type(e).__name__, e)}
            print('bench.py: B-scan part failed: {}'.format(e), file=sys.stderr, flush=True)
        try:
            ad = adjoint_line(args)
            forms['adjoint'] = {'value': ad['value'], 'unit': ad['unit'], 'forward_ms': ad['forward_ms'],
                                'backward_ms': ad['backward_ms'], 'kept_states': ad.get('kept_states')}
        except Exception as e:  # noqa: BLE001 - reported, not hidden
            forms['adjoint'] = {'error': '{}: {}'.format(type(e).__name__, e)}
            print('bench.py: adjoint part failed: {}'.format(e), file=sys.stderr, flush=True)
        line['forms'] = forms
    if not args.no_cpu_baseline:
        line['cpu_baseline'] = cpu_baseline(args.workload)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == '__main__':
    sys.exit(main())
