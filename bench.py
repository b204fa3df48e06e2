"""Benchmark: MAC/IB time steps at 128^3 with 1.5M points on one MI355X.

Workload (plain FP64): periodic unit cube 128^3, 1.5M
points on two nested ellipsoidal shells with 10% clustered in a 0.05 patch,
lognormal-magnitude forces. One step = bin -> spread -> u* = u + dt f ->
divergence -> multigrid V(2,2) red-black GS to rtol 1e-10 -> gradient
correction -> interpolation -> x += dt U. Inputs are resident in HBM before
the timed region; K steps are timed between two device synchronisations.

Prints one JSON line (rank 0, N=1). `roofline` is for the interpolation
kernel (one launch per step), timed with HIP events on the launch stream;
`cpu_baseline` times the FP64 NumPy oracle on a bounded sample on this host.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = 'IB spread+interp Mpts/s & MAC steps/s at 128³ vs plain; p50/p99 served latency'
HBM_PEAK = 8.0e12


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--points', type=int, default=1_500_000)
    ap.add_argument('--n', type=int, default=128)
    ap.add_argument('--no-cpu-baseline', action='store_true')
    return ap.parse_args()


def cpu_baseline(x, F, n, cycles, dt):
    """FP64 oracle on this host: spread+gather on a 100k-point sample and one
    V-cycle at n^3, scaled to the full step (bin is included in prepare)."""
    from oracle import ib3d, mac3d
    grid = ib3d.Grid3((n,) * 3, (1., 1., 1.), periodic=True)
    sample = min(100_000, len(x))
    xs, Fs = x[:sample], F[:sample]
    t0 = time.perf_counter()
    base, phi = ib3d.prepare_compact(xs, grid)
    f = ib3d.spread_grid(Fs, base, phi, grid)
    t1 = time.perf_counter()
    ib3d.gather_grid(f, base, phi, grid)
    t2 = time.perf_counter()
    mg = mac3d.Multigrid((n,) * 3, grid.spacing, mac3d.MGOptions(smooth=2, smoother='rbgs'), periodic=True)
    rhs = np.random.default_rng(0).normal(size=(n,) * 3)
    rhs -= rhs.mean()
    t3 = time.perf_counter()
    mg.cycle(0, np.zeros_like(rhs), rhs)
    t4 = time.perf_counter()
    scale = len(x) / sample
    spread_s, interp_s, cycle_s = (t1 - t0) * scale, (t2 - t1) * scale, t4 - t3
    step_s = spread_s + interp_s + cycles * cycle_s
    return dict(value=1 / step_s, unit='steps/s', cores=1, kind='port',
                sample=f'oracle prepare+spread and gather on {sample} of {len(x)} points scaled '
                       f'linearly, one RBGS V(2,2) cycle at {n}^3 times {cycles} cycles '
                       f'(the GPU solve count); single-threaded NumPy',
                spread_mpts=len(x) / spread_s / 1e6, interp_mpts=len(x) / interp_s / 1e6,
                vcycle_s=cycle_s)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps(dict(error='this build runs on one GPU; --gpus must be 1')))
        sys.exit(2)
    import torch
    from afsi_gpu_amd import workloads
    from afsi_gpu_amd.mac import MACGrid, MGOptions
    from afsi_gpu_amd.step import IBStep

    device = torch.device('cuda:0')
    n, dt = args.n, 1e-3
    x_np, F_np = workloads.config3(args.points)
    grid = MACGrid((n, n, n), (1., 1., 1.), periodic=True)
    opts = MGOptions(smooth=2, smoother='rbgs')
    stepper = IBStep(grid, len(x_np), dt=dt, device=device, options=opts)
    x = torch.tensor(x_np, device=device)
    F = torch.tensor(F_np, device=device)
    u = grid.zeros(device=device)
    p = None
    cycles = []
    for _ in range(args.warmup):
        u, p, U, info = stepper.step(x, F, u, p)
        cycles.append(info['pressure']['cycles'])
    torch.cuda.synchronize()

    phases = ('bin', 'spread', 'project', 'interp')
    timers = [{ph: (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for ph in phases} for _ in range(args.steps)]
    t0 = time.perf_counter()
    for k in range(args.steps):
        u, p, U, info = stepper.step(x, F, u, p, timers=timers[k])
        cycles.append(info['pressure']['cycles'])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    phase_s = {ph: float(np.mean([t[ph][0].elapsed_time(t[ph][1]) for t in timers])) / 1e3
               for ph in phases}
    interp_s = phase_s['interp']
    P, N = len(x_np), n ** 3
    interp_bytes = 48 * P + 24 * N
    achieved = interp_bytes / interp_s
    result = dict(
        metric=METRIC, value=args.steps / elapsed, unit='steps/s', n_gpus=1, steps=args.steps,
        warmup=args.warmup, ms_per_step=ms_per_step, higher_is_better=True, scaling='weak',
        vs_baseline=None, dtype='f64', data='synthetic',
        config=dict(workload=f'config 3: {n}^3 periodic unit cube, {P} points (two ellipsoidal '
                             'shells + 10% in a 0.05 patch), lognormal forces; plain FP64 step, '
                             'RBGS V(2,2) to rtol 1e-10', grid=n, points=P, precision='plain-fp64'),
        roofline=dict(bound='hbm', achieved=achieved / 1e9, peak=HBM_PEAK / 1e9, unit='GB/s',
                      frac=achieved / HBM_PEAK, traffic=None, kernel='interp_tiled_kernel',
                      algorithmic_bytes=interp_bytes, avg_us=interp_s * 1e6),
        interp_mpts=P / interp_s / 1e6, spread_mpts=P / phase_s['spread'] / 1e6,
        phases_ms={k: v * 1e3 for k, v in phase_s.items()}, mg_cycles_mean=float(np.mean(cycles)))
    if not args.no_cpu_baseline:
        result['cpu_baseline'] = cpu_baseline(x_np, F_np, n, int(round(np.mean(cycles))), dt)
    print(json.dumps(result))


if __name__ == '__main__':
    main()
