"""Benchmark: MAC/IB time steps at 128^3 with 1.5M points on one MI355X.

Workload (config 3): periodic unit cube 128^3, 1.5M points on two nested
ellipsoidal shells with 10% clustered in a 0.05 patch, lognormal-magnitude
forces. One step = bin -> spread -> u* = u + dt f -> divergence -> pressure
multigrid (red-black GS V(2,2)) to rtol 1e-10 -> gradient correction ->
interpolation -> x += dt U. Two precisions on identical inputs:
  plain : everything FP64;
  mixed : FP32 V-cycles inside FP64 iterative refinement (FP64 residual,
          mean removal, velocities, point positions and transfers).
Inputs are resident in HBM before the timed region; K steps per path are
timed between device synchronisations. `value` is the mixed path's steps/s.

`roofline` is for the dominant kernel of the mixed step, the level-0
red-black sweep (algorithmic bytes: read p and rhs, write p = 3 x 4 B per
cell in FP32), timed with HIP events on the stream it is launched on;
`traffic` is its measured HBM bytes per launch (rocprofv3 counters).
`cpu_baseline` times the FP64 NumPy oracle on a bounded sample.
"""
import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = 'IB spread+interp Mpts/s & MAC steps/s at 128³ vs plain; p50/p99 served latency'
HBM_PEAK = 8.0e12
# HBM bytes per launch of the roofline kernel, measured with rocprofv3
# FETCH_SIZE / WRITE_SIZE passes and scaled by a copy kernel of known size.
TRAFFIC_FILE = ROOT / 'profiles' / 'rb_sweep_f32_128cubed_hbm_traffic.md'


def measured_traffic(n):
    """traffic_bytes_per_launch from the counter summary (measured at
    128^3), or None for another grid or when the summary is absent."""
    if n != 128 or not TRAFFIC_FILE.exists():
        return None
    for line in TRAFFIC_FILE.read_text().splitlines():
        if line.startswith('- traffic_bytes_per_launch:'):
            return float(line.split(':')[1])
    return None
PHASES = ('bin', 'spread', 'project', 'interp')


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--points', type=int, default=1_500_000)
    ap.add_argument('--n', type=int, default=128)
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--kernels-only', type=int, default=0, metavar='R',
                    help='only launch the roofline sweep and a copy kernel R times each')
    return ap.parse_args()


def cpu_baseline(x, F, n, cycles):
    """FP64 oracle on this host: prepare+spread and gather on a 100k-point
    sample and one RBGS V(2,2) cycle at n^3, scaled to the full step."""
    from oracle import ib3d, mac3d
    grid = ib3d.Grid3((n,) * 3, (1., 1., 1.), periodic=True)
    sample = min(100_000, len(x))
    t0 = time.perf_counter()
    base, phi = ib3d.prepare_compact(x[:sample], grid)
    f = ib3d.spread_grid(F[:sample], base, phi, grid)
    t1 = time.perf_counter()
    ib3d.gather_grid(f, base, phi, grid)
    t2 = time.perf_counter()
    mg = mac3d.Multigrid((n,) * 3, grid.spacing, mac3d.MGOptions(smooth=2, smoother='rbgs'),
                         periodic=True)
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
                       f'(the GPU solve count); single-threaded NumPy FP64',
                spread_mpts=len(x) / spread_s / 1e6, interp_mpts=len(x) / interp_s / 1e6,
                vcycle_s=cycle_s)


def run_path(torch, stepper, x0, F, steps, warmup):
    """Warm up, then time `steps` steps; returns timings and the final state."""
    x = x0.clone()
    grid = stepper.grid
    u = grid.zeros(device=x.device)
    p = None
    cycles = []
    for _ in range(warmup):
        u, p, U, info = stepper.step(x, F, u, p)
        cycles.append(info['pressure']['cycles'])
    timers = [{ph: (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for ph in PHASES} for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        u, p, U, info = stepper.step(x, F, u, p, timers=timers[k])
        cycles.append(info['pressure']['cycles'])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    phase_ms = {ph: float(np.mean([t[ph][0].elapsed_time(t[ph][1]) for t in timers]))
                for ph in PHASES}
    return dict(steps_per_s=steps / elapsed, ms_per_step=1e3 * elapsed / steps,
                phases_ms=phase_ms, mg_cycles_mean=float(np.mean(cycles)))


def one_step(torch, stepper, x0, F):
    """One step from rest; returns (U, velocity, ||D u*||, ||D u||)."""
    from afsi_gpu_amd.mac.flow import divergence
    x = x0.clone()
    u = stepper.grid.zeros(device=x.device)
    u, p, U, info = stepper.step(x, F, u, None)
    tentative = float(torch.linalg.vector_norm(stepper.rhs)) * stepper.dt / stepper.rho
    after = float(torch.linalg.vector_norm(divergence(u, stepper.grid)))
    return U.clone(), tuple(c.clone() for c in u), tentative, after


def time_sweep(torch, solver, repeats=50, batches=5):
    """Average duration of the level-0 sweep kernel: HIP events on the
    stream the sweeps are launched on, around `repeats` back-to-back
    launches; the mean over `batches` such brackets, divided by `repeats`."""
    from afsi_gpu_amd import _abi
    name = f'afsi_mg3d_sweep_{solver.suffix}'
    _abi.call(name, solver.plan, 0, _abi.stream())
    stream = torch.cuda.current_stream()
    start = [torch.cuda.Event(enable_timing=True) for _ in range(batches)]
    end = [torch.cuda.Event(enable_timing=True) for _ in range(batches)]
    torch.cuda.synchronize()
    for i in range(batches):
        start[i].record(stream)
        for _ in range(repeats):
            _abi.call(name, solver.plan, 0, _abi.stream())
        end[i].record(stream)
    torch.cuda.synchronize()
    return float(np.mean([a.elapsed_time(b) for a, b in zip(start, end)])) / 1e3 / repeats


def kernels_only(torch, n, repeats):
    """For hardware-counter passes: `repeats` launches of the level-0 FP32
    red-black sweep at n^3 (the roofline kernel), then `repeats` launches of
    a plain copy kernel of known size (out = 1.0 * src over n^3 float64:
    8 n^3 bytes read, 8 n^3 bytes written), which scales the counters."""
    from afsi_gpu_amd.mac import MACGrid, MGOptions, MixedMultigrid
    grid = MACGrid((n, n, n), (1., 1., 1.), periodic=True)
    solver = MixedMultigrid(grid, device='cuda:0', options=MGOptions(smooth=2, smoother='rbgs'),
                            backend='workspace').inner
    gen = torch.Generator(device='cuda:0').manual_seed(0)
    solver.p[0].copy_(torch.randn(solver.shapes[0], device='cuda:0', generator=gen))
    solver.rhs[0].copy_(torch.randn(solver.shapes[0], device='cuda:0', generator=gen))
    sweep_s = time_sweep(torch, solver, repeats=repeats, batches=1)
    src = torch.randn(n ** 3, device='cuda:0', dtype=torch.float64, generator=gen)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(repeats):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(json.dumps(dict(kernels_only=True, n=n, repeats=repeats, sweep_us=sweep_s * 1e6,
                          sweep_algorithmic_bytes=12 * n ** 3, copy_bytes_read=8 * n ** 3,
                          copy_bytes_written=8 * n ** 3)))


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps(dict(error='this build runs on one GPU; --gpus must be 1')))
        sys.exit(2)
    import torch
    if args.kernels_only:
        kernels_only(torch, args.n, args.kernels_only)
        return
    from afsi_gpu_amd import workloads
    from afsi_gpu_amd.mac import MACGrid, MGOptions
    from afsi_gpu_amd.step import IBStep

    device = torch.device('cuda:0')
    n, dt = args.n, 1e-3
    x_np, F_np = workloads.config3(args.points)
    P, N = len(x_np), n ** 3
    grid = MACGrid((n, n, n), (1., 1., 1.), periodic=True)
    opts = MGOptions(smooth=2, smoother='rbgs')
    x0 = torch.tensor(x_np, device=device)
    F = torch.tensor(F_np, device=device)
    results = {}
    steppers = {}
    for precision in ('fp64', 'mixed'):
        stepper = IBStep(grid, P, dt=dt, device=device, options=opts, precision=precision)
        steppers[precision] = stepper
        results[precision] = run_path(torch, stepper, x0, F, args.steps, args.warmup)
    # accuracy of the mixed step against the plain FP64 step (same inputs)
    U64, u64, div0, div64 = one_step(torch, steppers['fp64'], x0, F)
    Umx, umx, _, divmx = one_step(torch, steppers['mixed'], x0, F)
    u_err = float(torch.linalg.vector_norm(Umx - U64) / torch.linalg.vector_norm(U64))
    face_err = max(float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
                   for a, b in zip(umx, u64))
    inner = steppers['mixed'].flow.pressure_solver.inner
    sweep_s = time_sweep(torch, inner)
    sweep_bytes = 3 * 4 * N
    achieved = sweep_bytes / sweep_s
    mixed, plain = results['mixed'], results['fp64']
    result = dict(
        metric=METRIC, value=mixed['steps_per_s'], unit='steps/s', n_gpus=1, steps=args.steps,
        warmup=args.warmup, ms_per_step=mixed['ms_per_step'], higher_is_better=True,
        scaling='weak', vs_baseline=None, dtype='f32+f64', data='synthetic',
        config=dict(workload=f'config 3: {n}^3 periodic unit cube, {P} points (two ellipsoidal '
                             'shells + 10% in a 0.05 patch), lognormal forces; IB step with '
                             'red-black GS V(2,2) to rtol 1e-10; mixed = FP32 V-cycles in FP64 '
                             'refinement', grid=n, points=P, precision='mixed'),
        roofline=dict(bound='hbm', achieved=achieved / 1e9, peak=HBM_PEAK / 1e9, unit='GB/s',
                      frac=achieved / HBM_PEAK, traffic=measured_traffic(n),
                      kernel='rb_rows_kernel<float,true,4,4> (level-0 sweep)',
                      algorithmic_bytes=sweep_bytes, avg_us=sweep_s * 1e6),
        plain_fp64=dict(steps_per_s=plain['steps_per_s'], ms_per_step=plain['ms_per_step'],
                        phases_ms=plain['phases_ms'], mg_cycles_mean=plain['mg_cycles_mean'],
                        spread_mpts=P / plain['phases_ms']['spread'] / 1e3,
                        interp_mpts=P / plain['phases_ms']['interp'] / 1e3),
        mixed=dict(steps_per_s=mixed['steps_per_s'], ms_per_step=mixed['ms_per_step'],
                   phases_ms=mixed['phases_ms'], mg_cycles_mean=mixed['mg_cycles_mean'],
                   spread_mpts=P / mixed['phases_ms']['spread'] / 1e3,
                   interp_mpts=P / mixed['phases_ms']['interp'] / 1e3),
        mixed_vs_plain=mixed['steps_per_s'] / plain['steps_per_s'],
        accuracy=dict(interp_velocity_rel_l2_mixed_vs_fp64=u_err,
                      face_velocity_rel_l2_mixed_vs_fp64=face_err,
                      divergence_rel_fp64=div64 / div0, divergence_rel_mixed=divmx / div0))
    if not args.no_cpu_baseline:
        result['cpu_baseline'] = cpu_baseline(x_np, F_np, n, int(round(plain['mg_cycles_mean'])))
    print(json.dumps(result))


if __name__ == '__main__':
    main()
