"""Benchmark: MAC/IB time steps at 128^3 with 1.5M points on one MI355X.

Workload: periodic unit cube 128^3, 1.5M points on two nested
ellipsoidal shells with 10% clustered in a 0.05 patch, lognormal-magnitude
forces. One step = bin -> spread -> u* = u + dt f -> divergence -> pressure
multigrid (red-black GS V(2,2)) to rtol 1e-10 -> gradient correction ->
interpolation -> x += dt U. Two precisions on identical inputs:
  plain : everything FP64;
  mixed : FP32 V-cycles inside FP64 iterative refinement (FP64 residual,
          mean removal, velocities, point positions and transfers).
Inputs are resident in HBM before the timed region; K steps per path are
timed between device synchronisations, with Python's cyclic garbage
collector paused inside timed regions. `value` is the mixed path's steps/s.

`batched`: 64 uneven instances stepped together (3D: one
batched mixed pressure solve, one launch per multigrid phase) against the
same instances stepped one at a time on the plain path. `serving`
(bounded sample): a Poisson request stream served with
continuous batching, p50/p99 latency.

`roofline` is for the dominant kernel of the mixed step, the level-0
red-black sweep (algorithmic bytes: read p and rhs, write p = 3 x 4 B per
cell in FP32), timed with HIP events on the stream it is launched on;
`traffic` is its measured HBM bytes per launch (rocprofv3 counters).
`cpu_baseline` times the reference's CPU tensor path (restated in torch,
FP64, all host threads) on full steps of the same workload.
"""
import argparse
import gc
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
SPREAD_TRAFFIC_FILE = ROOT / 'profiles' / 'spread_128cubed_1p5M_hbm_traffic.md'


def measured_traffic(n, path=TRAFFIC_FILE):
    """traffic_bytes_per_launch from a counter summary (measured at 128^3
    with 1.5M points), or None for another grid or when it is absent."""
    if n != 128 or not path.exists():
        return None
    for line in path.read_text().splitlines():
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
    ap.add_argument('--no-dropin', action='store_true', help='skip the reference-API drop-in sample')
    ap.add_argument('--kernels-only', type=int, default=0, metavar='R',
                    help='only launch the roofline sweep, the transfers and a copy kernel R times each')
    ap.add_argument('--serve-seconds', type=float, default=60.,
                    help='arrival window of the served-latency sample (0: skip)')
    ap.add_argument('--batch-instances', type=int, default=64,
                    help='instances of the batched sample (0: skip)')
    ap.add_argument('--fe-level', type=int, default=7,
                    help='sphere subdivision level of the FE interaction-point sample (0: skip)')
    ap.add_argument('--no-p2', action='store_true', help='skip the P2 fused-transfer sample')
    ap.add_argument('--no-assembled', action='store_true', help='skip the CSR-assembled P1 sample')
    ap.add_argument('--only', choices=('p2', 'fe', 'batched', 'serving', 'assembled'),
                    help='run one secondary sample alone and print its JSON (for profiling)')
    return ap.parse_args()


def host_cpu():
    """CPU model, sockets, cores and threads of this host (lscpu)."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(['lscpu'], capture_output=True, text=True, timeout=10).stdout
    except (OSError, subprocess.SubprocessError):
        return info
    for line in out.splitlines():
        key, _, value = line.partition(':')
        if key.strip() in ('Model name', 'Socket(s)', 'Core(s) per socket', 'Thread(s) per core', 'CPU(s)'):
            info[key.strip()] = value.strip()
    return info


def cpu_threads():
    """Threads the baseline may use: the CPUs this process may run on,
    capped by OMP_NUM_THREADS where the host sets it (a shared host's CPU
    share)."""
    import os
    cores = len(os.sched_getaffinity(0))
    env = os.environ.get('OMP_NUM_THREADS', '')
    return min(cores, int(env)) if env.isdigit() and int(env) > 0 else cores


def cpu_baseline(x, F, n, dt, warmup=2, repeats=5):
    """The reference's CPU tensor path (oracle/torch_cpu.py: chunked
    index_add_ spread and gather, Jacobi V(4,4) multigrid to rtol 1e-10
    checked every 2 cycles) on this host's cores: `warmup` untimed steps,
    then the median of `repeats` timed full steps on all points, each from
    the previous step's velocity and with the pressure solved from zero (as
    the GPU paths)."""
    import torch
    from oracle import torch_cpu
    threads = cpu_threads()
    torch.set_num_threads(threads)
    shape, spacing = (n,) * 3, (1. / n,) * 3
    mg = torch_cpu.JacobiMultigrid(shape, spacing, smooth=4)
    xt, Ft = torch.tensor(x), torch.tensor(F)
    u = tuple(torch.zeros(shape, dtype=torch.float64) for _ in range(3))
    p, times, cycles = None, [], []
    for k in range(warmup + repeats):
        t0 = time.perf_counter()
        _, u, p, _, c = torch_cpu.ib_step(xt, Ft, u, None, shape, spacing, mg, dt=dt)
        if k >= warmup:
            times.append(time.perf_counter() - t0)
            cycles.append(c)
    step_s = float(np.median(times))
    base, phi = torch_cpu.prepare(xt, shape, spacing)
    t0 = time.perf_counter()
    f = torch_cpu.spread(Ft, base, phi, shape, spacing[0] ** 3)
    t1 = time.perf_counter()
    torch_cpu.gather(f, base, phi, shape)
    t2 = time.perf_counter()
    return dict(value=1 / step_s, unit='steps/s', cores=threads, kind='port',
                sample=f'{repeats} full steps (median, after {warmup} warm-up steps) of the reference CPU '
                       f'tensor path restated in torch (oracle/torch_cpu.py) on all {len(x)} points at '
                       f'{n}^3: prepare, chunked index_add_ spread, Jacobi V(4,4) multigrid to rtol 1e-10, '
                       f'projection, gather; torch threads = {threads}',
                ms_per_step=1e3 * step_s, mg_cycles=cycles, spread_mpts=len(x) / (t1 - t0) / 1e6,
                interp_mpts=len(x) / (t2 - t1) / 1e6, host=host_cpu())


class paused_gc:
    """Collect, then keep Python's cyclic collector paused inside the timed
    region (a full collection in a torch-sized heap stalls the host for
    tens of milliseconds and idles the GPU; no GPU work is skipped)."""

    def __enter__(self):
        gc.collect()
        gc.disable()

    def __exit__(self, *exc):
        gc.enable()


def run_path(torch, stepper, x0, F, steps, warmup, phase_steps=5):
    """Warm up, then time `steps` steps; returns timings and the final state.
    Every step is the full IB step (binning, spread, projection, interpolation,
    x += dt U) with the pressure solved from zero, so each step runs the whole
    solve to tolerance; dt keeps the points where the workload puts them (the
    largest displacement over all steps is reported in cells). The timed steps
    run as the stepper runs them (HIP graphs when it has them); the phase split
    comes from `phase_steps` further steps with HIP events around each phase,
    which run as plain launches."""
    x = x0.clone()
    grid = stepper.grid
    u = grid.zeros(device=x.device)
    cycles = []
    for _ in range(warmup):
        u, p, U, info = stepper.step(x, F, u, None)
        cycles.append(info['pressure']['cycles'])
    with paused_gc():
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(steps):
            u, p, U, info = stepper.step(x, F, u, None)
            cycles.append(info['pressure']['cycles'])
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
    timers = [{ph: (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for ph in PHASES} for _ in range(phase_steps)]
    with paused_gc():
        for k in range(phase_steps):
            u, p, U, info = stepper.step(x, F, u, None, timers=timers[k])
        torch.cuda.synchronize()
    phase_ms = {ph: float(np.mean([t[ph][0].elapsed_time(t[ph][1]) for t in timers]))
                for ph in PHASES}
    moved = float((x - x0).abs().max()) / min(grid.spacing)
    return dict(steps_per_s=steps / elapsed, ms_per_step=1e3 * elapsed / steps,
                phases_ms=phase_ms, mg_cycles_mean=float(np.mean(cycles)), max_displacement_cells=moved)


def one_step(torch, stepper, x0, F):
    """One step from rest; returns (U, velocity, ||D u*||, ||D u||) on the host."""
    from afsi_gpu_amd.mac.flow import divergence
    x = x0.clone()
    u = stepper.grid.zeros(device=x.device)
    u, p, U, info = stepper.step(x, F, u, None)
    tentative = float(torch.linalg.vector_norm(stepper.last_rhs)) * stepper.dt / stepper.rho
    after = float(torch.linalg.vector_norm(divergence(u, stepper.grid)))
    return U.cpu(), tuple(c.cpu() for c in u), tentative, after


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


def time_spread(torch, stepper, x, F, repeats=20, batches=3):
    """Average duration of one tiled spread call (afsi_spread_tiled_f64: the
    zero fill of the three face arrays, the work-list scan and
    spread_tiled_kernel) on the stepper's binned points: HIP events on the
    launch stream around `repeats` back-to-back calls, the mean over
    `batches` brackets."""
    bins = stepper.binner(x)
    out = stepper.grid.zeros(device=x.device, dtype=torch.float64)
    stepper.transfer.spread(x, F, bins, out=out)
    stream = torch.cuda.current_stream()
    start = [torch.cuda.Event(enable_timing=True) for _ in range(batches)]
    end = [torch.cuda.Event(enable_timing=True) for _ in range(batches)]
    torch.cuda.synchronize()
    for i in range(batches):
        start[i].record(stream)
        for _ in range(repeats):
            stepper.transfer.spread(x, F, bins, out=out)
        end[i].record(stream)
    torch.cuda.synchronize()
    return float(np.mean([a.elapsed_time(b) for a, b in zip(start, end)])) / 1e3 / repeats


def serving_sample(torch, seconds, rate=200.):
    """Config 5 request stream (Poisson arrivals at `rate`/s; 60% 2D
    128x32..512x128 with 2k-20k points, 30% 3D 32^3-64^3 with 20k-200k,
    10% 128^3 with 0.5M-1.5M; 10 requests per instance) over a window of
    `seconds`, served with continuous batching (admission, launch and
    completion overlapped on separate host threads), once with the mixed
    batched 3D solve and once with plain FP64 per-instance solves: latency
    percentiles and every request's divergence ratio ||D u|| / ||D u*||."""
    from afsi_gpu_amd import workloads
    from afsi_gpu_amd.serve import Server
    times, inst, specs = workloads.request_stream(workloads.GEOMETRY_SEED + 5, rate=rate,
                                                  duration=seconds)
    warm_t, warm_i, warm_s = workloads.request_stream(workloads.GEOMETRY_SEED + 6, rate=rate,
                                                      duration=.2)
    out = {}
    for precision in ('mixed', 'fp64'):
        Server(device='cuda:0', precision=precision).run(warm_t, warm_i, warm_s)  # compile/cache warm-up
        with paused_gc():
            stats = Server(device='cuda:0', precision=precision).run(times, inst, specs)
        out[precision] = stats.summary()
        print(f'serving {precision}: {out[precision]["requests"]} requests, p99 {out[precision]["p99_ms"]:.2f} ms',
              file=sys.stderr, flush=True)
    out.update(rate_per_s=rate, window_s=seconds,
               sample=f'{len(times)} requests over {seconds:g} s of the 60 s stream, '
                      f'{len(specs)} instances, each precision')
    return out


def fe_sample(torch, n=128, level=7, steps=10, warmup=2):
    """FE interaction-point steps (FEIBStep, mixed-precision pressure): a P1
    sphere of radius 0.3 (icosahedron subdivided `level` times, three
    quadrature points per triangle) in an n^3 periodic unit cube, nodal
    forces of unit surface density turning between two random fields
    (F_k = F0 cos 0.3k + F1 sin 0.3k); consistent-mass Jacobi PCG on both
    transfers (rtol 1e-12, warm started)."""
    from afsi_gpu_amd import workloads
    from afsi_gpu_amd.mac import MACGrid, MGOptions
    from afsi_gpu_amd.mac.fe import FEGeometry
    from afsi_gpu_amd.step import FEIBStep
    device = torch.device('cuda:0')
    nodes, cells, values, weights = workloads.fe_sphere(level)
    t = lambda a, dt=torch.float64: torch.tensor(a, device=device, dtype=dt)
    geometry = FEGeometry(t(cells, torch.int64), t(values), t(weights), len(nodes))
    grid = MACGrid((n,) * 3, (1., 1., 1.), periodic=True)
    stepper = FEIBStep(grid, geometry, dt=1e-3, device=device, options=MGOptions(smooth=2, smoother='rbgs'),
                       precision='mixed', warm_start=True)
    x = t(nodes)
    # O(1) surface force density: nodal forces scale with the nodal area. The
    # force turns between two random fields from step to step, so each
    # warm-started mass solve has a new right-hand side to work on.
    gen = np.random.default_rng(7)
    scale = weights.sum() / len(nodes)
    F0, F1 = (t(gen.normal(size=nodes.shape) * scale) for _ in range(2))
    forces = [F0 * np.cos(.3 * k) + F1 * np.sin(.3 * k) for k in range(warmup + steps)]
    u = grid.zeros(device=device)
    p = None
    for k in range(warmup):
        x, u, p, U, info = stepper.step(x, forces[k], u, p)
    timers = [{ph: (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for ph in PHASES} for _ in range(steps)]
    iters = []
    with paused_gc():
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(steps):
            x, u, p, U, info = stepper.step(x, forces[warmup + k], u, p, timers=timers[k])
            iters.append((info['force_mass'].iterations, info['velocity_mass'].iterations))
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
    phases = {ph: float(np.mean([tm[ph][0].elapsed_time(tm[ph][1]) for tm in timers])) for ph in PHASES}
    points = cells.shape[0] * values.shape[0]
    return dict(steps_per_s=steps / elapsed, ms_per_step=1e3 * elapsed / steps, phases_ms=phases,
                nodes=len(nodes), elements=int(cells.shape[0]), interaction_points=int(points),
                mass_pcg_iterations_mean=[float(np.mean([a for a, _ in iters])), float(np.mean([b for _, b in iters]))],
                spread_with_mass_solve_mpts=points / phases['spread'] / 1e3,
                interp_with_mass_solve_mpts=points / phases['interp'] / 1e3,
                workload=f'P1 sphere r=0.3, level {level} ({len(nodes)} nodes, {points} interaction points) '
                         f'in {n}^3 periodic; FEIBStep (prepare, mass-PCG spread, mixed projection, '
                         f'interpolate + mass PCG), bin phase = B x + binning')


def p2_sample(torch, n=64, reps=10, warmup=2):
    """P2 fused cell-block transfer (FETransfer(fused=True)) on an idealised
    LV wall of straight P2 tetrahedra with the 64-point rule (60,480
    elements, 3,870,720 interaction points) in an n^3 closed unit box:
    prepare, spread with the force mass solve, interpolate with the velocity
    mass solve (cold start: every solve from zero), the mass solve alone, and
    the two fused kernels, against the
    unfused transfer (points, binning, tiled kernels) on the same inputs.
    Algorithmic bytes of the fused kernels: tables 240 B and weight 8 B per
    point, 8 B per touched face value (written by the spread, read by the
    gather) and 24 B per node."""
    from afsi_gpu_amd import workloads
    from afsi_gpu_amd.mac import MACGrid
    from afsi_gpu_amd.mac import p2_transfer as p2
    from afsi_gpu_amd.mac.fe import FEGeometry, FETransfer
    device = torch.device('cuda:0')
    nodes, cells, values, weights = workloads.lv_p2()
    t = lambda a, dt=torch.float64: torch.tensor(a, device=device, dtype=dt)
    geometry = FEGeometry(t(cells, torch.int64), t(values), t(weights), len(nodes))
    grid = MACGrid((n,) * 3, (1., 1., 1.))
    gen = workloads.rng(workloads.GEOMETRY_SEED + 7)
    x = t(nodes)
    F = t(gen.normal(size=nodes.shape) * (weights.sum() / len(nodes)))
    coords = [grid.coordinates(c, device=device) for c in range(3)]
    velocity = tuple(torch.sin(2 * np.pi * coords[c][..., (c + 1) % 3]) * torch.cos(np.pi * coords[c][..., c])
                     for c in range(3))

    def timed(fn):
        for _ in range(warmup):
            fn()
        start = [torch.cuda.Event(enable_timing=True) for _ in range(reps)]
        stop = [torch.cuda.Event(enable_timing=True) for _ in range(reps)]
        torch.cuda.synchronize()
        for k in range(reps):
            start[k].record()
            out = fn()
            stop[k].record()
        torch.cuda.synchronize()
        return float(np.mean([a.elapsed_time(b) for a, b in zip(start, stop)])), out

    out = {}
    for name, fused in (('fused', True), ('unfused', False)):
        tr = FETransfer(grid, geometry, fused=fused)
        ms = {}
        ms['prepare'], st = timed(lambda: tr.prepare(x))
        ms['spread_with_mass'], (faces, finfo) = timed(lambda: tr.spread(F, st))
        ms['interp_with_mass'], (U, uinfo) = timed(lambda: tr.interpolate(velocity, st))
        ms['mass_solve'], _ = timed(lambda: tr.solve_mass(F, None))
        coefficient, _ = tr.solve_mass(F, None)
        if fused:
            ms['spread_kernel'], faces = timed(lambda: p2.spread(tr, coefficient, st))
            ms['gather_kernel'], _ = timed(lambda: p2.assemble_velocity(tr, velocity, st))
        else:
            force_q = tr.weighted_force(coefficient)
            ms['spread_kernel'], _ = timed(lambda: tr.spread_grid(force_q, st))
            ms['gather_kernel'], _ = timed(lambda: tr.gather_grid(velocity, st))
        out[name] = dict(ms=ms, pcg_iterations=[finfo.iterations, uinfo.iterations])
        if fused:
            touched = int(sum(int((f != 0).sum()) for f in faces))
            del st
        else:
            del st, force_q
        del tr
        torch.cuda.empty_cache()
    P = cells.shape[0] * values.shape[0]
    fused_ms = out['fused']['ms']
    spread_bytes = 248 * P + 8 * touched + 24 * len(nodes)
    roof = lambda b, ms: dict(achieved=b / (ms * 1e-3) / 1e9, frac=b / (ms * 1e-3) / HBM_PEAK,
                              algorithmic_bytes=b)
    return dict(workload=f'idealised LV, P2 tetrahedra, 64-point rule: {cells.shape[0]} elements, {len(nodes)} '
                         f'nodes, {P} interaction points in {n}^3 closed unit box',
                points=P, elements=int(cells.shape[0]), nodes=len(nodes), touched_faces=touched,
                fused=out['fused'], unfused=out['unfused'],
                spread_mpts=P / fused_ms['spread_kernel'] / 1e3, gather_mpts=P / fused_ms['gather_kernel'] / 1e3,
                spread_roofline=roof(spread_bytes, fused_ms['spread_kernel']),
                gather_roofline=roof(spread_bytes, fused_ms['gather_kernel']))


def assembled_sample(torch, n=128, reps=5, warmup=1):
    """CSR-assembled P1 operator (AssembledTransfer) on the idealised LV wall
    as P1 tetrahedra with the 64-point rule (60,480 elements, 3,870,720
    interaction points) in an n^3 closed unit box: the CSR build (prepare:
    points, compact tables, per-cell contraction into the hash table, sort,
    B and B^T), spread and interpolate with the mass solves, against the
    tiled FE transfer on the same inputs."""
    from afsi_gpu_amd import workloads
    from afsi_gpu_amd.mac import MACGrid
    from afsi_gpu_amd.mac.assembled import AssembledTransfer
    from afsi_gpu_amd.mac.fe import FEGeometry, FETransfer
    device = torch.device('cuda:0')
    nodes, cells, values, weights = workloads.lv_p1()
    t = lambda a, dt=torch.float64: torch.tensor(a, device=device, dtype=dt)
    geometry = FEGeometry(t(cells, torch.int64), t(values), t(weights), len(nodes))
    grid = MACGrid((n,) * 3, (1., 1., 1.))
    gen = workloads.rng(workloads.GEOMETRY_SEED + 8)
    x = t(nodes)
    F = t(gen.normal(size=nodes.shape) * (weights.sum() / len(nodes)))
    coords = [grid.coordinates(c, device=device) for c in range(3)]
    velocity = tuple(torch.sin(2 * np.pi * coords[c][..., (c + 1) % 3]) for c in range(3))

    def timed(fn):
        for _ in range(warmup):
            out = fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            out = fn()
        torch.cuda.synchronize()
        return 1e3 * (time.perf_counter() - t0) / reps, out

    out = {}
    for name, make in (('assembled', lambda: AssembledTransfer(grid, geometry, max_entries=128_000_000)),
                       ('tiled', lambda: FETransfer(grid, geometry))):
        tr = make()
        ms = {}
        ms['prepare'], st = timed(lambda: tr.prepare(x))
        ms['spread_with_mass'], _ = timed(lambda: tr.spread(F, st))
        ms['interp_with_mass'], _ = timed(lambda: tr.interpolate(velocity, st))
        ms['mass_solve'], _ = timed(lambda: tr.solve_mass(F, None))
        out[name] = dict(ms=ms)
        if name == 'assembled':
            out[name]['nnz'] = st.assembly['total_nnz']
            out[name]['hash_attempts'] = [c['hash_attempts'] for c in st.assembly['components']]
            out[name]['csr_bytes'] = st.storage_bytes
        del st, tr
        torch.cuda.empty_cache()
    P = cells.shape[0] * values.shape[0]
    return dict(workload=f'idealised LV, P1 tetrahedra, 64-point rule: {cells.shape[0]} elements, {len(nodes)} '
                         f'nodes, {P} interaction points in {n}^3 closed unit box', points=P, **out)


def batched_sample(torch, instances, steps=3, warmup=1):
    """Config 4: `instances` independent instances, grids uniform over
    {32^3, 48^3, 64^3, 96^3, 128x32, 256x64, 512x128}, points log-uniform in
    [1e4, 3e5] (random shells / sinusoidal curves), N(0,1) forces. Batched:
    the 3D instances with one batched mixed pressure solve and the 2D
    channels with one batched FP64 step (one launch per phase each);
    against the same instances stepped one at a time on the plain FP64
    paths."""
    from afsi_gpu_amd import workloads
    from afsi_gpu_amd.batch import BatchedIBStep, ConcurrentGroups
    from afsi_gpu_amd.mac import MACGrid, MGOptions
    from afsi_gpu_amd.mac.multigrid import MIXED_INNER_CYCLES, MIXED_RELAX
    from afsi_gpu_amd.mac2d import ChannelGrid, ChannelIBStep
    from afsi_gpu_amd.mac2d.batch import BatchedChannelStep
    from afsi_gpu_amd.step import IBStep
    dev = torch.device('cuda:0')
    opts = MGOptions(smooth=2, smoother='rbgs')
    gen = workloads.rng(workloads.GEOMETRY_SEED + 4)
    specs = [workloads.instance_spec(gen, 'batch') for _ in range(instances)]
    data = [workloads.instance_points(gen, shape, count) for shape, count in specs]
    i3 = [i for i, (s, _) in enumerate(specs) if len(s) == 3]
    i2 = [i for i, (s, _) in enumerate(specs) if len(s) == 2]
    points = sum(c for _, c in specs)
    g3 = [MACGrid(specs[i][0], (1., 1., 1.), periodic=True) for i in i3]
    g2 = [ChannelGrid(specs[i][0], (4., 1.), periodic=True) for i in i2]
    s2 = [ChannelIBStep(g, dt=1e-3, device=dev, options=opts) for g in g2]

    def fresh():
        xs = [torch.tensor(data[i][0], device=dev) for i in range(instances)]
        fs = [torch.tensor(data[i][1], device=dev) for i in range(instances)]
        return xs, fs

    def run(step_all):
        xs, fs = fresh()
        state = {}
        for k in range(warmup):
            step_all(xs, fs, state)
        with paused_gc():
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for k in range(steps):
                step_all(xs, fs, state)
            torch.cuda.synchronize()
            return (time.perf_counter() - t0) / steps

    relaxed = MGOptions(smooth=2, smoother='rbgs', relax=MIXED_RELAX, inner_cycles=MIXED_INNER_CYCLES)
    batch3 = BatchedIBStep(g3, [specs[i][1] for i in i3], dt=1e-3, device=dev, options=relaxed)
    batch2 = BatchedChannelStep(g2, [specs[i][1] for i in i2], dt=1e-3, device=dev, options=opts) if i2 else None

    groups = ConcurrentGroups(dev)

    def batched(xs, fs, state):
        v3 = state.setdefault('v3', [g.zeros(device=dev) for g in g3])
        p3 = state.get('p3')
        step3 = lambda: batch3.step([xs[i] for i in i3], [fs[i] for i in i3], v3, p3)
        if batch2 is None:
            state['p3'] = step3()[1]
            return
        if 'b2' not in state:
            # the 2D instances live in the batch's own buffers
            batch2.reset()
            for j, i in enumerate(i2):
                batch2.x[j].copy_(xs[i])
                batch2.force[j].copy_(fs[i])
            state['b2'] = True
        # the two groups step concurrently, each on its own stream and host thread
        (v3, p3, _, _), _ = groups.run(step3, batch2.step)
        state['p3'] = p3

    plain3 = [IBStep(g, specs[i][1], dt=1e-3, device=dev, options=opts, precision='fp64')
              for g, i in zip(g3, i3)]

    def sequential(xs, fs, state):
        for j, i in enumerate(i3):
            v = state.setdefault(('v3', j), g3[j].zeros(device=dev))
            v, p, _, _ = plain3[j].step(xs[i], fs[i], v, state.get(('p3', j)))
            state[('v3', j)], state[('p3', j)] = v, p
        for j, i in enumerate(i2):
            v = state.setdefault(('v2', j), g2[j].zeros(device=dev))
            state[('v2', j)], state[('p2', j)], _, _ = s2[j].step(xs[i], fs[i], v, state.get(('p2', j)))

    t_batch = run(batched)
    groups.close()
    t_seq = run(sequential)
    return dict(instances=instances, instances_3d=len(i3), instances_2d=len(i2), points=points,
                batched=dict(steps_per_s=1 / t_batch, instance_steps_per_s=instances / t_batch,
                             mpts_per_s=points / t_batch / 1e6, ms_per_step=t_batch * 1e3),
                sequential_plain=dict(steps_per_s=1 / t_seq, instance_steps_per_s=instances / t_seq,
                                      mpts_per_s=points / t_seq / 1e6, ms_per_step=t_seq * 1e3),
                speedup=t_seq / t_batch,
                note='one step = every instance advanced once; batched = 3D instances in mixed precision '
                     f'(FP32 V-cycles with red-black over-relaxation {MIXED_RELAX}, {MIXED_INNER_CYCLES} per FP64 '
                     'refinement step) and 2D channels in FP64, '
                     'each group with one launch per phase over all its instances, the two groups '
                     'concurrently on their own streams and host threads; sequential = the same '
                     'instances stepped one at a time on the plain FP64 paths')


def dropin_sample(torch, x0, F, n, reps=5, warmup=2):
    """The reference-API drop-in path (``mac.cuda_ib.compact_spread`` /
    ``compact_gather`` on a ``CompactStencil`` from ``mac.transfer.from_points``,
    with the reference's error check: one flag read-back per call) on the same
    points in the reference's closed n^3 box, timed with HIP events on the
    current stream; the tiled path on the same closed box beside it."""
    from afsi_gpu_amd.mac import MACGrid, cuda_ib
    from afsi_gpu_amd.mac.transfer import TiledTransfer, bin_points, from_points
    grid = MACGrid((n,) * 3, (1., 1., 1.))
    P = x0.shape[0]

    def timed(fn):
        for _ in range(warmup):
            out = fn()
        start = [torch.cuda.Event(enable_timing=True) for _ in range(reps)]
        stop = [torch.cuda.Event(enable_timing=True) for _ in range(reps)]
        torch.cuda.synchronize()
        for k in range(reps):
            start[k].record()
            out = fn()
            stop[k].record()
        torch.cuda.synchronize()
        return float(np.median([a.elapsed_time(b) for a, b in zip(start, stop)])), out

    ms = {}
    ms['prepare'], st = timed(lambda: from_points(x0, grid))
    ms['compact_spread'], f = timed(lambda: cuda_ib.compact_spread(grid, F, st))
    ms['compact_gather'], _ = timed(lambda: cuda_ib.compact_gather(grid, f, st))
    del st
    torch.cuda.empty_cache()
    tr = TiledTransfer(grid, x0.device)
    ms['tiled_bin'], bins = timed(lambda: bin_points(x0, grid))
    ms['tiled_spread'], f2 = timed(lambda: tr.spread(x0, F, bins))
    ms['tiled_interp'], _ = timed(lambda: tr.interpolate(x0, f2, bins))
    diff = max(float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b)) for a, b in zip(f2, f))
    return dict(workload=f'{P} points of the clustered-shells workload in the closed {n}^3 unit box',
                ms=ms, compact_spread_mpts=P / ms['compact_spread'] / 1e3,
                compact_gather_mpts=P / ms['compact_gather'] / 1e3,
                tiled_spread_mpts=P / ms['tiled_spread'] / 1e3, tiled_interp_mpts=P / ms['tiled_interp'] / 1e3,
                tiled_vs_compact_spread_rel_l2=diff)


def oracle_accuracy(torch, x_np, F_np, n, dt, states):
    """Each path's one step from rest against the FP64 NumPy oracle's step on
    the same inputs (oracle/step.ib_step, red-black GS V(2,2), rtol 1e-10):
    l2-relative error of the interpolated point velocity and of the projected
    face velocity, and the post-projection divergence relative to the
    tentative field's."""
    from oracle import ib3d, mac3d
    from oracle.step import ib_step
    grid = ib3d.Grid3((n,) * 3, (1., 1., 1.), periodic=True)
    mg = mac3d.Multigrid(grid.shape, grid.spacing, mac3d.MGOptions(smooth=2, smoother='rbgs'), periodic=True)
    u0 = tuple(np.zeros(grid.face_shape(c)) for c in range(3))
    ref = ib_step(x_np, F_np, u0, None, grid, mg, dt=dt)
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    out = dict(oracle_cycles=int(ref['info']['cycles']))
    for name, (U, u, div_star, div) in states.items():
        out[name] = dict(interp_velocity_rel_l2_vs_oracle=rel(U.cpu().numpy(), ref['point_velocity']),
                         face_velocity_rel_l2_vs_oracle=max(rel(u[c].cpu().numpy(), ref['velocity'][c])
                                                            for c in range(3)),
                         divergence_rel=div / div_star)
    return out


def kernels_only(torch, n, repeats, points):
    """For hardware-counter passes: `repeats` launches of the level-0 FP32
    red-black sweep at n^3 (the roofline kernel), `repeats` binnings, spreads
    and interpolations of the clustered-shells workload (`points` points),
    then `repeats` launches of a plain copy kernel of known size (out = 1.0 *
    src over n^3 float64: 8 n^3 bytes read, 8 n^3 bytes written), which
    scales the counters."""
    from afsi_gpu_amd import workloads
    from afsi_gpu_amd.mac import MACGrid, MGOptions, MixedMultigrid
    from afsi_gpu_amd.mac.transfer import Binner, TiledTransfer
    grid = MACGrid((n, n, n), (1., 1., 1.), periodic=True)
    solver = MixedMultigrid(grid, device='cuda:0', options=MGOptions(smooth=2, smoother='rbgs'),
                            backend='workspace').inner
    gen = torch.Generator(device='cuda:0').manual_seed(0)
    solver.p[0].copy_(torch.randn(solver.shapes[0], device='cuda:0', generator=gen))
    solver.rhs[0].copy_(torch.randn(solver.shapes[0], device='cuda:0', generator=gen))
    sweep_s = time_sweep(torch, solver, repeats=repeats, batches=1)
    x_np, F_np = workloads.clustered_shells_128(points)
    x = torch.tensor(x_np, device='cuda:0')
    F = torch.tensor(F_np, device='cuda:0')
    binner = Binner(grid, len(x_np), x.device)
    transfer = TiledTransfer(grid, x.device, flags=binner.flags)
    for _ in range(repeats):
        bins = binner(x)
        f = transfer.spread(x, F, bins)
        transfer.interpolate(x, f, bins)
    src = torch.randn(n ** 3, device='cuda:0', dtype=torch.float64, generator=gen)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(repeats):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    P, N = len(x_np), n ** 3
    print(json.dumps(dict(kernels_only=True, n=n, repeats=repeats, sweep_us=sweep_s * 1e6,
                          sweep_algorithmic_bytes=12 * N, spread_algorithmic_bytes=48 * P + 24 * N,
                          interp_algorithmic_bytes=48 * P + 24 * N, copy_bytes_read=8 * N,
                          copy_bytes_written=8 * N)))


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps(dict(error='this build runs on one GPU; --gpus must be 1')))
        sys.exit(2)
    import torch
    if args.kernels_only:
        kernels_only(torch, args.n, args.kernels_only, args.points)
        return
    if args.only:
        run = dict(p2=lambda: p2_sample(torch), fe=lambda: fe_sample(torch, args.n, args.fe_level),
                   assembled=lambda: assembled_sample(torch),
                   batched=lambda: batched_sample(torch, args.batch_instances),
                   serving=lambda: serving_sample(torch, args.serve_seconds))[args.only]
        print(json.dumps({args.only: run()}))
        return
    from afsi_gpu_amd import workloads
    from afsi_gpu_amd.mac import MACGrid, MGOptions
    from afsi_gpu_amd.mac.multigrid import MIXED_INNER_CYCLES, MIXED_RELAX
    from afsi_gpu_amd.step import IBStep

    device = torch.device('cuda:0')
    # dt: with LogNormal(0, 1) point forces at this density, dt = 1e-3 moves the
    # points by many cells per step and scatters the clustered set after one
    # step; at 2e-7 they stay where the workload puts them (the displacement is
    # reported). The work of a step does not depend on dt.
    n, dt = args.n, 2e-7
    x_np, F_np = workloads.clustered_shells_128(args.points)
    P, N = len(x_np), n ** 3
    grid = MACGrid((n, n, n), (1., 1., 1.), periodic=True)
    opts = MGOptions(smooth=2, smoother='rbgs')
    relaxed = MGOptions(smooth=2, smoother='rbgs', relax=MIXED_RELAX, inner_cycles=MIXED_INNER_CYCLES)
    x0 = torch.tensor(x_np, device=device)
    F = torch.tensor(F_np, device=device)
    results = {}
    steppers = {}
    # fp64: the reference's red-black composition in FP64; mixed: the
    # production path; fp64_relaxed: FP64 with the mixed path's relaxation,
    # which separates the gain of the relaxation from that of FP32.
    for name, precision, options in (('fp64', 'fp64', opts), ('mixed', 'mixed', relaxed),
                                     ('fp64_relaxed', 'fp64', relaxed)):
        stepper = IBStep(grid, P, dt=dt, device=device, options=options, precision=precision,
                         resident_pressure=True, graphs=True)
        steppers[name] = stepper
        results[name] = run_path(torch, stepper, x0, F, args.steps, args.warmup)
    # accuracy of every path against the FP64 oracle (same inputs, one step from rest)
    states = {name: one_step(torch, steppers[key], x0, F)
              for name, key in (('plain_fp64', 'fp64'), ('mixed', 'mixed'), ('fp64_relaxed', 'fp64_relaxed'))}
    inner = steppers['mixed'].flow.pressure_solver.inner
    sweep_s = time_sweep(torch, inner)
    sweep_bytes = 3 * 4 * N
    achieved = sweep_bytes / sweep_s
    # the mixed step's largest kernel by time (one launch per step): the spread
    spread_s = time_spread(torch, steppers['mixed'], x0, F)
    spread_bytes = 48 * P + 24 * N
    mixed, plain = results['mixed'], results['fp64']
    result = dict(
        metric=METRIC, value=mixed['steps_per_s'], unit='steps/s', n_gpus=1, steps=args.steps,
        warmup=args.warmup, ms_per_step=mixed['ms_per_step'], higher_is_better=True,
        scaling='weak', vs_baseline=None, dtype='f32+f64', data='synthetic',
        config=dict(workload=f'{n}^3 periodic unit cube, {P} points (two ellipsoidal '
                             'shells + 10% in a 0.05 patch), lognormal forces; IB step (bin, spread, '
                             'projection with the pressure solved from zero by red-black GS V(2,2) to '
                             f'rtol 1e-10, interpolation, x += dt U, dt {dt:g}); mixed = FP32 V-cycles with '
                             f'red-black over-relaxation {MIXED_RELAX}, {MIXED_INNER_CYCLES} per FP64 '
                             'refinement step',
                    grid=n, points=P, precision='mixed', relax=MIXED_RELAX, inner_cycles=MIXED_INNER_CYCLES,
                    execution='every path: each step as two HIP graphs around the host convergence checks '
                              '(V-cycle blocks as graphs of their own); phases_ms from 5 further steps '
                              'as plain launches with HIP events per phase'),
        roofline=dict(bound='hbm', achieved=spread_bytes / spread_s / 1e9, peak=HBM_PEAK / 1e9, unit='GB/s',
                      frac=spread_bytes / spread_s / HBM_PEAK, traffic=measured_traffic(n, SPREAD_TRAFFIC_FILE),
                      kernel='afsi_spread_tiled_f64 (zero fill, work-list scan, spread_tiled_kernel<8>)',
                      algorithmic_bytes=spread_bytes, avg_us=spread_s * 1e6,
                      note='the largest kernel of the mixed step by time (one launch per step); '
                           'algorithmic bytes 48 per point + 24 per cell'),
        roofline_sweep=dict(bound='hbm', achieved=achieved / 1e9, peak=HBM_PEAK / 1e9, unit='GB/s',
                            frac=achieved / HBM_PEAK, traffic=measured_traffic(n),
                            kernel='rb_rows_kernel<float,true,4,4> (level-0 sweep, 24 launches per mixed step)',
                            algorithmic_bytes=sweep_bytes, avg_us=sweep_s * 1e6),
        plain_fp64=dict(steps_per_s=plain['steps_per_s'], ms_per_step=plain['ms_per_step'],
                        phases_ms=plain['phases_ms'], mg_cycles_mean=plain['mg_cycles_mean'],
                        max_displacement_cells=plain['max_displacement_cells'],
                        spread_mpts=P / plain['phases_ms']['spread'] / 1e3,
                        interp_mpts=P / plain['phases_ms']['interp'] / 1e3),
        mixed=dict(steps_per_s=mixed['steps_per_s'], ms_per_step=mixed['ms_per_step'],
                   phases_ms=mixed['phases_ms'], mg_cycles_mean=mixed['mg_cycles_mean'],
                   max_displacement_cells=mixed['max_displacement_cells'],
                   spread_mpts=P / mixed['phases_ms']['spread'] / 1e3,
                   interp_mpts=P / mixed['phases_ms']['interp'] / 1e3),
        plain_fp64_relaxed=dict(steps_per_s=results['fp64_relaxed']['steps_per_s'],
                                ms_per_step=results['fp64_relaxed']['ms_per_step'],
                                phases_ms=results['fp64_relaxed']['phases_ms'],
                                mg_cycles_mean=results['fp64_relaxed']['mg_cycles_mean'], relax=MIXED_RELAX),
        mixed_vs_plain=mixed['steps_per_s'] / plain['steps_per_s'],
        mixed_vs_plain_relaxed=mixed['steps_per_s'] / results['fp64_relaxed']['steps_per_s'],
        accuracy=oracle_accuracy(torch, x_np, F_np, n, dt, states))
    del states
    if not args.no_dropin:
        result['dropin'] = dropin_sample(torch, x0, F, n)
    if args.batch_instances:
        result['batched'] = batched_sample(torch, args.batch_instances)
    if args.serve_seconds > 0:
        result['serving'] = serving_sample(torch, args.serve_seconds)
    if args.fe_level > 0:
        result['fe'] = fe_sample(torch, n, args.fe_level)
    if not args.no_p2:
        result['p2'] = p2_sample(torch)
    if not args.no_assembled:
        result['assembled'] = assembled_sample(torch)
    if not args.no_cpu_baseline:
        result['cpu_baseline'] = cpu_baseline(x_np, F_np, n, dt)
    print(json.dumps(result))


if __name__ == '__main__':
    main()
