"""Serving spread-project-interpolate steps as requests arrive (build-own).

Each request advances one instance by one IB step. Instance state (grid
fields, multigrid hierarchy, points, forces, velocity, pressure) stays
resident in HBM across the instance's requests and is released after its
last one. Admission is against an HBM budget: an instance's first request
waits until its estimated footprint fits.

Three host threads keep the device busy:
* the arrival thread (the caller's) queues requests as they arrive and
  admits new instances: allocation, setup and upload run on a stream of
  their own, so they overlap the steps in flight; an instance becomes
  eligible once its upload event has completed;
* the executor takes every eligible request (at most one per instance, in
  arrival order) and launches them as one iteration: the 3D steps share
  one batched pressure solve (mixed precision: one launch per multigrid
  phase over all of them; plain FP64: one solve per instance), the 2D
  channel steps run on one stepper per channel shape. It does not wait for
  the iteration to finish: it records an event and takes the next batch
  (an instance's next step queues behind its previous one on the stream);
* the completer polls the events of iterations in flight and stamps their
  requests complete (latency = completion - arrival), reads each request's
  divergence ratio ||D u|| / ||D u*|| and releases finished instances.
The solves still read their convergence flags on the host, so the executor
itself waits inside a solve; everything else overlaps.
"""
from collections import deque
from dataclasses import dataclass, field
import threading
import time

import numpy as np
import torch

from . import _abi, workloads
from .batch import BatchedMixedMultigrid, MixedState
from .mac import MACGrid, MGOptions
from .mac.flow import correct, divergence
from .mac.multigrid import MIXED_RELAX, GeometricMultigrid
from .mac.transfer import DEFAULT_TILE, Binner, TiledTransfer
from .mac2d import ChannelGrid, ChannelIBStep
from .mac2d.flow import divergence as divergence2d

HBM_BUDGET = 288 * 2 ** 30


def footprint(shape, count):
    """Estimated resident bytes of an instance (fields, hierarchy, points)."""
    cells = int(np.prod(shape))
    if len(shape) == 3:
        fields = 3 * cells * 8 * 2 + cells * 8           # velocity, force density, rhs
        mg = int(cells * 8 / 7) * 4 * 3 + 3 * cells * 8  # FP32 p/rhs/tmp, FP64 x[2] and b
        points = count * 8 * 3 * 3 + count * 12           # x, force, U; keys and perm
        bins = (cells + 1) * 4
        return fields + mg + points + bins
    fields = 2 * cells * 8 * 3 + cells * 8
    mg = int(cells * 4 / 3) * 8 * 3
    points = count * 8 * 2 * 3 + count * 16 * 16 * 2     # x, force, U; ids and weights
    return fields + mg + points


class Instance3D:
    """A resident 3D instance. ``solver`` is its own ``MixedState`` (mixed)
    or a plain FP64 ``GeometricMultigrid`` shared by the instances of its
    shape (fp64); the tiled transfer is shared per shape as well."""

    def __init__(self, shape, x, force, *, device, dt, options, precision='mixed', shared=None,
                 tile=DEFAULT_TILE):
        self.grid = MACGrid(tuple(shape), (1., 1., 1.), periodic=True)
        self.dt = dt
        self.precision = precision
        self.x = torch.tensor(x, device=device)
        self.force = torch.tensor(force, device=device)
        self.velocity = list(self.grid.zeros(device=device))
        self.pressure = None
        shared = shared if shared is not None else {}
        if precision == 'mixed':
            self.state = MixedState(self.grid, device=device, options=options)
            self.plain = None
        else:
            self.state = None
            self.plain = shared.get('plain') or GeometricMultigrid(self.grid, device=device, options=options,
                                                                   backend='workspace')
            shared['plain'] = self.plain
        self.transfer = shared.get('transfer') or TiledTransfer(self.grid, device, tile)
        shared['transfer'] = self.transfer
        self.binner = Binner(self.grid, len(x), device, tile)
        self.density = self.grid.zeros(device=device)
        self.rhs = torch.empty(self.grid.shape, device=device, dtype=torch.float64)
        self.div = torch.empty(self.grid.shape, device=device, dtype=torch.float64)
        self.U = torch.empty_like(self.x)
        self.bins = None
        self.ready = None

    def before_solve(self):
        self.bins = self.binner(self.x)
        f = self.transfer.spread(self.x, self.force, self.bins, out=self.density)
        for uc, fc in zip(self.velocity, f):
            _abi.call('afsi_axpy_f64', _abi.ptr(uc), _abi.ptr(fc), uc.numel(), self.dt, _abi.stream())
        return divergence(self.velocity, self.grid, -1 / self.dt, out=self.rhs)

    def after_solve(self, p):
        """Correction, interpolation and advance; returns the device scalar
        ||D u|| / ||D u*|| of this step."""
        correct(self.velocity, p, self.grid, self.dt)
        self.pressure = p
        U = self.transfer.interpolate(self.x, self.velocity, self.bins, out=self.U)
        _abi.call('afsi_axpy_f64', _abi.ptr(self.x), _abi.ptr(U), self.x.numel(), self.dt, _abi.stream())
        after = torch.linalg.vector_norm(divergence(self.velocity, self.grid, out=self.div))
        return after / (torch.linalg.vector_norm(self.rhs) * self.dt)


class Instance2D:
    """A resident 2D channel; the stepper (fields, multigrid workspace) is
    shared by the instances of its shape, the state passed per step."""

    def __init__(self, shape, x, force, *, device, dt, options, shared=None):
        self.grid = ChannelGrid(tuple(shape), (4., 1.), periodic=True)
        shared = shared if shared is not None else {}
        self.stepper = shared.get('channel') or ChannelIBStep(self.grid, dt=dt, device=device, options=options,
                                                              pressure_backend='workspace')
        shared['channel'] = self.stepper
        self.dt = dt
        self.x = torch.tensor(x, device=device)
        self.force = torch.tensor(force, device=device)
        self.velocity = self.grid.zeros(device=device)
        self.pressure = None
        self.div = torch.empty(self.grid.shape, device=device, dtype=torch.float64)
        self.ready = None

    def step(self):
        """One step; returns the device scalar ||D u|| / ||D u*||."""
        self.velocity, self.pressure, _, _ = self.stepper.step(self.x, self.force, self.velocity,
                                                               self.pressure)
        before = torch.linalg.vector_norm(self.stepper.flow._rhs) * self.dt
        after = torch.linalg.vector_norm(divergence2d(self.velocity, self.grid, out=self.div))
        return after / before


@dataclass
class Stats:
    latencies: list = field(default_factory=list)
    iterations: int = 0
    batch_sizes: list = field(default_factory=list)
    peak_bytes: int = 0
    admitted: int = 0
    waits_for_memory: int = 0
    wall: float = 0.
    admit_seconds: list = field(default_factory=list)  # instance creation (allocation, setup, upload)
    launch_seconds: list = field(default_factory=list)  # host time to launch one iteration
    waits: list = field(default_factory=list)  # arrival to the launch of the request's iteration
    divergence: list = field(default_factory=list)  # ||D u|| / ||D u*|| per request
    kinds: list = field(default_factory=list)  # '2d' / '3d' / '128' per request

    def summary(self):
        lat = np.asarray(self.latencies) * 1e3
        div = np.asarray(self.divergence)
        out = dict(requests=len(lat), p50_ms=float(np.percentile(lat, 50)) if len(lat) else None,
                   p99_ms=float(np.percentile(lat, 99)) if len(lat) else None,
                   mean_ms=float(lat.mean()) if len(lat) else None,
                   throughput_rps=len(lat) / self.wall if self.wall else None,
                   iterations=self.iterations,
                   mean_batch=float(np.mean(self.batch_sizes)) if self.batch_sizes else 0.,
                   max_batch=int(max(self.batch_sizes)) if self.batch_sizes else 0,
                   instances=self.admitted, peak_resident_gib=self.peak_bytes / 2 ** 30,
                   memory_waits=self.waits_for_memory, wall_s=self.wall,
                   admit_ms_mean=1e3 * float(np.mean(self.admit_seconds)) if self.admit_seconds else 0.,
                   admit_ms_max=1e3 * float(max(self.admit_seconds)) if self.admit_seconds else 0.,
                   launch_ms_p50=1e3 * float(np.percentile(self.launch_seconds, 50)) if self.launch_seconds else 0.,
                   launch_ms_p99=1e3 * float(np.percentile(self.launch_seconds, 99)) if self.launch_seconds else 0.,
                   wait_ms_p99=1e3 * float(np.percentile(self.waits, 99)) if self.waits else 0.,
                   divergence_max=float(div.max()) if len(div) else None,
                   divergence_p99=float(np.percentile(div, 99)) if len(div) else None)
        if len(lat) and self.kinds:
            kinds = np.asarray(self.kinds)
            out['p99_ms_by_kind'] = {k: float(np.percentile(lat[kinds == k], 99)) for k in np.unique(kinds)}
        return out


def _kind(shape):
    return '2d' if len(shape) == 2 else ('128' if shape[0] >= 128 else '3d')


class Server:
    """Runs a request stream (arrival times, instance ids, instance specs)
    against the GPU in real time. ``precision``: 'mixed' (batched FP32
    V-cycles in FP64 refinement for the 3D instances) or 'fp64' (plain FP64
    multigrid per 3D instance)."""

    def __init__(self, *, device='cuda', dt=1e-3, budget=HBM_BUDGET, options=None,
                 max_batch=64, precision='mixed'):
        if precision not in ('mixed', 'fp64'):
            raise ValueError("precision must be 'mixed' or 'fp64'")
        self.device = torch.device(device)
        self.dt = dt
        self.budget = budget
        self.options = options or MGOptions(smooth=2, smoother='rbgs',
                                            relax=MIXED_RELAX if precision == 'mixed' else 1.)
        self.max_batch = max_batch
        self.precision = precision
        self.resident = {}
        self.resident_bytes = {}
        self.shared = {}          # per shape: steppers, plain solvers, transfers
        self.admit_stream = torch.cuda.Stream(device=self.device)

    def _create(self, spec, data):
        shape = tuple(spec[0])
        x, F = data
        shared = self.shared.setdefault(shape, {})
        if len(shape) == 3:
            return Instance3D(shape, x, F, device=self.device, dt=self.dt, options=self.options,
                              precision=self.precision, shared=shared)
        return Instance2D(shape, x, F, device=self.device, dt=self.dt, options=self.options, shared=shared)

    def snapshot(self, path):
        """Write every resident instance's state (points, forces, face
        velocities, pressure) to one checkpoint."""
        from . import checkpoint
        torch.cuda.synchronize(self.device)
        instances = {key: dict(kind='3d' if isinstance(inst, Instance3D) else '2d', shape=inst.grid.shape,
                               x=inst.x, force=inst.force, velocity=tuple(inst.velocity), pressure=inst.pressure)
                     for key, inst in self.resident.items()}
        checkpoint.save_resident(path, instances, dict(dt=self.dt))

    def restore(self, path):
        """Make the instances of a snapshot resident again, with their state;
        returns their ids."""
        from . import checkpoint
        instances, settings = checkpoint.load_resident(path)
        if abs(settings['dt'] - self.dt) > 1e-15:
            raise ValueError('snapshot time step differs from the server\'s')
        for key, saved in instances.items():
            inst = self._create((saved['shape'],), (saved['x'].numpy(), saved['force'].numpy()))
            for target, value in zip(inst.velocity, saved['velocity']):
                if target.shape != value.shape:
                    raise ValueError('invalid serving checkpoint index')
                target.copy_(value)
            inst.pressure = None if saved['pressure'] is None else saved['pressure'].to(self.device)
            self.resident[key] = inst
            self.resident_bytes[key] = footprint(saved['shape'], len(saved['x']))
        return sorted(instances)

    def _prepare_shapes(self, shapes):
        """Per-shape state shared by the instances of a shape, built before
        serving starts (like loading a model): the channel stepper and, in
        FP64, the 3D pressure solver with their V-cycle blocks captured as
        HIP graphs once per shape (capturing while other threads launch work
        would not be safe), and the 3D tiled transfer."""
        for shape in shapes:
            shared = self.shared.setdefault(shape, {})
            if len(shape) == 2:
                if 'channel' not in shared:
                    grid = ChannelGrid(shape, (4., 1.), periodic=True)
                    shared['channel'] = ChannelIBStep(grid, dt=self.dt, device=self.device, options=self.options,
                                                      pressure_backend='graph')
            else:
                grid = MACGrid(shape, (1., 1., 1.), periodic=True)
                if self.precision == 'fp64' and 'plain' not in shared:
                    shared['plain'] = GeometricMultigrid(grid, device=self.device, options=self.options,
                                                         backend='graph')
                if 'transfer' not in shared:
                    shared['transfer'] = TiledTransfer(grid, self.device, DEFAULT_TILE)
        torch.cuda.synchronize(self.device)

    def _execute(self, batch):
        """Launch one iteration on the current stream; returns the (n,)
        device tensor of the requests' divergence ratios (no host wait
        beyond the solves' own convergence checks)."""
        ratios = [None] * len(batch)
        three = [(j, inst) for j, inst in enumerate(batch) if isinstance(inst, Instance3D)]
        rhs = [inst.before_solve() for _, inst in three]
        for j, inst in enumerate(batch):
            if isinstance(inst, Instance2D):
                ratios[j] = inst.step()
        if three:
            if self.precision == 'mixed':
                solver = BatchedMixedMultigrid(states=[inst.state for _, inst in three],
                                               options=self.options, device=self.device, graphs=False)
                ps, _ = solver.solve(rhs, [inst.pressure for _, inst in three])
            else:
                ps = [inst.plain.solve(r, inst.pressure)[0] for (_, inst), r in zip(three, rhs)]
            for (j, inst), p in zip(three, ps):
                ratios[j] = inst.after_solve(p)
        return torch.stack(ratios)

    def run(self, times, inst_ids, specs, *, keep=(), clock=time.perf_counter):
        """Serve every request; returns Stats. Instances in ``keep`` stay
        resident after their last request (for inspection). The instances'
        synthetic points are generated before the clock starts (they stand
        for the client's upload); admission copies them to the device."""
        stats = Stats()
        wanted = set(int(i) for i in inst_ids)
        data = {iid: workloads.instance_points(workloads.rng(sp[2]), sp[0], sp[1])
                for iid, sp in specs.items() if iid in wanted}
        self._prepare_shapes(sorted({tuple(sp[0]) for iid, sp in specs.items() if iid in wanted}))
        order = np.argsort(times, kind='stable')
        times, inst_ids = np.asarray(times)[order], np.asarray(inst_ids)[order]
        remaining = {int(i): int(c) for i, c in zip(*np.unique(inst_ids, return_counts=True))}
        lock = threading.Lock()
        eligible = deque()      # (arrival, instance id), instance resident
        inflight = []           # (done event, [(arrival, iid, launch time)], ratios)
        shared = dict(used=0, arrived=0, launched=0, error=None)
        total = len(times)
        t0 = clock()
        # Channel and 3D requests run on their own streams from their own
        # executor threads, so a channel step never queues behind a 128^3
        # solve; both streams have the default priority.
        streams = dict(two=torch.cuda.current_stream(self.device), three=torch.cuda.Stream(device=self.device))
        totals = dict(two=sum(1 for i in inst_ids if len(specs[int(i)][0]) == 2))
        totals['three'] = total - totals['two']

        def fail(exc):
            with lock:
                if shared['error'] is None:
                    shared['error'] = exc

        def executor(kind):
            launched = 0
            stream = streams[kind]
            try:
                with torch.cuda.stream(stream):
                    while True:
                        with lock:
                            if shared['error'] is not None:
                                return
                            if launched == totals[kind]:
                                return
                            batch, reqs, busy, keep_q = [], [], set(), deque()
                            while eligible:
                                arr, iid = eligible.popleft()
                                inst = self.resident[iid]
                                mine = isinstance(inst, Instance2D) == (kind == 'two')
                                ok = mine and iid not in busy and len(batch) < self.max_batch and (
                                    inst.ready is None or inst.ready.query())
                                if ok:
                                    busy.add(iid)
                                    batch.append(inst)
                                    reqs.append((arr, iid))
                                else:
                                    keep_q.append((arr, iid))
                            eligible.extendleft(reversed(keep_q))
                        if not batch:
                            time.sleep(0.0002)
                            continue
                        start = clock() - t0
                        ratios = self._execute(batch)
                        host = torch.empty(ratios.shape, dtype=ratios.dtype, pin_memory=True)
                        host.copy_(ratios, non_blocking=True)
                        done = torch.cuda.Event()
                        done.record(stream)
                        stats.launch_seconds.append(clock() - t0 - start)
                        stats.batch_sizes.append(len(batch))
                        launched += len(batch)
                        with lock:
                            stats.iterations += 1
                            inflight.append((done, [(a, i, start) for a, i in reqs], host))
                            shared['launched'] += len(batch)
            except BaseException as exc:  # surfaced by run()
                fail(exc)

        def completer():
            try:
                finished = 0
                while finished < total:
                    with lock:
                        if shared['error'] is not None:
                            return
                        pending = list(inflight)
                    head = next((entry for entry in pending if entry[0].query()), None)
                    if head is None:
                        time.sleep(0.0002)
                        continue
                    now = clock() - t0
                    done, reqs, host = head
                    values = host.tolist()
                    with lock:
                        inflight.remove(head)
                        for (arr, iid, start), ratio in zip(reqs, values):
                            stats.latencies.append(now - arr)
                            stats.waits.append(start - arr)
                            stats.divergence.append(ratio)
                            stats.kinds.append(_kind(specs[iid][0]))
                            remaining[iid] -= 1
                            if remaining[iid] == 0 and iid not in keep:
                                del self.resident[iid]
                                shared['used'] -= self.resident_bytes.pop(iid)
                        finished += len(reqs)
            except BaseException as exc:
                fail(exc)

        threads = [threading.Thread(target=executor, args=(kind,), daemon=True) for kind in ('two', 'three')
                   if totals[kind]]
        threads.append(threading.Thread(target=completer, daemon=True))
        for th in threads:
            th.start()
        waiting = deque()       # arrived, instance not yet admitted (budget)
        nxt = 0
        try:
            while nxt < total or waiting:
                with lock:
                    if shared['error'] is not None:
                        break
                now = clock() - t0
                while nxt < total and times[nxt] <= now:
                    waiting.append((float(times[nxt]), int(inst_ids[nxt])))
                    nxt += 1
                admitted_any = False
                held = deque()
                while waiting:
                    arr, iid = waiting.popleft()
                    with lock:
                        resident = iid in self.resident
                        used = shared['used']
                    if not resident:
                        need = footprint(specs[iid][0], specs[iid][1])
                        if used + need > self.budget:
                            stats.waits_for_memory += 1
                            held.append((arr, iid))
                            continue
                        a0 = clock()
                        with torch.cuda.stream(self.admit_stream):
                            inst = self._create(specs[iid], data.pop(iid))
                            inst.ready = torch.cuda.Event()
                            inst.ready.record(self.admit_stream)
                        stats.admit_seconds.append(clock() - a0)
                        with lock:
                            self.resident[iid] = inst
                            self.resident_bytes[iid] = need
                            shared['used'] += need
                            stats.peak_bytes = max(stats.peak_bytes, shared['used'])
                        stats.admitted += 1
                        admitted_any = True
                    with lock:
                        eligible.append((arr, iid))
                waiting.extend(held)
                if not admitted_any:
                    wake = times[nxt] - (clock() - t0) if nxt < total else 0.0005
                    time.sleep(max(0., min(wake, 0.0005)))
        finally:
            for th in threads:
                th.join()
        stats.wall = clock() - t0
        if shared['error'] is not None:
            raise shared['error']
        torch.cuda.synchronize(self.device)
        return stats
