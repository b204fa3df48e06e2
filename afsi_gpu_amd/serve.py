"""Serving spread-project-interpolate steps as requests arrive (build-own).

Each request advances one instance by one IB step. Instance state (grid
fields, multigrid hierarchy, points, forces, velocity, pressure) stays
resident in HBM across the instance's requests and is released after its
last one. Admission is against an HBM budget: an instance's first request
waits until its estimated footprint fits.

Continuous batching at step granularity: every GPU iteration takes all
requests that have arrived and whose instance is resident and idle (at
most one per instance, in arrival order) and runs them together; requests
that arrive meanwhile join the next iteration. The 3D requests of an
iteration share one batched mixed-precision pressure solve (one launch per
multigrid phase over all of them); their transfers and the 2D channel
steps are issued on the same stream. Latency = completion - arrival.
"""
from dataclasses import dataclass, field
import time

import numpy as np
import torch

from . import _abi, workloads
from .batch import BatchedMixedMultigrid, MixedState
from .mac import MACGrid, MGOptions
from .mac.flow import correct, divergence
from .mac.transfer import DEFAULT_TILE, Binner, TiledTransfer
from .mac2d import ChannelGrid, ChannelIBStep

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
    def __init__(self, shape, x, force, *, device, dt, options, tile=DEFAULT_TILE):
        self.grid = MACGrid(tuple(shape), (1., 1., 1.), periodic=True)
        self.dt = dt
        self.x = torch.tensor(x, device=device)
        self.force = torch.tensor(force, device=device)
        self.velocity = list(self.grid.zeros(device=device))
        self.pressure = None
        self.state = MixedState(self.grid, device=device, options=options)
        self.binner = Binner(self.grid, len(x), device, tile)
        self.transfer = TiledTransfer(self.grid, device, tile)
        self.density = self.grid.zeros(device=device)
        self.rhs = torch.empty(self.grid.shape, device=device, dtype=torch.float64)
        self.U = torch.empty_like(self.x)
        self.bins = None

    def before_solve(self):
        self.bins = self.binner(self.x)
        f = self.transfer.spread(self.x, self.force, self.bins, out=self.density)
        for uc, fc in zip(self.velocity, f):
            _abi.call('afsi_axpy_f64', _abi.ptr(uc), _abi.ptr(fc), uc.numel(), self.dt, _abi.stream())
        return divergence(self.velocity, self.grid, -1 / self.dt, out=self.rhs)

    def after_solve(self, p):
        correct(self.velocity, p, self.grid, self.dt)
        self.pressure = p
        U = self.transfer.interpolate(self.x, self.velocity, self.bins, out=self.U)
        _abi.call('afsi_axpy_f64', _abi.ptr(self.x), _abi.ptr(U), self.x.numel(), self.dt, _abi.stream())


class Instance2D:
    def __init__(self, shape, x, force, *, device, dt, options):
        self.grid = ChannelGrid(tuple(shape), (4., 1.), periodic=True)
        self.stepper = ChannelIBStep(self.grid, dt=dt, device=device, options=options,
                                     pressure_backend='workspace')
        self.x = torch.tensor(x, device=device)
        self.force = torch.tensor(force, device=device)
        self.velocity = self.grid.zeros(device=device)
        self.pressure = None

    def step(self):
        self.velocity, self.pressure, _, _ = self.stepper.step(self.x, self.force, self.velocity,
                                                               self.pressure)


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
    execute_seconds: list = field(default_factory=list)  # one batched GPU iteration
    waits: list = field(default_factory=list)  # arrival to the start of the request's iteration

    def summary(self):
        lat = np.asarray(self.latencies) * 1e3
        return dict(requests=len(lat), p50_ms=float(np.percentile(lat, 50)) if len(lat) else None,
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
                    execute_ms_p50=1e3 * float(np.percentile(self.execute_seconds, 50)) if self.execute_seconds else 0.,
                    execute_ms_p99=1e3 * float(np.percentile(self.execute_seconds, 99)) if self.execute_seconds else 0.,
                    wait_ms_p99=1e3 * float(np.percentile(self.waits, 99)) if self.waits else 0.)


class Server:
    """Runs a request stream (arrival times, instance ids, instance specs)
    against the GPU in real time."""

    def __init__(self, *, device='cuda', dt=1e-3, budget=HBM_BUDGET, options=None,
                 max_batch=64):
        self.device = torch.device(device)
        self.dt = dt
        self.budget = budget
        self.options = options or MGOptions(smooth=2, smoother='rbgs')
        self.max_batch = max_batch
        self.resident = {}
        self.resident_bytes = {}

    def _create(self, spec, data):
        shape = spec[0]
        x, F = data
        if len(shape) == 3:
            return Instance3D(shape, x, F, device=self.device, dt=self.dt, options=self.options)
        return Instance2D(shape, x, F, device=self.device, dt=self.dt, options=self.options)

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

    def _execute(self, batch):
        three = [inst for inst in batch if isinstance(inst, Instance3D)]
        rhs = [inst.before_solve() for inst in three]
        for inst in batch:
            if isinstance(inst, Instance2D):
                inst.step()
        if three:
            solver = BatchedMixedMultigrid(states=[inst.state for inst in three],
                                           options=self.options, device=self.device, graphs=False)
            ps, _ = solver.solve(rhs, [inst.pressure for inst in three])
            for inst, p in zip(three, ps):
                inst.after_solve(p)
        torch.cuda.synchronize(self.device)

    def run(self, times, inst_ids, specs, *, keep=(), clock=time.perf_counter):
        """Serve every request; returns Stats. Instances in ``keep`` stay
        resident after their last request (for inspection). The instances'
        synthetic points are generated before the clock starts (they stand
        for the client's upload); admission copies them to the device."""
        stats = Stats()
        wanted = set(int(i) for i in inst_ids)
        data = {iid: workloads.instance_points(workloads.rng(sp[2]), sp[0], sp[1])
                for iid, sp in specs.items() if iid in wanted}
        order = np.argsort(times, kind='stable')
        times, inst_ids = np.asarray(times)[order], np.asarray(inst_ids)[order]
        remaining = {int(i): int(c) for i, c in zip(*np.unique(inst_ids, return_counts=True))}
        pending = []            # (arrival, instance id)
        nxt = 0
        used = 0
        t0 = clock()
        while nxt < len(times) or pending:
            now = clock() - t0
            while nxt < len(times) and times[nxt] <= now:
                pending.append((float(times[nxt]), int(inst_ids[nxt])))
                nxt += 1
            if not pending:
                time.sleep(max(0., min(times[nxt] - now, 0.002)))
                continue
            batch, chosen, busy = [], [], set()
            for j, (arr, iid) in enumerate(pending):
                if iid in busy or len(batch) >= self.max_batch:
                    continue
                if iid not in self.resident:
                    need = footprint(specs[iid][0], specs[iid][1])
                    if used + need > self.budget:
                        stats.waits_for_memory += 1
                        busy.add(iid)
                        continue
                    a0 = clock()
                    self.resident[iid] = self._create(specs[iid], data.pop(iid))
                    stats.admit_seconds.append(clock() - a0)
                    self.resident_bytes[iid] = need
                    used += need
                    stats.admitted += 1
                    stats.peak_bytes = max(stats.peak_bytes, used)
                busy.add(iid)
                batch.append(self.resident[iid])
                chosen.append(j)
            if not batch:
                time.sleep(0.0005)
                continue
            start = clock() - t0
            self._execute(batch)
            done = clock() - t0
            stats.execute_seconds.append(done - start)
            stats.iterations += 1
            stats.batch_sizes.append(len(batch))
            for j in chosen:
                arr, iid = pending[j]
                stats.latencies.append(done - arr)
                stats.waits.append(start - arr)
                remaining[iid] -= 1
                if remaining[iid] == 0 and iid not in keep:
                    del self.resident[iid]
                    used -= self.resident_bytes.pop(iid)
            for j in sorted(chosen, reverse=True):
                pending.pop(j)
        stats.wall = clock() - t0
        return stats
