"""Serving: resident instances, HBM-budget admission, continuous batching
with admission, launch and completion on separate host threads.

Bars: every request is served; an instance's state after its requests is
the same alone and in a mix with other instances within 1e-10 relative
(FP64 transfers accumulate with atomics, so only their summation order may
differ; the batched pressure solve is per-instance deterministic); the
resident footprint never exceeds the budget; at the request-stream mix
(2D channels, 32^3-64^3 and 128^3 with up to 1.5M points) every request's
projected divergence ||D u|| / ||D u*|| is <= 1e-9 and the kept instances
match their single-instance steppers (IBStep mixed, ChannelIBStep), which
the oracle pins; latency percentiles are reported.
"""
import numpy as np
import pytest
import torch

from afsi_gpu_amd import workloads
from afsi_gpu_amd.mac import MACGrid, MGOptions
from afsi_gpu_amd.mac2d import ChannelGrid, ChannelIBStep
from afsi_gpu_amd.serve import Instance3D, Server, footprint
from afsi_gpu_amd.step import IBStep

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


def small_stream():
    specs = {0: ((32, 32, 32), 20000, 11), 1: ((256, 64), 4000, 12),
             2: ((48, 48, 48), 50000, 13), 3: ((128, 32), 2000, 14)}
    gen = np.random.default_rng(5)
    times = np.sort(gen.uniform(0, .5, size=40))
    inst = np.arange(40) % 4
    return times, inst, specs


def test_same_alone_and_in_the_mix(gpu):
    times, inst, specs = small_stream()
    mix = Server(device=gpu)
    stats = mix.run(times, inst, specs, keep=(0, 1))
    assert stats.summary()['requests'] == 40
    for k in (0, 1):
        alone = Server(device=gpu)
        sel = inst == k
        alone.run(times[sel], inst[sel], {k: specs[k]}, keep=(k,))
        a, b = alone.resident[k], mix.resident[k]
        assert rel(b.x, a.x) < 1e-12
        for u, v in zip(b.velocity, a.velocity):
            assert rel(u, v) < 1e-10
        assert rel(b.pressure, a.pressure) < 1e-10


def test_a_burst_is_batched(gpu):
    """Forty requests arriving at once: later iterations take several
    instances each, with the same results as alone."""
    times, inst, specs = small_stream()
    times = np.zeros_like(times)
    mix = Server(device=gpu)
    stats = mix.run(times, inst, specs, keep=(0, 2))
    assert stats.summary()['requests'] == 40
    assert max(stats.batch_sizes) >= 2
    for k in (0, 2):
        alone = Server(device=gpu)
        sel = inst == k
        alone.run(times[sel], inst[sel], {k: specs[k]}, keep=(k,))
        a, b = alone.resident[k], mix.resident[k]
        assert rel(b.x, a.x) < 1e-12
        assert rel(b.pressure, a.pressure) < 1e-10


def test_admission_respects_budget(gpu):
    times, inst, specs = small_stream()
    budget = footprint(*specs[2][:2]) + footprint(*specs[0][:2]) // 2
    srv = Server(device=gpu, budget=budget)
    stats = srv.run(times, inst, specs)
    s = stats.summary()
    assert s['requests'] == 40 and s['instances'] == 4
    assert stats.peak_bytes <= budget
    assert stats.waits_for_memory > 0
    assert not srv.resident


def test_latency_percentiles_on_a_poisson_stream(gpu):
    times, inst, specs = workloads.request_stream(3, rate=100., duration=1.,
                                                  mix=((.7, 'serve-2d'), (.3, 'serve-3d')))
    stats = Server(device=gpu).run(times, inst, specs)
    s = stats.summary()
    assert s['requests'] == len(times)
    assert 0 < s['p50_ms'] <= s['p99_ms'] < 5000
    print('serving summary:', s)


def test_snapshot_restore_continues_the_same_trajectory(gpu, tmp_path):
    """Resident instances written by Server.snapshot and made resident again
    by Server.restore (in a fresh server) step on exactly as the originals:
    the restored state is bit-identical, and one more step of each agrees
    to the transfers' atomic-order roundoff."""
    times, inst, specs = small_stream()
    first = Server(device=gpu)
    first.run(times, inst, specs, keep=(0, 1, 2, 3))
    path = tmp_path / 'serving.npz'
    first.snapshot(path)
    second = Server(device=gpu)
    assert second.restore(path) == [0, 1, 2, 3]
    for k in range(4):
        a, b = first.resident[k], second.resident[k]
        assert torch.equal(a.x, b.x) and torch.equal(a.pressure, b.pressure)
        assert all(torch.equal(u, v) for u, v in zip(a.velocity, b.velocity))
    first._execute([first.resident[k] for k in range(4)])
    second._execute([second.resident[k] for k in range(4)])
    for k in range(4):
        a, b = first.resident[k], second.resident[k]
        assert rel(b.x, a.x) < 1e-12
        for u, v in zip(b.velocity, a.velocity):
            assert rel(u, v) < 1e-10


@pytest.mark.parametrize('precision', ['mixed', 'fp64'])
def test_request_stream_mix_with_the_128_class(gpu, precision):
    """Two seconds of the request stream (60% channels, 30% 32^3-64^3, 10%
    128^3 with up to 1.5M points; ten requests per instance): every request
    served, every projection divergence-free to the solve tolerance."""
    times, inst, specs = workloads.request_stream(3, rate=200., duration=2.)
    assert any(len(sp[0]) == 3 and sp[0][0] == 128 for sp in specs.values())
    stats = Server(device=gpu, precision=precision).run(times, inst, specs)
    s = stats.summary()
    print(precision, 'stream summary:', s)
    assert s['requests'] == len(times)
    assert s['divergence_max'] <= 1e-9
    assert 0 < s['p50_ms'] <= s['p99_ms']


@pytest.mark.parametrize('precision', ['mixed', 'fp64'])
def test_served_steps_match_single_instance_steps(gpu, precision):
    """One request per instance at the stream mix: the kept instance of each
    class (channel, 32^3-64^3, 128^3) equals its single-instance stepper
    (IBStep / ChannelIBStep, pinned by the oracle) after its step within
    1e-9. One step: these forces move a 128^3 instance's points by tens of
    cells per step, so trajectories that start 1e-10 apart separate by
    about that factor per step."""
    times, inst, specs = workloads.request_stream(5, rate=200., duration=.6, per_instance=1)
    first = {}
    for iid in sorted(specs):
        shape = specs[iid][0]
        first.setdefault('2d' if len(shape) == 2 else ('128' if shape[0] >= 128 else '3d'), iid)
    keep = tuple(first.values())
    assert len(keep) == 3
    srv = Server(device=gpu, precision=precision)
    stats = srv.run(times, inst, specs, keep=keep)
    assert stats.summary()['requests'] == len(times)
    opts = srv.options  # the server's (over-relaxed red-black for the mixed path)
    for iid in keep:
        shape, count, seed = specs[iid]
        x0, F = workloads.instance_points(workloads.rng(seed), shape, count)
        x = torch.tensor(x0, device=gpu)
        f = torch.tensor(F, device=gpu)
        if len(shape) == 3:
            g = MACGrid(shape, (1., 1., 1.), periodic=True)
            single = IBStep(g, count, dt=srv.dt, device=gpu, options=opts,
                            precision='mixed' if precision == 'mixed' else 'fp64')
        else:
            g = ChannelGrid(shape, (4., 1.), periodic=True)
            single = ChannelIBStep(g, dt=srv.dt, device=gpu, options=opts, pressure_backend='workspace')
        v, _, _, _ = single.step(x, f, g.zeros(device=gpu), None)
        got = srv.resident[iid]
        x0t = torch.tensor(x0, device=gpu)
        dx = rel(got.x - x0t, x - x0t)
        du = max(rel(u, w) for u, w in zip(got.velocity, v))
        print(precision, shape, count, 'displacement', dx, 'velocity', du)
        assert dx < 1e-9 and du < 1e-9, (shape, count)
