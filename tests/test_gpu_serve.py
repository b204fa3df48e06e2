"""Serving: resident instances, HBM-budget admission, continuous batching.

Bars: every request is served; an instance's state after its requests is
the same alone and in a mix with other instances within 1e-10 relative
(FP64 transfers accumulate with atomics, so only their summation order may
differ; the batched pressure solve is per-instance deterministic); the
resident footprint never exceeds the budget; latency percentiles are
reported.
"""
import numpy as np
import pytest
import torch

from afsi_gpu_amd import workloads
from afsi_gpu_amd.serve import Instance3D, Server, footprint

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
    assert max(stats.batch_sizes) >= 2


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
