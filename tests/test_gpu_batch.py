"""Batched mixed-precision pressure solve and batched IB steps.

Bars: every instance of a batch of uneven grids meets its FP64 tolerance
and agrees with the single-instance mixed solver within 1e-10 relative
(same V-cycle arithmetic; only reduction order differs) with the same
iteration count; an instance's result is bit-identical alone and in the
mix (per-instance work and reductions do not depend on the other
instances); graph replay equals eager launches; the batched IB step's
point velocities agree with per-instance IB steps within 1e-9 relative.
"""
import numpy as np
import pytest
import torch

from afsi_gpu_amd import workloads
from afsi_gpu_amd.batch import BatchedIBStep, BatchedMixedMultigrid
from afsi_gpu_amd.mac import MACGrid, MGOptions, MixedMultigrid
from afsi_gpu_amd.step import IBStep

pytestmark = pytest.mark.gpu

SHAPES = [(16, 16, 16), (32, 32, 32), (24, 24, 24), (48, 48, 48), (64, 32, 48)]


def rhs_for(shape, seed, gpu):
    gen = np.random.default_rng(seed)
    r = gen.normal(size=shape)
    r -= r.mean()
    return torch.tensor(r, device=gpu)


def rel(a, b):
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


@pytest.mark.parametrize('relax,inner', [(1., 1), (1.3, 3)])
@pytest.mark.parametrize('periodic', [True, False])
def test_batched_solve_matches_single_instance(gpu, periodic, relax, inner):
    """Also with the mixed path's settings: over-relaxation 1.3 and three
    FP32 V-cycles per refinement iteration (later cycles of an iteration
    continue from the correction so far on level 0)."""
    opts = MGOptions(smooth=2, smoother='rbgs', relax=relax, inner_cycles=inner)
    grids = [MACGrid(s, (1., 1.5, 2.), periodic=periodic) for s in SHAPES]
    rhs = [rhs_for(g.shape, i, gpu) for i, g in enumerate(grids)]
    batch = BatchedMixedMultigrid(grids, device=gpu, options=opts, graphs=False)
    ps, infos = batch.solve(rhs)
    for g, r, p, info in zip(grids, rhs, ps, infos):
        single = MixedMultigrid(g, device=gpu, options=opts, backend='workspace')
        q, sinfo = single.solve(r)
        assert info['residual_norm'] <= info['tolerance']
        assert info['cycles'] == sinfo['cycles'], (g.shape, info, sinfo)
        assert rel(p, q) < 1e-10, g.shape
        assert abs(float(p.mean())) < 1e-12 * float(p.abs().max())


def test_alone_and_in_the_mix_are_identical(gpu):
    opts = MGOptions(smooth=2, smoother='rbgs')
    grids = [MACGrid(s, (1., 1., 1.), periodic=True) for s in SHAPES]
    rhs = [rhs_for(g.shape, 10 + i, gpu) for i, g in enumerate(grids)]
    mix = BatchedMixedMultigrid(grids, device=gpu, options=opts)
    ps, infos = mix.solve(rhs)
    for k in (1, 3):
        alone = BatchedMixedMultigrid([grids[k]], device=gpu, options=opts)
        (q,), (qi,) = alone.solve([rhs[k]])
        assert torch.equal(q, ps[k])
        assert qi['cycles'] == infos[k]['cycles']
    # graph replay equals eager launches
    eager = BatchedMixedMultigrid(grids, device=gpu, options=opts, graphs=False)
    es, _ = eager.solve(rhs)
    for a, b in zip(es, ps):
        assert torch.equal(a, b)
    # a second solve on the same objects (warm start) stays consistent
    ps2, infos2 = mix.solve(rhs, ps)
    assert all(i['cycles'] <= 2 for i in infos2)
    for a, b in zip(ps2, ps):
        assert rel(a, b) < 1e-9


def test_batched_solve_errors(gpu):
    grids = [MACGrid((16, 16, 16), (1., 1., 1.), periodic=True),
             MACGrid((32, 32, 32), (1., 1., 1.), periodic=True)]
    batch = BatchedMixedMultigrid(grids, device=gpu, graphs=False)
    good = rhs_for((16, 16, 16), 1, gpu)
    with pytest.raises(ValueError, match='incompatible'):
        batch.solve([good, torch.ones((32, 32, 32), device=gpu, dtype=torch.float64)])
    bad = good.clone()
    bad[1, 2, 3] = float('nan')
    with pytest.raises(ValueError, match='invalid pressure RHS'):
        batch.solve([bad, rhs_for((32, 32, 32), 2, gpu)])
    with pytest.raises(ValueError, match='one right-hand side'):
        batch.solve([good])


def test_batched_ib_step_matches_per_instance_steps(gpu):
    gen = np.random.default_rng(21)
    specs = [((32, 32, 32), 20000), ((48, 48, 48), 60000), ((64, 64, 64), 100000)]
    grids = [MACGrid(s, (1., 1., 1.), periodic=True) for s, _ in specs]
    pts, forces = [], []
    for (s, count) in specs:
        x = workloads.shell(gen, (.3, .2, .2), count)
        pts.append(x)
        forces.append(gen.normal(size=x.shape))
    opts = MGOptions(smooth=2, smoother='rbgs')
    batch = BatchedIBStep(grids, [len(x) for x in pts], dt=1e-3, device=gpu, options=opts)
    xs = [torch.tensor(x, device=gpu) for x in pts]
    fs = [torch.tensor(f, device=gpu) for f in forces]
    vels = [g.zeros(device=gpu) for g in grids]
    vels, ps, Us, infos = batch.step(xs, fs, vels)
    for i, g in enumerate(grids):
        single = IBStep(g, len(pts[i]), dt=1e-3, device=gpu, options=opts, precision='mixed')
        x1 = torch.tensor(pts[i], device=gpu)
        v1 = g.zeros(device=gpu)
        v1, p1, U1, info1 = single.step(x1, fs[i], v1)
        assert infos[i]['cycles'] == info1['pressure']['cycles']
        assert rel(Us[i], U1) < 1e-9
        assert rel(xs[i], x1) < 1e-14


@pytest.mark.parametrize('periodic', [True, False])
def test_batched_transfer_matches_per_instance(gpu, periodic):
    from afsi_gpu_amd.batch import BatchedTransfer
    from afsi_gpu_amd.mac.transfer import TiledTransfer, bin_points
    gen = np.random.default_rng(31)
    specs = [((32, 32, 32), 15000), ((48, 40, 56), 40000), ((16, 16, 16), 3000), ((64, 64, 64), 80000)]
    grids = [MACGrid(s, (1., 1., 1.), periodic=periodic) for s, _ in specs]
    xs = [torch.tensor(workloads.shell(gen, (.3, .2, .2), c), device=gpu) for _, c in specs]
    fs = [torch.randn((c, 3), device=gpu, dtype=torch.float64) for _, c in specs]
    faces = [g.zeros(device=gpu) for g in grids]
    outs = [torch.empty((c, 3), device=gpu, dtype=torch.float64) for _, c in specs]
    bt = BatchedTransfer(grids, [c for _, c in specs], device=gpu).bind(xs, fs, faces, outs)
    bt.bin()
    bt.spread()
    vel = [tuple(torch.randn(g.face_shape(c), device=gpu, dtype=torch.float64) for c in range(3))
           for g in grids]
    gt = BatchedTransfer(grids, [c for _, c in specs], device=gpu).bind(xs, fs, vel, outs)
    gt.share_bins(bt)
    gt.interpolate()
    assert int(bt.flags.item()) == 0
    cell = bt.cell_ptr.cpu().numpy()
    perm = bt.perm.cpu().numpy()
    pbeg, kbeg = 0, 0
    for i, (g, x) in enumerate(zip(grids, xs)):
        bins = bin_points(x, g)
        ks = bins.cell_ptr.numel() - 1
        seg = cell[kbeg:kbeg + ks + 1]
        assert np.array_equal(seg - seg[0], bins.cell_ptr.cpu().numpy())
        assert np.array_equal(perm[seg[0]:seg[-1]] - pbeg, bins.perm.cpu().numpy())
        tr = TiledTransfer(g, gpu)
        f = tr.spread(x, fs[i], bins)
        for a, b in zip(faces[i], f):
            assert rel(a, b) < 1e-12
        U = tr.interpolate(x, vel[i], bins)
        assert rel(outs[i], U) < 1e-12
        pbeg += x.shape[0]
        kbeg += ks
