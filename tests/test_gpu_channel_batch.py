"""Batched channel instances (mac2d/batch.py) against the oracle: each
instance of an uneven batch (128x32, 256x64 and 512x128 periodic-x channels
and two reference channels with inlet/outlet, 64x32 and 128x64) advanced two
steps with one launch per phase, against oracle/step.ib_step_2d on the same
inputs: the solve's cycle count equal, projected velocity and interpolated
point velocity within 1e-10 relative, positions within 1e-14. The levels
above 4096 cells of the 256x64, 512x128 and 128x64 instances run as batched
launches around a one-workgroup tail; the other instances cycle in one
workgroup each."""
from ctypes import byref

import numpy as np
import pytest
import torch

from oracle import mac2d, step as ostep
from oracle.mac3d import MGOptions as OOptions

from afsi_gpu_amd import _abi, workloads
from afsi_gpu_amd.mac import MGOptions, cuda_ib
from afsi_gpu_amd.mac2d import ChannelGrid
from afsi_gpu_amd.mac2d.batch import BatchedChannelStep

pytestmark = pytest.mark.gpu
DT = 1e-3


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_channel_batch_matches_oracle(gpu):
    gen = np.random.default_rng(21)
    specs = [((128, 32), True, 3000), ((512, 128), True, 20000), ((256, 64), True, 9000),
             ((64, 32), False, 1500), ((128, 32), True, 1), ((128, 64), False, 4000)]
    grids, chans, xs, Fs = [], [], [], []
    for shape, periodic, count in specs:
        lengths = (4., 1.) if periodic else (2., 1.)
        grids.append(ChannelGrid(shape, lengths, periodic))
        chans.append(mac2d.Channel(shape, lengths, periodic))
        if periodic:
            x = workloads.random_curves(gen, count, shape)
        else:
            x = np.stack([gen.uniform(.3, 1.7, count), gen.uniform(.2, .8, count)], 1)
        xs.append(x)
        Fs.append(gen.normal(size=x.shape))
    opts = MGOptions(smooth=2, smoother='rbgs')
    batch = BatchedChannelStep(grids, [len(x) for x in xs], dt=DT, device=gpu, options=opts)
    for i, (x, F) in enumerate(zip(xs, Fs)):
        batch.x[i].copy_(torch.tensor(x))
        batch.force[i].copy_(torch.tensor(F))
    refs = [dict(x=x, velocity=tuple(np.zeros(ch.face_shape(c)) for c in range(2)), p=None)
            for x, ch in zip(xs, chans)]
    mgs = [mac2d.ChannelMultigrid(ch.shape, ch.spacing, OOptions(smooth=2, smoother='rbgs'), ch.periodic)
           for ch in chans]
    for step in range(2):
        U, info = batch.step()
        for i, (ch, mg, F) in enumerate(zip(chans, mgs, Fs)):
            r = ostep.ib_step_2d(refs[i]['x'], F, refs[i]['velocity'], refs[i]['p'], ch, mg, dt=DT)
            refs[i] = dict(x=r['x'], velocity=r['velocity'], p=r['pressure'])
            a, b = batch.begins[i], batch.begins[i + 1]
            print(step, specs[i][0], 'cycles', info['cycles'][i], r['info']['cycles'])
            assert info['cycles'][i] == r['info']['cycles']
            assert rel(U[a:b].cpu().numpy(), r['point_velocity']) < 1e-10
            for got, want in zip(batch.velocity(i), r['velocity']):
                assert rel(got.cpu().numpy(), want) < 1e-10
            np.testing.assert_allclose(batch.x[i].cpu().numpy(), r['x'], rtol=0, atol=1e-14)


def test_channel_batch_support_error(gpu):
    grid = ChannelGrid((64, 32), (2., 1.), False)
    batch = BatchedChannelStep([grid], [2], dt=DT, device=gpu)
    batch.x[0].copy_(torch.tensor([[.01, .5], [1., .5]]))
    with pytest.raises(ValueError, match='support region'):
        batch.step()


def test_tiled_spread_matches_indexed_spread(gpu):
    """afsi_chb_spread_f64 (points binned by stencil-base tile, LDS boxes,
    one atomic per face and tile) against the per-component indexed spread
    over the same stencils: periodic channels with points outside [0, Lx)
    (wrapped bases) and near both walls (reflected rows), an inlet/outlet
    channel at the support limits, an empty instance, and a dense cluster
    that fills several chunks of one tile."""
    gen = np.random.default_rng(33)
    shapes = [((128, 32), True), ((64, 32), False), ((8, 8), True), ((256, 64), True), ((128, 32), True)]
    grids, xs = [], []
    for k, (shape, periodic) in enumerate(shapes):
        lengths = (4., 1.) if periodic else (2., 1.)
        g = ChannelGrid(shape, lengths, periodic)
        hx, hy = lengths[0] / shape[0], lengths[1] / shape[1]
        if k == 4:
            x = np.zeros((0, 2))
        elif periodic:
            x = np.stack([gen.uniform(-2 * lengths[0], 3 * lengths[0], 5000),
                          gen.uniform(-hy, lengths[1] + hy, 5000)], 1)
            if k == 3:  # 3000 points in one cell
                x = np.concatenate([x, np.stack([gen.uniform(1., 1. + hx, 3000),
                                                 gen.uniform(.5, .5 + hy, 3000)], 1)])
        else:
            x = np.stack([gen.uniform(2 * hx, lengths[0] - 2 * hx, 4000), gen.uniform(-hy, lengths[1] + hy, 4000)], 1)
            x[:4] = [[2 * hx, -hy], [lengths[0] - 2 * hx, lengths[1] + hy], [2 * hx, lengths[1]], [1., 0.]]
        grids.append(g)
        xs.append(x)
    opts = MGOptions(smooth=2, smoother='rbgs')
    tiled = BatchedChannelStep(grids, [len(x) for x in xs], dt=DT, device=gpu, options=opts)
    for i, x in enumerate(xs):
        tiled.x[i].copy_(torch.tensor(x))
    F = torch.tensor(gen.normal(size=(sum(len(x) for x in xs), 2)), device=gpu)
    tiled.flags.zero_()
    _abi.call('afsi_chb_stencil_f64', byref(tiled.batch), _abi.ptr(tiled.ids[0]), _abi.ptr(tiled.w[0]),
              _abi.ptr(tiled.ids[1]), _abi.ptr(tiled.w[1]), _abi.ptr(tiled.flags), _abi.stream())
    _abi.call('afsi_chb_spread_f64', byref(tiled.batch), _abi.ptr(F), _abi.ptr(tiled.fu), _abi.ptr(tiled.fv),
              _abi.ptr(tiled.spread_workspace), tiled.spread_workspace.numel(), _abi.ptr(tiled.flags), _abi.stream())
    assert int(tiled.flags.item()) == 0
    want_u = cuda_ib.indexed_spread(F[:, :1].contiguous(), tiled.ids[0], tiled.w[0], tiled.nu).view(-1)
    want_v = cuda_ib.indexed_spread(F[:, 1:].contiguous(), tiled.ids[1], tiled.w[1], tiled.nv).view(-1)
    for got, want in ((tiled.fu, want_u), (tiled.fv, want_v)):
        err = float((got - want).abs().max() / want.abs().max())
        print('tiled vs indexed spread', err)
        assert err < 1e-13
