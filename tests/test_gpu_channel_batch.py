"""Batched channel instances (mac2d/batch.py) against the oracle: each
instance of an uneven batch (128x32, 256x64 and 512x128 periodic-x channels
and two reference channels with inlet/outlet, 64x32 and 128x64) advanced two
steps with one launch per phase, against oracle/step.ib_step_2d on the same
inputs: the solve's cycle count equal, projected velocity and interpolated
point velocity within 1e-10 relative, positions within 1e-14. The levels
above 4096 cells of the 256x64, 512x128 and 128x64 instances run as batched
launches around a one-workgroup tail; the other instances cycle in one
workgroup each."""
import numpy as np
import pytest
import torch

from oracle import mac2d, step as ostep
from oracle.mac3d import MGOptions as OOptions
from afsi_gpu_amd import workloads
from afsi_gpu_amd.mac import MGOptions
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
