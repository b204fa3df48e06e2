"""The 64-instance batch of uneven 3D and 2D instances (grids drawn from
{32^3, 48^3, 64^3, 96^3, 128x32, 256x64, 512x128}, 1e4-3e5 points each,
random shells / sinusoidal curves, N(0,1) forces), advanced two steps with
one launch per phase: the 3D instances by ``BatchedIBStep`` (mixed
precision, pressures passed back as warm starts), the 2D channels by
``BatchedChannelStep``.

Every instance is checked against its single-instance path, which the
oracle pins elsewhere at these sizes (``IBStep`` mixed up to 128^3,
``ChannelIBStep`` on the channel grids): the same cycle counts (the
warm-started 3D step within one: the single solve checks every cycle past
its previous count, the batch every check_every cycles), point
velocities within 1e-9 (3D: FP32 V-cycles in both, summation order of the
spread differs) and 1e-10 (2D), and the displacements x - x0 to the same
bounds. The smallest 3D
and 2D instances are also stepped by the oracle itself (interpolated
velocity within 1e-8 of the FP64 oracle for the mixed 3D path, 1e-10 for
the 2D channel)."""
import numpy as np
import pytest
import torch

from oracle import ib3d, mac2d as omac2d, mac3d
from oracle import step as ostep
from oracle.mac3d import MGOptions as OOptions
from afsi_gpu_amd import workloads
from afsi_gpu_amd.batch import BatchedIBStep, ConcurrentGroups
from afsi_gpu_amd.mac import MACGrid, MGOptions
from afsi_gpu_amd.mac2d import ChannelGrid, ChannelIBStep
from afsi_gpu_amd.mac2d.batch import BatchedChannelStep
from afsi_gpu_amd.step import IBStep

pytestmark = pytest.mark.gpu
DT = 1e-3


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize('concurrent', [False, True])
def test_batch_of_64_mixed_instances(gpu, concurrent):
    """``concurrent``: the 3D and 2D groups step on their own streams and
    host threads (``ConcurrentGroups``), to the same bars."""
    gen = workloads.rng(workloads.GEOMETRY_SEED + 4)
    specs = [workloads.instance_spec(gen, 'batch') for _ in range(64)]
    data = [workloads.instance_points(gen, shape, count) for shape, count in specs]
    i3 = [i for i, (s, _) in enumerate(specs) if len(s) == 3]
    i2 = [i for i, (s, _) in enumerate(specs) if len(s) == 2]
    assert i3 and i2 and any(specs[i][0] == (96, 96, 96) for i in i3)
    opts = MGOptions(smooth=2, smoother='rbgs')
    g3 = [MACGrid(specs[i][0], (1., 1., 1.), periodic=True) for i in i3]
    g2 = [ChannelGrid(specs[i][0], (4., 1.), periodic=True) for i in i2]

    # batched: 3D
    xs3 = [torch.tensor(data[i][0], device=gpu) for i in i3]
    fs3 = [torch.tensor(data[i][1], device=gpu) for i in i3]
    vel3 = [g.zeros(device=gpu) for g in g3]
    batch3 = BatchedIBStep(g3, [specs[i][1] for i in i3], dt=DT, device=gpu, options=opts)
    # batched: 2D
    batch2 = BatchedChannelStep(g2, [specs[i][1] for i in i2], dt=DT, device=gpu, options=opts)
    for j, i in enumerate(i2):
        batch2.x[j].copy_(torch.tensor(data[i][0]))
        batch2.force[j].copy_(torch.tensor(data[i][1]))
    got3, got2 = [], []
    p3 = None
    groups = ConcurrentGroups(gpu) if concurrent else None
    for _ in range(2):
        step3 = lambda: batch3.step(xs3, fs3, vel3, p3)
        if groups is not None:
            (vel3, p3, U3, info3), (U2, info2) = groups.run(step3, batch2.step)
        else:
            vel3, p3, U3, info3 = step3()
            U2, info2 = batch2.step()
        got3.append(([U.cpu().numpy() for U in U3], [inf['cycles'] for inf in info3]))
        got2.append((U2.cpu().numpy(), list(info2['cycles'])))
    if groups is not None:
        groups.close()

    # single-instance paths
    worst3 = worst2 = 0.
    for j, (i, g) in enumerate(zip(i3, g3)):
        single = IBStep(g, specs[i][1], dt=DT, device=gpu, options=opts, precision='mixed')
        x = torch.tensor(data[i][0], device=gpu)
        f = torch.tensor(data[i][1], device=gpu)
        v, p = g.zeros(device=gpu), None
        for k in range(2):
            v, p, U, info = single.step(x, f, v, p)
            # from rest both check every check_every cycles; warm-started, the
            # single solve checks after every cycle past its previous count
            slack = 0 if k == 0 else 1
            assert abs(got3[k][1][j] - info['pressure']['cycles']) <= slack, (specs[i], k)
            worst3 = max(worst3, rel(got3[k][0][j], U.cpu().numpy()))
        assert rel(xs3[j].cpu().numpy() - data[i][0], x.cpu().numpy() - data[i][0]) < 1e-9
    for j, (i, g) in enumerate(zip(i2, g2)):
        single = ChannelIBStep(g, dt=DT, device=gpu, options=opts, pressure_backend='workspace')
        x = torch.tensor(data[i][0], device=gpu)
        f = torch.tensor(data[i][1], device=gpu)
        v, p = g.zeros(device=gpu), None
        a, b = batch2.begins[j], batch2.begins[j + 1]
        for k in range(2):
            v, p, U, info = single.step(x, f, v, p)
            assert got2[k][1][j] == info['pressure']['cycles'], (specs[i], k)
            worst2 = max(worst2, rel(got2[k][0][a:b], U.cpu().numpy()))
        assert rel(batch2.x[j].cpu().numpy() - data[i][0], x.cpu().numpy() - data[i][0]) < 1e-10
    print(f'{len(i3)} 3D + {len(i2)} 2D instances; worst point-velocity difference 3D {worst3:.2e}, 2D {worst2:.2e}')
    assert worst3 < 1e-9
    assert worst2 < 1e-10

    # the oracle on the smallest instance of each kind, first step from rest
    j3 = min(range(len(i3)), key=lambda j: np.prod(specs[i3[j]][0]) * 1e9 + specs[i3[j]][1])
    i = i3[j3]
    grid = ib3d.Grid3(specs[i][0], (1., 1., 1.), periodic=True)
    mg = mac3d.Multigrid(grid.shape, grid.spacing, OOptions(smooth=2, smoother='rbgs'), periodic=True)
    u0 = tuple(np.zeros(grid.face_shape(c)) for c in range(3))
    ref = ostep.ib_step(data[i][0], data[i][1], u0, None, grid, mg, dt=DT)
    err3 = rel(got3[0][0][j3], ref['point_velocity'])
    j2 = min(range(len(i2)), key=lambda j: specs[i2[j]][1])
    i = i2[j2]
    ch = omac2d.Channel(specs[i][0], (4., 1.), True)
    mg2 = omac2d.ChannelMultigrid(ch.shape, ch.spacing, OOptions(smooth=2, smoother='rbgs'), ch.periodic)
    r2 = ostep.ib_step_2d(data[i][0], data[i][1], tuple(np.zeros(ch.face_shape(c)) for c in range(2)), None,
                          ch, mg2, dt=DT)
    a, b = batch2.begins[j2], batch2.begins[j2 + 1]
    err2 = rel(got2[0][0][a:b], r2['point_velocity'])
    print('vs oracle: 3D', specs[i3[j3]], err3, '2D', specs[i], err2)
    assert err3 < 1e-8
    assert err2 < 1e-10
    assert got2[0][1][j2] == r2['info']['cycles']
