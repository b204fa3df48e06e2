"""Mixed precision (FP32 V-cycles inside FP64 iterative refinement).

Bars: the FP64 residual meets the FP64 tolerance max(atol, 1e-10 ||b||);
the pressure agrees with the FP64 oracle's solution within 1e-8 relative
(both are 1e-10-residual solutions of the same system); the projected
velocity's divergence is <= 1e-10 of the tentative one's; the outer
iteration count is within two of the FP64 solver's cycle count.
"""
from math import pi
import numpy as np
import pytest
import torch

from oracle import mac3d
from afsi_gpu_amd.mac import GeometricMultigrid, MACFlow, MACGrid, MGOptions, MixedMultigrid

pytestmark = pytest.mark.gpu


def manufactured(shape, lengths):
    h = tuple(L / n for L, n in zip(lengths, shape))
    axes = [(np.arange(n) + .5) * hh for n, hh in zip(shape, h)]
    x, y, z = np.meshgrid(*axes, indexing='ij')
    exact = np.cos(pi * x / 2) * np.cos(2 * pi * y / 3) * np.cos(pi * z / 4)
    return h, exact - exact.mean()


@pytest.mark.parametrize('backend', ['workspace', 'graph'])
@pytest.mark.parametrize('periodic', [False, True])
def test_mixed_solve_reaches_fp64_tolerance(gpu, backend, periodic):
    shape, lengths = (64, 32, 48), (2., 3., 4.)
    h, exact = manufactured(shape, lengths)
    rhs = mac3d.negative_laplacian(exact, h, periodic)
    grid = MACGrid(shape, lengths, periodic=periodic)
    opts = MGOptions(smooth=2, smoother='rbgs')
    mixed = MixedMultigrid(grid, device=gpu, options=opts, backend=backend)
    plain = GeometricMultigrid(grid, device=gpu, options=opts, backend=backend)
    b = torch.tensor(rhs, device=gpu)
    p, info = mixed.solve(b)
    p64, info64 = plain.solve(b)
    ref = mac3d.Multigrid(shape, h, mac3d.MGOptions(smooth=2, smoother='rbgs'), periodic)
    p_ref, _ = ref.solve(rhs)
    got = p.cpu().numpy()
    true_res = np.linalg.norm(rhs - rhs.mean() - mac3d.negative_laplacian(got, h, periodic))
    assert true_res <= info['tolerance'] * 1.0001
    assert np.linalg.norm(got - p_ref) / np.linalg.norm(p_ref) < 1e-8
    assert abs(got.mean()) < 1e-12 * np.abs(got).max()
    assert info['cycles'] <= info64['cycles'] + 2
    with pytest.raises(ValueError, match='incompatible'):
        mixed.solve(torch.ones_like(b))


@pytest.mark.parametrize('periodic', [False, True])
def test_mixed_projection_divergence_and_velocity(gpu, periodic):
    shape, lengths = (32, 32, 32), (2., 3., 4.)
    h = tuple(L / n for L, n in zip(lengths, shape))
    grid = MACGrid(shape, lengths, periodic=periodic)
    gen = np.random.default_rng(8)
    u = [gen.normal(size=grid.face_shape(c)) for c in range(3)]
    if not periodic:
        for c in range(3):
            u[c][mac3d.slab(c, 0, 1)] = 0
            u[c][mac3d.slab(c, -1, None)] = 0
    ref = mac3d.Multigrid(shape, h, mac3d.MGOptions(smooth=2, smoother='rbgs'), periodic)
    v_ref, _, _ = mac3d.project(tuple(u), h, ref, dt=1e-3)
    flow = MACFlow(grid, dt=1e-3, device=gpu, options=MGOptions(smooth=2, smoother='rbgs'),
                   precision='mixed', predictor=False)
    res = flow.project(tuple(torch.tensor(a, device=gpu) for a in u))
    vel = tuple(a.cpu().numpy() for a in res.velocity)
    div = mac3d.divergence(vel, h, periodic)
    div0 = mac3d.divergence(tuple(u), h, periodic)
    assert np.linalg.norm(div) <= 1e-10 * np.linalg.norm(div0) * 1.01
    for a, b in zip(vel, v_ref):
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-9
