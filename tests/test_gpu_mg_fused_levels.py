"""Fused small multigrid levels (vdown_kernel: two red-black sweeps, the
residual and its restriction in one launch; vup_kernel: the prolongation
and two sweeps in another) against the same driver with every level
unfused (plan.fuse_cells = -1), and against the oracle.

The fused kernels form every value with the expressions and summation order
of the unfused sweep, restriction and prolongation kernels, and the library
is compiled without floating-point contraction, so the two drivers must
agree bit for bit; the oracle comparison pins both."""
from math import pi

import numpy as np
import pytest
import torch

from oracle import mac3d
from afsi_gpu_amd.mac import MACGrid, MGOptions
from afsi_gpu_amd.mac.multigrid import GeometricMultigrid, MixedMultigrid

pytestmark = pytest.mark.gpu
LENGTHS = (1., 1.5, 2.)


def manufactured(shape, periodic):
    h = tuple(L / n for L, n in zip(LENGTHS, shape))
    axes = [(np.arange(n) + .5) * hh for n, hh in zip(shape, h)]
    x, y, z = np.meshgrid(*axes, indexing='ij')
    exact = np.cos(2 * pi * x) * np.cos(2 * pi * y / 1.5) * np.sin(pi * z) + .3 * np.cos(2 * pi * z / 2.)
    exact -= exact.mean()
    rhs = mac3d.negative_laplacian(exact, h, periodic)
    rhs -= rhs.mean()
    start = .2 * exact + .05 * np.random.default_rng(1).normal(size=shape)
    return h, rhs, start - start.mean()


def solver(gpu, shape, periodic, dtype, fuse):
    mg = GeometricMultigrid(MACGrid(shape, LENGTHS, periodic=periodic), device=gpu, dtype=dtype,
                            options=MGOptions(smooth=2, smoother='rbgs', check_every=1), backend='workspace')
    mg.plan.fuse_cells = 0 if fuse else -1
    return mg


def fused_levels(shapes):
    """Levels the driver fuses by default (z extent 32/64, <= 64^3 cells,
    above the LDS coarse levels)."""
    return [s for s in shapes[:-1] if s[2] in (32, 64) and np.prod(s) <= 64 ** 3 and np.prod(s) > 4096]


@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
@pytest.mark.parametrize('periodic', [False, True])
@pytest.mark.parametrize('shape', [(64, 64, 64), (32, 32, 32), (16, 24, 64), (128, 128, 128)])
def test_fused_levels_match_unfused_bitwise(gpu, shape, periodic, dtype):
    _, rhs, start = manufactured(shape, periodic)
    out = []
    for fuse in (True, False):
        mg = solver(gpu, shape, periodic, dtype, fuse)
        assert fused_levels(mg.shapes), 'shape reaches no fused level'
        mg.initialize(torch.tensor(start, device=gpu, dtype=dtype), torch.tensor(rhs, device=gpu, dtype=dtype))
        for _ in range(3):
            mg._cycles(1)
        torch.cuda.synchronize()
        out.append(mg.p[0].clone())
    diff = (out[0] - out[1]).abs().max().item()
    print(shape, periodic, dtype, 'max |fused - unfused|', diff)
    assert torch.equal(out[0], out[1])


@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
@pytest.mark.parametrize('periodic', [False, True])
@pytest.mark.parametrize('shape', [(32, 32, 32), (16, 24, 64)])
def test_fused_cycle_matches_oracle(gpu, shape, periodic, dtype):
    h, rhs, start = manufactured(shape, periodic)
    ref = mac3d.Multigrid(shape, h, mac3d.MGOptions(smooth=2, smoother='rbgs'), periodic)
    expected = ref.cycle(0, start, rhs)
    expected -= expected.mean()
    mg = solver(gpu, shape, periodic, dtype, True)
    mg.initialize(torch.tensor(start, device=gpu, dtype=dtype), torch.tensor(rhs, device=gpu, dtype=dtype))
    got = mg.cycle().double().cpu().numpy()
    got -= got.mean()
    err = np.linalg.norm(got - expected) / np.linalg.norm(expected)
    print(shape, periodic, dtype, 'one-cycle error vs oracle', err)
    assert err <= (2e-12 if dtype == torch.float64 else 2e-6)


@pytest.mark.parametrize('periodic', [False, True])
def test_mixed_solve_fused_matches_unfused(gpu, periodic):
    """The mixed solve's FP32 cycles start from a zero iterate
    (zero_initial): the fused down kernel then reads no p at level 0."""
    shape = (64, 64, 64)
    _, rhs, _ = manufactured(shape, periodic)
    results = []
    for fuse in (True, False):
        mg = MixedMultigrid(MACGrid(shape, LENGTHS, periodic=periodic), device=gpu,
                            options=MGOptions(smooth=2, smoother='rbgs', check_every=1), backend='workspace')
        mg.inner.plan.fuse_cells = 0 if fuse else -1
        p, info = mg.solve(torch.tensor(rhs, device=gpu))
        results.append((p, info['cycles'], info['residual_norm']))
    assert results[0][1] == results[1][1]
    assert results[0][2] == results[1][2]
    assert torch.equal(results[0][0], results[1][0])
