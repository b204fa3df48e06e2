"""The red-black sweep specialisations the 128^3 and 64^3 levels run
(rb_rows_kernel with the z extent 128 or 64 as a compile-time constant, for
2x2, 4x4, 8x4 and 8x8 row tiles) against the oracle, on small grids whose
levels reach them: (32, 32, 128) and (16, 16, 64), both boundary types, FP64
and FP32. One V(2,2) cycle from a manufactured start, then the per-cycle
residual reduction of a solve (FP64 within 1e-10 of the oracle's, same cycle
count; FP32 cycles within FP32 rounding of the FP64 oracle cycle)."""
from math import pi

import numpy as np
import pytest
import torch

from oracle import mac3d
from afsi_gpu_amd.mac import MACGrid, MGOptions
from afsi_gpu_amd.mac.multigrid import GeometricMultigrid

pytestmark = pytest.mark.gpu
LENGTHS = (1., 1.5, 2.)
TILES = {2: (2, 2), 4: (4, 4), 84: (8, 4), 8: (8, 8)}


def manufactured(shape):
    h = tuple(L / n for L, n in zip(LENGTHS, shape))
    axes = [(np.arange(n) + .5) * hh for n, hh in zip(shape, h)]
    x, y, z = np.meshgrid(*axes, indexing='ij')
    exact = np.cos(2 * pi * x) * np.cos(2 * pi * y / 1.5) * np.sin(pi * z) + .3 * np.cos(2 * pi * z / 2.)
    return h, exact - exact.mean()


def solver(gpu, shape, periodic, tile, dtype, check_every=2, relax=1.):
    mg = GeometricMultigrid(MACGrid(shape, LENGTHS, periodic=periodic), device=gpu, dtype=dtype,
                            options=MGOptions(smooth=2, smoother='rbgs', check_every=check_every, relax=relax),
                            backend='workspace')
    for level, s in enumerate(mg.shapes):
        tx, ty = TILES[tile]
        if s[0] % tx == 0 and s[1] % ty == 0:
            mg.plan.sweep_tile[level] = tile
    return mg


@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
@pytest.mark.parametrize('tile', [2, 4, 84, 8])
@pytest.mark.parametrize('periodic', [False, True])
@pytest.mark.parametrize('shape', [(32, 32, 128), (16, 16, 64)])
def test_cycle_with_forced_sweep_tiles(gpu, shape, periodic, tile, dtype):
    h, exact = manufactured(shape)
    rhs = mac3d.negative_laplacian(exact, h, periodic)
    rhs -= rhs.mean()
    start = .2 * exact + .05 * np.random.default_rng(0).normal(size=shape)
    start -= start.mean()
    ref = mac3d.Multigrid(shape, h, mac3d.MGOptions(smooth=2, smoother='rbgs'), periodic)
    expected = ref.cycle(0, start, rhs)
    mg = solver(gpu, shape, periodic, tile, dtype)
    mg.initialize(torch.tensor(start, device=gpu, dtype=dtype), torch.tensor(rhs, device=gpu, dtype=dtype))
    got = mg.cycle().double().cpu().numpy()
    got -= got.mean()
    err = np.linalg.norm(got - (expected - expected.mean())) / np.linalg.norm(expected)
    print(shape, periodic, tile, dtype, 'one-cycle error', err)
    assert err <= (2e-12 if dtype == torch.float64 else 2e-6)


@pytest.mark.parametrize('tile', [4, 8])
@pytest.mark.parametrize('periodic', [False, True])
@pytest.mark.parametrize('shape', [(32, 32, 128), (16, 16, 64)])
def test_solve_history_with_forced_sweep_tiles(gpu, shape, periodic, tile):
    h, exact = manufactured(shape)
    rhs = mac3d.negative_laplacian(exact, h, periodic)
    rhs -= rhs.mean()
    ref = mac3d.Multigrid(shape, h, mac3d.MGOptions(smooth=2, smoother='rbgs', check_every=1), periodic)
    _, info_ref = ref.solve(rhs)
    mg = solver(gpu, shape, periodic, tile, torch.float64, check_every=1)
    hist = []
    p, info = mg.solve(torch.tensor(rhs, device=gpu), history=hist)
    rho = np.array(hist) / hist[0]
    rho_ref = np.array(info_ref['history']) / info_ref['history'][0]
    assert info['cycles'] == info_ref['cycles']
    assert np.abs(rho - rho_ref).max() <= 1e-10


@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
@pytest.mark.parametrize('periodic', [False, True])
@pytest.mark.parametrize('shape', [(32, 32, 128), (16, 16, 64)])
def test_over_relaxed_red_black(gpu, shape, periodic, dtype):
    """Over-relaxed red-black smoothing (relax 1.3) on every level kind the
    fused driver has (row sweeps at z 128/64, fused small levels, the LDS
    coarse cycle): one cycle against the oracle's over-relaxed cycle, and in
    FP64 the per-cycle residual reduction of a solve within 1e-10 with the
    same (smaller) cycle count."""
    relax = 1.3
    h, exact = manufactured(shape)
    rhs = mac3d.negative_laplacian(exact, h, periodic)
    rhs -= rhs.mean()
    start = .2 * exact + .05 * np.random.default_rng(0).normal(size=shape)
    start -= start.mean()
    ref = mac3d.Multigrid(shape, h, mac3d.MGOptions(smooth=2, smoother='rbgs', relax=relax), periodic)
    expected = ref.cycle(0, start, rhs)
    mg = solver(gpu, shape, periodic, 4, dtype, relax=relax)
    mg.initialize(torch.tensor(start, device=gpu, dtype=dtype), torch.tensor(rhs, device=gpu, dtype=dtype))
    got = mg.cycle().double().cpu().numpy()
    got -= got.mean()
    err = np.linalg.norm(got - (expected - expected.mean())) / np.linalg.norm(expected)
    print(shape, periodic, dtype, 'over-relaxed one-cycle error', err)
    assert err <= (2e-12 if dtype == torch.float64 else 2e-6)
    if dtype != torch.float64:
        return
    ref = mac3d.Multigrid(shape, h, mac3d.MGOptions(smooth=2, smoother='rbgs', check_every=1, relax=relax), periodic)
    _, info_ref = ref.solve(rhs)
    plain = mac3d.Multigrid(shape, h, mac3d.MGOptions(smooth=2, smoother='rbgs', check_every=1), periodic)
    _, info_plain = plain.solve(rhs)
    mg = solver(gpu, shape, periodic, 4, torch.float64, check_every=1, relax=relax)
    hist = []
    mg.solve(torch.tensor(rhs, device=gpu), history=hist)
    rho = np.array(hist) / hist[0]
    rho_ref = np.array(info_ref['history']) / info_ref['history'][0]
    print('cycles over-relaxed', len(hist) - 1, 'plain', info_plain['cycles'])
    assert len(hist) - 1 == info_ref['cycles'] < info_plain['cycles']
    assert np.abs(rho - rho_ref).max() <= 1e-10
