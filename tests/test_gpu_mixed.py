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


@pytest.mark.parametrize('relax,inner', [(1., 1), (1.3, 1), (1.3, 3)])
@pytest.mark.parametrize('backend', ['workspace', 'graph'])
@pytest.mark.parametrize('periodic', [False, True])
def test_mixed_solve_reaches_fp64_tolerance(gpu, backend, periodic, relax, inner):
    """relax 1.3: over-relaxed red-black FP32 cycles (the benchmark's mixed
    path, with 3 FP32 cycles per FP64 refinement step) reach the same FP64
    tolerance and oracle solution, in fewer cycles than the plain FP64 solve."""
    shape, lengths = (64, 32, 48), (2., 3., 4.)
    h, exact = manufactured(shape, lengths)
    rhs = mac3d.negative_laplacian(exact, h, periodic)
    grid = MACGrid(shape, lengths, periodic=periodic)
    opts = MGOptions(smooth=2, smoother='rbgs', relax=relax, inner_cycles=inner)
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
    assert info['cycles'] <= info64['cycles'] + 2 * inner
    if relax != 1:
        _, info_gs = GeometricMultigrid(grid, device=gpu, options=MGOptions(smooth=2, smoother='rbgs'),
                                        backend=backend).solve(b)
        print('mixed over-relaxed cycles', info['cycles'], 'plain Gauss-Seidel FP64', info_gs['cycles'])
        assert info['cycles'] < info_gs['cycles']
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


def seven_point(x, w, periodic):
    """A x = sum_a w_a (2 x - x(+e_a) - x(-e_a)), Neumann neighbours omitted."""
    a = torch.zeros_like(x)
    for ax in range(3):
        for step in (1, -1):
            term = w[ax] * (x - torch.roll(x, -step, ax))
            if not periodic:
                edge = [slice(None)] * 3
                edge[ax] = -1 if step == 1 else 0
                term[tuple(edge)] = 0
            a += term
    return a


@pytest.mark.parametrize('periodic', [False, True])
@pytest.mark.parametrize('shape', [(8, 12, 16), (6, 4, 128), (6, 10, 130)])
def test_refine_step_is_update_then_residual(gpu, periodic, shape):
    """afsi_refine_step_3d forms x_out = x + e bit for bit and r = b - A x_out
    (FP32 result within one FP32 rounding of torch's FP64 residual), with the
    block reduction [sum r, sum r^2, sum |r|, non-finite] within 1e-12."""
    from afsi_gpu_amd import _abi
    g = torch.Generator(device='cpu').manual_seed(5)
    x = torch.randn(shape, generator=g, dtype=torch.float64).to(gpu)
    e = torch.randn(shape, generator=g, dtype=torch.float64).float().to(gpu)
    b = torch.randn(shape, generator=g, dtype=torch.float64).to(gpu)
    inv_h2 = (3.0, 1.7, 0.6)
    xo = torch.empty_like(x)
    r32 = torch.empty(shape, device=gpu, dtype=torch.float32)
    partials = torch.zeros(4 * ((shape[0] * shape[1] + 3) // 4) + 64, device=gpu, dtype=torch.float64)
    result = torch.zeros(4, device=gpu, dtype=torch.float64)
    _abi.call('afsi_refine_step_3d', _abi.ptr(x), _abi.ptr(e), _abi.ptr(b), _abi.ptr(xo), _abi.ptr(r32),
              _abi.i64x3(shape), _abi.f64x3(inv_h2), int(periodic), _abi.ptr(partials), partials.numel(),
              _abi.ptr(result), _abi.stream())
    xs = x + e.double()
    assert torch.equal(xo, xs)
    r = b - seven_point(xs, inv_h2, periodic)
    ulp = torch.finfo(torch.float32).eps * r.abs().float() + torch.finfo(torch.float32).tiny
    assert bool(((r32 - r.float()).abs() <= ulp).all())
    want = torch.stack([r.sum(), (r * r).sum(), r.abs().sum(), torch.zeros((), device=gpu, dtype=torch.float64)])
    assert torch.allclose(result, want, rtol=1e-12, atol=1e-12 * float(r.abs().sum()))
