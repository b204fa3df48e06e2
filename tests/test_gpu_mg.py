"""GPU parity of the multigrid kernels, V-cycle, solve and projection.

Bars: elementwise kernels (residual, Jacobi, red-black GS colour sweeps,
restriction, fused residual-restriction, prolongation) bit-identical to the
oracle in FP64 (same IEEE operation sequence, no contraction); one V-cycle
within 2e-11 (tests/test_mac_multigrid_fused.py:62 of the reference);
per-cycle residual reduction r_k/r_0 within 1e-10 absolute of the oracle's;
projected velocity within 1e-10 relative (l2) of the oracle's.
"""
from math import pi
import numpy as np
import pytest
import torch

from oracle import mac3d
from afsi_gpu_amd import _abi
from afsi_gpu_amd.mac import GeometricMultigrid, MACFlow, MACGrid, MGOptions

pytestmark = pytest.mark.gpu


def level(shape, lengths):
    h = tuple(L / n for L, n in zip(lengths, shape))
    return h, _abi.i64x3(shape), _abi.f64x3([a * a for a in h]), _abi.f64x3([1 / a ** 2 for a in h])


def field(gen, shape):
    return gen.normal(size=shape)


@pytest.mark.parametrize('periodic', [False, True])
@pytest.mark.parametrize('shape', [(6, 10, 14), (8, 12, 16)])
def test_elementwise_kernels_bit_exact(gpu, periodic, shape):
    gen = np.random.default_rng(1)
    h, n, h2, w = level(shape, (2., 3., 5.))
    p, rhs = field(gen, shape), field(gen, shape)
    tp, tr = torch.tensor(p, device=gpu), torch.tensor(rhs, device=gpu)
    out = torch.full_like(tp, float('nan'))
    s = _abi.stream()
    _abi.call('afsi_mg_residual_3d_f64', _abi.ptr(tp), _abi.ptr(tr), _abi.ptr(out), n, h2, int(periodic), s)
    assert np.array_equal(out.cpu().numpy(), rhs - mac3d.negative_laplacian(p, h, periodic))
    mg = mac3d.Multigrid(shape, h, mac3d.MGOptions(smooth=1), periodic) if min(shape) > 4 else None
    _abi.call('afsi_mg_jacobi_3d_f64', _abi.ptr(tp), _abi.ptr(tr), _abi.ptr(out), n, h2, w, 2 / 3, int(periodic), s)
    diag = mac3d.diagonal(shape, h, periodic)
    expected = p + (2 / 3) * (rhs - mac3d.negative_laplacian(p, h, periodic)) / diag
    assert np.array_equal(out.cpu().numpy(), expected)
    # red then black Gauss-Seidel colours
    gs = tp.clone()
    _abi.call('afsi_mg_rbgs_3d_f64', _abi.ptr(gs), _abi.ptr(tr), n, h2, w, 0, int(periodic), s)
    _abi.call('afsi_mg_rbgs_3d_f64', _abi.ptr(gs), _abi.ptr(tr), n, h2, w, 1, int(periodic), s)
    assert mg is not None
    assert np.array_equal(gs.cpu().numpy(), mg.rbgs(0, p, rhs))
    coarse = torch.full(tuple(k // 2 for k in shape), float('nan'), device=gpu, dtype=torch.float64)
    _abi.call('afsi_mg_restrict_3d_f64', _abi.ptr(tp), _abi.ptr(coarse), _abi.i64x3(coarse.shape), s)
    assert np.array_equal(coarse.cpu().numpy(), mac3d.restrict(p))
    _abi.call('afsi_mg_residual_restrict_3d_f64', _abi.ptr(tp), _abi.ptr(tr), _abi.ptr(coarse), n, h2, int(periodic), s)
    assert np.array_equal(coarse.cpu().numpy(), mac3d.restrict(rhs - mac3d.negative_laplacian(p, h, periodic)))
    fine = tp.clone()
    _abi.call('afsi_mg_prolong_add_3d_f64', _abi.ptr(coarse), _abi.ptr(fine), n, int(periodic), s)
    assert np.array_equal(fine.cpu().numpy(), p + mac3d.prolong(coarse.cpu().numpy(), shape, periodic))
    # A.const = 0 exactly (tests/test_mac_multigrid_fused.py:40-42)
    const = torch.full_like(tp, 7.)
    zero = torch.zeros_like(tp)
    _abi.call('afsi_mg_residual_3d_f64', _abi.ptr(const), _abi.ptr(zero), _abi.ptr(out), n, h2, int(periodic), s)
    assert out.count_nonzero().item() == 0


def test_f32_kernels_close(gpu):
    shape = (8, 12, 16)
    gen = np.random.default_rng(2)
    h, n, h2, w = level(shape, (2., 3., 5.))
    p, rhs = field(gen, shape), field(gen, shape)
    tp = torch.tensor(p, device=gpu, dtype=torch.float32)
    tr = torch.tensor(rhs, device=gpu, dtype=torch.float32)
    out = torch.empty_like(tp)
    _abi.call('afsi_mg_residual_3d_f32', _abi.ptr(tp), _abi.ptr(tr), _abi.ptr(out), n, h2, 0, _abi.stream())
    ref = rhs - mac3d.negative_laplacian(p, h)
    np.testing.assert_allclose(out.cpu().numpy(), ref, rtol=2e-5, atol=2e-5 * np.abs(ref).max())


def test_reduction_deterministic(gpu):
    gen = np.random.default_rng(3)
    x = torch.tensor(gen.normal(size=1_000_003), device=gpu)
    nparts = _abi.load().afsi_reduce_partials(x.numel())
    partials = torch.empty(4 * nparts, device=gpu, dtype=torch.float64)
    res = torch.empty(4, device=gpu, dtype=torch.float64)
    outs = []
    for _ in range(3):
        _abi.call('afsi_reduce_f64', _abi.ptr(x), x.numel(), _abi.ptr(partials), _abi.ptr(res), _abi.stream())
        outs.append(res.cpu().numpy().copy())
    xs = x.cpu().numpy()
    np.testing.assert_allclose(outs[0][:3], [xs.sum(), (xs * xs).sum(), np.abs(xs).sum()], rtol=1e-12)
    assert all(np.array_equal(outs[0], o) for o in outs[1:])


def manufactured(shape, lengths):
    h = tuple(L / n for L, n in zip(lengths, shape))
    axes = [(np.arange(n) + .5) * hh for n, hh in zip(shape, h)]
    x, y, z = np.meshgrid(*axes, indexing='ij')
    exact = np.cos(pi * x / 2) * np.cos(2 * pi * y / 3) * np.cos(pi * z / 4)
    exact -= exact.mean()
    return h, exact


@pytest.mark.parametrize('backend', ['workspace', 'unfused'])
@pytest.mark.parametrize('smoother', ['jacobi', 'rbgs'])
@pytest.mark.parametrize('periodic', [False, True])
@pytest.mark.parametrize('shape', [(4, 4, 4), (8, 12, 16), (16, 16, 16), (32, 32, 32), (64, 32, 48)])
def test_v_cycle_matches_oracle(gpu, backend, smoother, periodic, shape):
    # (16,16,16) runs entirely in the single-workgroup LDS kernel; (32,32,32)
    # and (64,32,48) add fine levels (partial z tiles for 48 and 24).
    lengths = (2., 3., 4.)
    h, exact = manufactured(shape, lengths)
    rhs = mac3d.negative_laplacian(exact, h, periodic)
    rhs -= rhs.mean()
    start = .1 * exact
    opts = MGOptions(smooth=3, smoother=smoother)
    ref = mac3d.Multigrid(shape, h, mac3d.MGOptions(smooth=3, smoother=smoother), periodic)
    expected = ref.cycle(0, start - start.mean(), rhs)
    mg = GeometricMultigrid(MACGrid(shape, lengths, periodic=periodic), device=gpu, options=opts,
                            backend=backend)
    mg.initialize(torch.tensor(start - start.mean(), device=gpu), torch.tensor(rhs, device=gpu))
    got = mg.cycle().cpu().numpy()
    np.testing.assert_allclose(got, expected, rtol=2e-11, atol=2e-12 * np.abs(expected).max())


@pytest.mark.parametrize('backend', ['workspace', 'graph', 'unfused'])
@pytest.mark.parametrize('smoother', ['jacobi', 'rbgs'])
@pytest.mark.parametrize('periodic', [False, True])
def test_solve_residual_history_matches_oracle(gpu, backend, smoother, periodic):
    shape, lengths = (64, 32, 48), (2., 3., 4.)
    h, exact = manufactured(shape, lengths)
    rhs = mac3d.negative_laplacian(exact, h, periodic)
    options = MGOptions(smooth=2, smoother=smoother, check_every=1)
    ref = mac3d.Multigrid(shape, h, mac3d.MGOptions(smooth=2, smoother=smoother, check_every=1), periodic)
    p_ref, info_ref = ref.solve(rhs)
    mg = GeometricMultigrid(MACGrid(shape, lengths, periodic=periodic), device=gpu, options=options,
                            backend=backend)
    hist = []
    p, info = mg.solve(torch.tensor(rhs, device=gpu), history=hist)
    r0, r0_ref = hist[0], info_ref['history'][0]
    rho = np.array(hist) / r0
    rho_ref = np.array(info_ref['history']) / r0_ref
    print(smoother, periodic, 'cycles', info['cycles'], 'reduction', rho[-1])
    assert info['cycles'] == info_ref['cycles']
    assert np.abs(rho - rho_ref).max() <= 1e-10
    assert abs(r0 / r0_ref - 1) < 1e-12
    np.testing.assert_allclose(p.cpu().numpy(), p_ref, rtol=0, atol=1e-10 * np.abs(p_ref).max())
    np.testing.assert_allclose(p.cpu().numpy(), exact, rtol=0, atol=1e-7 * np.abs(exact).max())


@pytest.mark.parametrize('backend', ['workspace', 'graph', 'unfused'])
@pytest.mark.parametrize('periodic', [False, True])
def test_over_relaxed_solve_history_matches_oracle(gpu, backend, periodic):
    """relax = 1.3 red-black smoothing in every 3D driver against the
    oracle's over-relaxed V-cycle: same cycle count, per-cycle residual
    reduction within 1e-10."""
    shape, lengths = (64, 32, 48), (2., 3., 4.)
    h, exact = manufactured(shape, lengths)
    rhs = mac3d.negative_laplacian(exact, h, periodic)
    ref = mac3d.Multigrid(shape, h, mac3d.MGOptions(smooth=2, smoother='rbgs', check_every=1, relax=1.3), periodic)
    p_ref, info_ref = ref.solve(rhs)
    mg = GeometricMultigrid(MACGrid(shape, lengths, periodic=periodic), device=gpu, backend=backend,
                            options=MGOptions(smooth=2, smoother='rbgs', check_every=1, relax=1.3))
    hist = []
    p, info = mg.solve(torch.tensor(rhs, device=gpu), history=hist)
    rho = np.array(hist) / hist[0]
    rho_ref = np.array(info_ref['history']) / info_ref['history'][0]
    print(backend, periodic, 'cycles', info['cycles'])
    assert info['cycles'] == info_ref['cycles']
    assert np.abs(rho - rho_ref).max() <= 1e-10
    np.testing.assert_allclose(p.cpu().numpy(), p_ref, rtol=0, atol=1e-10 * np.abs(p_ref).max())


def test_solve_errors_and_ownership(gpu):
    shape, lengths = (16, 16, 16), (2., 2., 2.)
    h, exact = manufactured(shape, lengths)
    rhs = torch.tensor(mac3d.negative_laplacian(exact, h), device=gpu)
    mg = GeometricMultigrid(MACGrid(shape, lengths), device=gpu, backend='graph')
    p, info = mg.solve(rhs)
    saved = p.clone()
    p2, _ = mg.solve(-rhs, p)
    assert torch.equal(p, saved)
    np.testing.assert_allclose(p2.cpu().numpy(), -exact, rtol=0, atol=2e-8)
    zero, zinfo = mg.solve(torch.zeros_like(rhs))
    assert zinfo['cycles'] == 0 and zero.count_nonzero().item() == 0
    with pytest.raises(ValueError, match='incompatible'):
        mg.solve(torch.ones_like(rhs))
    with pytest.raises(ValueError, match='invalid pressure RHS'):
        mg.solve(torch.full_like(rhs, float('nan')))
    tight = GeometricMultigrid(MACGrid(shape, lengths), device=gpu, backend='graph',
                               options=MGOptions(max_cycles=2, rtol=0, atol=1e-30, check_every=3))
    with pytest.raises(RuntimeError, match='did not converge'):
        tight.solve(rhs)
    assert set(tight.graphs) == {1, 2, 3}


@pytest.mark.parametrize('periodic', [False, True])
def test_projection_matches_oracle(gpu, periodic):
    shape, lengths = (32, 32, 32), (2., 3., 4.)
    h = tuple(L / n for L, n in zip(lengths, shape))
    grid = MACGrid(shape, lengths, periodic=periodic)
    gen = np.random.default_rng(8)
    u = [gen.normal(size=grid.face_shape(c)) for c in range(3)]
    if not periodic:
        for c in range(3):
            u[c][mac3d.slab(c, 0, 1)] = 0
            u[c][mac3d.slab(c, -1, None)] = 0
    options = MGOptions(smooth=2, smoother='jacobi')
    ref = mac3d.Multigrid(shape, h, mac3d.MGOptions(smooth=2), periodic)
    v_ref, p_ref, _ = mac3d.project(tuple(u), h, ref, dt=1e-3)
    flow = MACFlow(grid, dt=1e-3, device=gpu, options=options, predictor=False)
    res = flow.project(tuple(torch.tensor(a, device=gpu) for a in u))
    for a, b in zip(res.velocity, v_ref):
        err = np.linalg.norm(a.cpu().numpy() - b) / np.linalg.norm(b)
        assert err < 1e-10
    div = mac3d.divergence(tuple(a.cpu().numpy() for a in res.velocity), h, periodic)
    div0 = mac3d.divergence(tuple(u), h, periodic)
    assert np.linalg.norm(div) <= 1e-10 * np.linalg.norm(div0) * 1.01
    if not periodic:
        with pytest.raises(ValueError, match='normal velocity'):
            bad = [torch.tensor(a, device=gpu) for a in u]
            bad[0][0] = 1.
            flow.project(tuple(bad))


def test_autotuned_sweeps_keep_results(gpu):
    """Tuned tile/grid choices change only speed: a tuned solve equals the
    default one bit for bit."""
    from afsi_gpu_amd import autotune
    grid = MACGrid((64, 64, 64), (1., 1., 1.), periodic=True)
    opts = MGOptions(smooth=2, smoother='rbgs')
    gen = np.random.default_rng(3)
    rhs = gen.normal(size=grid.shape)
    rhs -= rhs.mean()
    b = torch.tensor(rhs, device=gpu)
    ref = GeometricMultigrid(grid, device=gpu, options=opts, backend='workspace')
    p0, i0 = ref.solve(b)
    tuned = GeometricMultigrid(grid, device=gpu, options=opts, backend='workspace')
    chosen = autotune.tune_sweeps(tuned)
    assert 0 in chosen and chosen[0][0] in (2, 4)
    p1, i1 = tuned.solve(b)
    assert i0['cycles'] == i1['cycles']
    assert torch.equal(p0, p1)
    cache = autotune.export_cache()
    autotune.import_cache(cache)
    assert autotune.export_cache() == cache
