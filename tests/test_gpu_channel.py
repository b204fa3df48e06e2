"""GPU parity of the 2D channel path (reference BCs and periodic-x).

Bars: reflected stencil ids and weights bit-identical to the oracle;
spread and gather within 1e-12 relative; elementwise multigrid kernels
bit-identical in FP64; one V-cycle within 1e-12 relative (the reference
channel cycle has no reductions; the periodic one removes means); solve
histories r_k/r_0 within 1e-10 of the oracle's; projected velocity within
1e-10 relative; the channel IB step's point velocity within 1e-10 relative.
"""
import numpy as np
import pytest
import torch

from oracle import mac2d, step as ostep
from oracle.mac3d import MGOptions as OOptions
from afsi_gpu_amd import _abi, workloads
from afsi_gpu_amd.mac import MGOptions
from afsi_gpu_amd.mac2d import (ChannelFlow, ChannelGrid, ChannelIBStep, ChannelMultigrid,
                                gather, prepare, spread)

pytestmark = pytest.mark.gpu
MODES = [False, True]


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def channel_pair(shape, lengths, periodic):
    return ChannelGrid(shape, lengths, periodic), mac2d.Channel(shape, lengths, periodic)


def wall_points(ch, count, seed):
    gen = np.random.default_rng(seed)
    lo, hi = (0., ch.lengths[0]) if ch.periodic else (2.2 * ch.spacing[0], ch.lengths[0] - 2.2 * ch.spacing[0])
    x = gen.uniform(lo, hi, count)
    y = np.where(gen.random(count) < .5, gen.uniform(-.9, 2., count) * ch.spacing[1],
                 gen.uniform(0, ch.lengths[1], count))
    y = np.where(gen.random(count) < .3, ch.lengths[1] - y, y)
    return np.stack([x, y], 1)


@pytest.mark.parametrize('periodic', MODES)
def test_prepare_bit_exact_and_transfers(gpu, periodic):
    grid, ch = channel_pair((64, 32), (4., 1.), periodic)
    pts = wall_points(ch, 3000, 1)
    ids_ref, w_ref = mac2d.prepare(pts, ch)
    st = prepare(grid, torch.tensor(pts, device=gpu))
    for c in range(2):
        assert np.array_equal(st.indices[c].cpu().numpy(), ids_ref[c])
        assert np.array_equal(st.weights[c].cpu().numpy(), w_ref[c])
    F = np.random.default_rng(2).normal(size=pts.shape)
    dens = spread(grid, torch.tensor(F, device=gpu), st)
    dens_ref = mac2d.spread(F, ids_ref, w_ref, ch)
    for a, b in zip(dens, dens_ref):
        assert a.shape == b.shape and rel(a.cpu().numpy(), b) < 1e-12
    vel = tuple(np.sin((c + 1) * ch.coordinates(c).sum(-1)) for c in range(2))
    U = gather(grid, tuple(torch.tensor(v, device=gpu) for v in vel), st)
    assert rel(U.cpu().numpy(), mac2d.gather(vel, ids_ref, w_ref)) < 1e-12


def test_prepare_rejects_points_outside_support(gpu):
    grid = ChannelGrid((32, 16), (4., 1.))
    hx, hy = grid.spacing
    for bad in ([1.9 * hx, .5], [1., -1.01 * hy], [1., float('nan')]):
        with pytest.raises(ValueError, match='support'):
            prepare(grid, torch.tensor([[1., .5], bad], device=gpu, dtype=torch.float64))
    empty = prepare(grid, torch.zeros((0, 2), device=gpu, dtype=torch.float64))
    assert empty.indices[0].shape == (0, 16)


def level(shape, lengths):
    h = tuple(L / n for L, n in zip(lengths, shape))
    n = (_abi.c_int64 * 2)(*shape)
    h2 = (_abi.c_double * 2)(*[a ** 2 for a in h])
    w = (_abi.c_double * 2)(*[1 / a ** 2 for a in h])
    return h, n, h2, w


@pytest.mark.parametrize('periodic', MODES)
@pytest.mark.parametrize('shape', [(16, 8), (12, 6), (10, 5)])
def test_elementwise_kernels_bit_exact(gpu, periodic, shape):
    gen = np.random.default_rng(3)
    h, n, h2, w = level(shape, (2., 1.))
    p, rhs = gen.normal(size=shape), gen.normal(size=shape)
    tp, tr = torch.tensor(p, device=gpu), torch.tensor(rhs, device=gpu)
    out = torch.full_like(tp, float('nan'))
    s = _abi.stream()
    A = mac2d.negative_laplacian(p, h, periodic)
    _abi.call('afsi_mg_residual_2d_f64', _abi.ptr(tp), _abi.ptr(tr), _abi.ptr(out), n, h2, int(periodic), s)
    assert np.array_equal(out.cpu().numpy(), rhs - A)
    _abi.call('afsi_mg_jacobi_2d_f64', _abi.ptr(tp), _abi.ptr(tr), _abi.ptr(out), n, h2, w, 2 / 3,
              int(periodic), s)
    mg = mac2d.ChannelMultigrid(shape, h, OOptions(smooth=1), periodic) if shape[0] % 2 == 0 else None
    diag = mac2d.diagonal(shape, h, periodic)
    assert np.array_equal(out.cpu().numpy(), p + (2 / 3) * (rhs - A) / diag)
    gs = tp.clone()
    _abi.call('afsi_mg_rbgs_2d_f64', _abi.ptr(gs), _abi.ptr(tr), n, h2, w, int(periodic), s)
    parity = np.add.outer(np.arange(shape[0]), np.arange(shape[1])) % 2
    expected = p.copy()
    for color in (0, 1):
        r = rhs - mac2d.negative_laplacian(expected, h, periodic)
        m = parity == color
        expected[m] = expected[m] + r[m] / diag[m]
    assert np.array_equal(gs.cpu().numpy(), expected)
    if mg is not None:
        assert np.array_equal(gs.cpu().numpy(), mg.rbgs(0, p, rhs))
    if shape[0] % 2 == 0 and shape[1] % 2 == 0:
        coarse = torch.full((shape[0] // 2, shape[1] // 2), float('nan'), device=gpu, dtype=torch.float64)
        _abi.call('afsi_mg_residual_restrict_2d_f64', _abi.ptr(tp), _abi.ptr(tr), _abi.ptr(coarse), n,
                  h2, int(periodic), s)
        assert np.array_equal(coarse.cpu().numpy(), mac2d.restrict(rhs - A))
        fine = tp.clone()
        _abi.call('afsi_mg_prolong_add_2d_f64', _abi.ptr(coarse), _abi.ptr(fine), n, int(periodic), s)
        assert np.array_equal(fine.cpu().numpy(), p + mac2d.prolong(coarse.cpu().numpy(), shape, periodic))
    # divergence and gradient correction
    grid, ch = channel_pair(shape, (2., 1.), periodic)
    u = gen.normal(size=ch.face_shape(0))
    v = gen.normal(size=ch.face_shape(1))
    tu, tv = torch.tensor(u, device=gpu), torch.tensor(v, device=gpu)
    hh = (_abi.c_double * 2)(*h)
    _abi.call('afsi_divergence_2d_f64', _abi.ptr(tu), _abi.ptr(tv), _abi.ptr(out), n, hh, -7.,
              int(periodic), s)
    assert np.array_equal(out.cpu().numpy(), -7. * mac2d.divergence((u, v), h, periodic))
    _abi.call('afsi_grad_correct_2d_f64', _abi.ptr(tu), _abi.ptr(tv), _abi.ptr(tp), n, hh, .3,
              int(periodic), s)
    gx, gy = mac2d.gradient(p, h, periodic)
    assert np.array_equal(tu.cpu().numpy(), u - .3 * gx)
    assert np.array_equal(tv.cpu().numpy(), v - .3 * gy)


@pytest.mark.parametrize('periodic', MODES)
@pytest.mark.parametrize('smoother', ['jacobi', 'rbgs'])
def test_v_cycle_matches_oracle(gpu, periodic, smoother):
    shape, lengths = (64, 32), (4., 1.)
    grid, ch = channel_pair(shape, lengths, periodic)
    opts = MGOptions(smooth=2, smoother=smoother)
    mg = ChannelMultigrid(grid, device=gpu, options=opts, backend='workspace')
    ref = mac2d.ChannelMultigrid(shape, ch.spacing, OOptions(smooth=2, smoother=smoother), periodic)
    gen = np.random.default_rng(4)
    p, rhs = gen.normal(size=shape), gen.normal(size=shape)
    if periodic:
        p, rhs = p - p.mean(), rhs - rhs.mean()
    got = mg.cycle(torch.tensor(p, device=gpu), torch.tensor(rhs, device=gpu)).cpu().numpy()
    assert rel(got, ref.cycle(0, p, rhs)) < 1e-12


@pytest.mark.parametrize('periodic', MODES)
@pytest.mark.parametrize('smoother', ['jacobi', 'rbgs'])
@pytest.mark.parametrize('backend', ['workspace', 'graph'])
def test_solve_history_matches_oracle(gpu, periodic, smoother, backend):
    shape, lengths = (128, 32), (4., 1.)
    grid, ch = channel_pair(shape, lengths, periodic)
    opts = MGOptions(smooth=2, smoother=smoother, check_every=1)
    mg = ChannelMultigrid(grid, device=gpu, options=opts, backend=backend)
    ref = mac2d.ChannelMultigrid(shape, ch.spacing, OOptions(smooth=2, smoother=smoother,
                                                             check_every=1), periodic)
    x = ch.coordinates()
    rhs = np.cos(np.pi * x[..., 0] / 2) * np.cos(2 * np.pi * x[..., 1]) + .1 * x[..., 1]
    if periodic:
        rhs = rhs - rhs.mean()
    hist, hist_ref = [], []
    p, info = mg.solve(torch.tensor(rhs, device=gpu), history=hist)
    p_ref, info_ref = ref.solve(rhs, history=hist_ref)
    assert info['cycles'] == info_ref['cycles']
    ratios = np.array(hist) / hist[0]
    ratios_ref = np.array(hist_ref) / hist_ref[0]
    assert np.abs(ratios - ratios_ref).max() < 1e-10
    assert rel(p.cpu().numpy(), p_ref) < 1e-9
    assert info['residual_norm'] <= info['tolerance']


def test_solve_errors_and_ownership(gpu):
    grid = ChannelGrid((32, 16), (2., 1.))
    mg = ChannelMultigrid(grid, device=gpu, options=MGOptions(smooth=2))
    with pytest.raises(ValueError, match='invalid channel pressure RHS'):
        mg.solve(torch.zeros((16, 16), device=gpu, dtype=torch.float64))
    bad = torch.zeros(grid.shape, device=gpu, dtype=torch.float64)
    bad[3, 3] = float('nan')
    with pytest.raises(ValueError, match='finite'):
        mg.solve(bad)
    rhs = torch.ones(grid.shape, device=gpu, dtype=torch.float64)
    p, _ = mg.solve(rhs)
    q, _ = mg.solve(2 * rhs)
    torch.testing.assert_close(q, 2 * p, rtol=1e-8, atol=1e-10)   # p not overwritten
    per = ChannelMultigrid(ChannelGrid((32, 16), (2., 1.), True), device=gpu)
    with pytest.raises(ValueError, match='incompatible'):
        per.solve(rhs)


@pytest.mark.parametrize('periodic', MODES)
def test_projection_matches_oracle(gpu, periodic):
    shape, lengths = (64, 32), (2., 1.)
    grid, ch = channel_pair(shape, lengths, periodic)
    gen = np.random.default_rng(5)
    u = gen.normal(size=ch.face_shape(0))
    v = gen.normal(size=ch.face_shape(1))
    v[:, 0] = v[:, -1] = 0
    opts = MGOptions(smooth=2, smoother='rbgs')
    flow = ChannelFlow(grid, dt=1e-3, device=gpu, options=opts)
    res = flow.project((torch.tensor(u, device=gpu), torch.tensor(v, device=gpu)))
    ref = mac2d.ChannelMultigrid(shape, ch.spacing, OOptions(smooth=2, smoother='rbgs'), periodic)
    (ru, rv), _, _ = mac2d.project((u, v), ch.spacing, ref, dt=1e-3)
    assert rel(res.velocity[0].cpu().numpy(), ru) < 1e-10
    assert rel(res.velocity[1].cpu().numpy(), rv) < 1e-10
    div = mac2d.divergence(tuple(a.cpu().numpy() for a in res.velocity), ch.spacing, periodic)
    assert np.linalg.norm(div) <= 1e-9 * np.linalg.norm(mac2d.divergence((u, v), ch.spacing, periodic))
    v[4, 0] = 1.
    with pytest.raises(ValueError, match='normal wall'):
        flow.project((torch.tensor(u, device=gpu), torch.tensor(v, device=gpu)))


def test_channel_step_matches_oracle(gpu):
    """256 x 64 on [0,4] x [0,1], periodic in x, walls in y; two curves of
    2,048 points; N(0,1) forces; red-black GS V(2,2) to 1e-10."""
    gen = np.random.default_rng(11)
    x = workloads.curves2d(gen)
    F = gen.normal(size=x.shape)
    grid, ch = channel_pair((256, 64), (4., 1.), True)
    opts = MGOptions(smooth=2, smoother='rbgs')
    stepper = ChannelIBStep(grid, dt=1e-3, device=gpu, options=opts)
    vel = grid.zeros(device=gpu)
    tx = torch.tensor(x, device=gpu)
    vel, p, U, info = stepper.step(tx, torch.tensor(F, device=gpu), vel)
    ref = ostep.ib_step_2d(x, F, tuple(np.zeros(ch.face_shape(c)) for c in range(2)), None, ch,
                           mac2d.ChannelMultigrid(ch.shape, ch.spacing,
                                                  OOptions(smooth=2, smoother='rbgs'), True),
                           dt=1e-3)
    assert len(mg_levels := stepper.flow.pressure_solver.shapes) == 5, mg_levels
    assert info['pressure']['cycles'] == ref['info']['cycles']
    assert rel(U.cpu().numpy(), ref['point_velocity']) < 1e-10
    for a, b in zip(vel, ref['velocity']):
        assert rel(a.cpu().numpy(), b) < 1e-10
    np.testing.assert_allclose(tx.cpu().numpy(), ref['x'], rtol=0, atol=1e-14)
