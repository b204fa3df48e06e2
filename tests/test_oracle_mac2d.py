"""Pins the 2D channel oracle (reference BCs and the periodic-x mode).

Checks that do not retype the oracle's formulas: the reference's own
channel tests (tests/test_valve_mac.py:60-118: A = -D G, solve of A p,
force/torque/power, wall reflection adjointness), scalar loops, dense
matrices, scipy's linear spline sampling of the ghost-padded coarse grid,
and a brute-force transfer operator from distances to every face and its
mirror images.
"""
from math import pi
import numpy as np
import pytest
from scipy import ndimage

from oracle import mac2d
from oracle.mac3d import MGOptions

MODES = [False, True]


def scalar_laplacian(p, spacing, periodic):
    hx, hy = spacing
    nx, ny = p.shape
    out = np.zeros_like(p)
    for i in range(nx):
        for j in range(ny):
            a = 0.
            if periodic or i + 1 < nx:
                a -= (p[(i + 1) % nx, j] - p[i, j]) / hx ** 2
            if periodic or i > 0:
                a += (p[i, j] - p[(i - 1) % nx, j]) / hx ** 2
            if j + 1 < ny:
                a -= (p[i, j + 1] - p[i, j]) / hy ** 2
            if j > 0:
                a += (p[i, j] - p[i, j - 1]) / hy ** 2
            if not periodic and i == nx - 1:
                a += 2 * p[i, j] / hx ** 2
            out[i, j] = a
    return out


def field(channel, seed=0):
    x = channel.coordinates()
    return np.sin(.7 * x.sum(-1)) + .4 + 0 * seed


@pytest.mark.parametrize('periodic', MODES)
def test_operator_scalar_loop_and_minus_div_grad(periodic):
    ch = mac2d.Channel((32, 16), periodic=periodic)
    p = field(ch)
    A = mac2d.negative_laplacian(p, ch.spacing, periodic)
    np.testing.assert_array_equal(A, scalar_laplacian(p, ch.spacing, periodic))
    # tests/test_valve_mac.py:62-64
    np.testing.assert_allclose(A, -mac2d.divergence(mac2d.gradient(p, ch.spacing, periodic),
                                                    ch.spacing, periodic), rtol=2e-13, atol=2e-13)
    dense = mac2d.dense_matrix(ch.shape, ch.spacing, periodic)
    np.testing.assert_allclose(dense @ p.reshape(-1), A.reshape(-1), rtol=1e-13, atol=1e-11)
    np.testing.assert_array_equal(np.diag(dense).reshape(ch.shape),
                                  mac2d.diagonal(ch.shape, ch.spacing, periodic))
    # symmetry; the periodic operator annihilates constants, the channel's does not
    np.testing.assert_allclose(dense, dense.T, rtol=0, atol=1e-9)
    const = np.abs(mac2d.negative_laplacian(np.ones(ch.shape), ch.spacing, periodic)).max()
    assert (const == 0) if periodic else (const > 0)


def test_outlet_gradient_and_shapes():
    ch = mac2d.Channel((8, 4))
    p = np.arange(32.).reshape(8, 4)
    gx, gy = mac2d.gradient(p, ch.spacing)
    assert gx.shape == ch.face_shape(0) and gy.shape == ch.face_shape(1)
    np.testing.assert_array_equal(gx[-1], -2 * p[-1] / ch.spacing[0])
    assert not gx[0].any() and not gy[:, 0].any() and not gy[:, -1].any()
    chp = mac2d.Channel((8, 4), periodic=True)
    gx, _ = mac2d.gradient(p, chp.spacing, True)
    assert gx.shape == chp.face_shape(0) == (8, 4)
    np.testing.assert_allclose(gx[0], (p[0] - p[-1]) / chp.spacing[0])


def test_restriction_is_the_mean_of_four_children():
    fine = np.random.default_rng(1).normal(size=(16, 8))
    expected = fine.reshape(8, 2, 4, 2).mean(axis=(1, 3))
    np.testing.assert_allclose(mac2d.restrict(fine), expected, rtol=1e-15, atol=1e-15)


@pytest.mark.parametrize('periodic', MODES)
def test_prolongation_is_linear_spline_of_ghost_padded_grid(periodic):
    coarse = np.random.default_rng(2).normal(size=(8, 5))
    if periodic:
        padded = np.concatenate((coarse[-1:], coarse, coarse[:1]), 0)
    else:
        padded = np.concatenate((coarse[:1], coarse, -coarse[-1:]), 0)
    padded = np.concatenate((padded[:, :1], padded, padded[:, -1:]), 1)
    fi = np.arange(16)[:, None] * np.ones((1, 10))
    fj = np.ones((16, 1)) * np.arange(10)[None]
    # align_corners=False: fine cell i sits at coarse coordinate (i+.5)/2-.5
    coords = np.stack([(fi + .5) / 2 - .5 + 1, (fj + .5) / 2 - .5 + 1])
    expected = ndimage.map_coordinates(padded, coords, order=1, mode='nearest')
    np.testing.assert_allclose(mac2d.prolong(coarse, (16, 10), periodic), expected,
                               rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize('periodic', MODES)
@pytest.mark.parametrize('smoother', ['jacobi', 'rbgs'])
def test_smoothers_match_dense_restatements(periodic, smoother):
    shape = (16, 8)
    ch = mac2d.Channel(shape, (2., 1.), periodic=periodic)
    mg = mac2d.ChannelMultigrid(shape, ch.spacing, MGOptions(smooth=1, smoother=smoother), periodic)
    A = mac2d.dense_matrix(shape, ch.spacing, periodic)
    d = np.diag(A)
    gen = np.random.default_rng(3)
    p, b = gen.normal(size=shape).reshape(-1), gen.normal(size=shape).reshape(-1)
    if smoother == 'jacobi':
        expected = p + (2 / 3) * (b - A @ p) / d
        got = mg.jacobi(0, p.reshape(shape), b.reshape(shape))
    else:
        expected = p.copy()
        parity = (np.arange(shape[0])[:, None] + np.arange(shape[1])[None]).reshape(-1) % 2
        for color in (0, 1):
            for k in np.flatnonzero(parity == color):       # sequential GS in colour order
                expected[k] += (b[k] - A[k] @ expected) / d[k]
        got = mg.rbgs(0, p.reshape(shape), b.reshape(shape))
    np.testing.assert_allclose(got.reshape(-1), expected, rtol=1e-12, atol=1e-10)


def test_hierarchy_and_coarse_inverse():
    shapes, spacings = mac2d.hierarchy((256, 64), (4 / 256, 1 / 64))
    assert shapes == [(256, 64), (128, 32), (64, 16), (32, 8), (16, 4)]
    with pytest.raises(ValueError, match='512'):
        mac2d.hierarchy((90, 30), (1., 1.))
    for periodic in MODES:
        mg = mac2d.ChannelMultigrid((32, 8), (1 / 32, 1 / 8), periodic=periodic)
        A = mac2d.dense_matrix(mg.shapes[-1], mg.spacings[-1], periodic)
        b = np.random.default_rng(4).normal(size=A.shape[0])
        if periodic:
            b -= b.mean()
        x = mg.inverse @ b
        np.testing.assert_allclose(A @ x, b, rtol=1e-10, atol=1e-10)
        if periodic:
            assert abs(x.mean()) < 1e-12


@pytest.mark.parametrize('periodic', MODES)
@pytest.mark.parametrize('smoother', ['jacobi', 'rbgs'])
def test_solve_recovers_field(periodic, smoother):
    # tests/test_valve_mac.py:60-68 for the channel; zero-mean field when periodic
    ch = mac2d.Channel((32, 16), periodic=periodic)
    p = field(ch)
    if periodic:
        p = p - p.mean()
    A = mac2d.negative_laplacian(p, ch.spacing, periodic)
    mg = mac2d.ChannelMultigrid(ch.shape, ch.spacing, MGOptions(smoother=smoother), periodic)
    history = []
    result, info = mg.solve(A, history=history)
    assert info['residual_norm'] <= info['tolerance']
    np.testing.assert_allclose(result, p, rtol=1e-8, atol=1e-8)
    dense = np.linalg.solve(mac2d.dense_matrix(ch.shape, ch.spacing, periodic)
                            + (np.ones((512, 512)) / 512 if periodic else 0), A.reshape(-1))
    np.testing.assert_allclose(result.reshape(-1), dense, rtol=1e-8, atol=1e-8)
    assert all(b < a for a, b in zip(history, history[1:]))


def test_periodic_solve_rejects_incompatible_rhs():
    mg = mac2d.ChannelMultigrid((16, 8), (1 / 16, 1 / 8), periodic=True)
    with pytest.raises(ValueError, match='incompatible'):
        mg.solve(np.ones((16, 8)))


@pytest.mark.parametrize('periodic', MODES)
def test_projection_divergence_idempotence_and_gradient_fields(periodic):
    ch = mac2d.Channel((32, 16), (2., 1.), periodic=periodic)
    gen = np.random.default_rng(5)
    u = gen.normal(size=ch.face_shape(0))
    v = gen.normal(size=ch.face_shape(1))
    v[:, 0] = v[:, -1] = 0
    mg = mac2d.ChannelMultigrid(ch.shape, ch.spacing, MGOptions(rtol=1e-12, atol=1e-14), periodic)
    (pu, pv), p, _ = mac2d.project((u, v), ch.spacing, mg, dt=1e-3)
    div0 = np.linalg.norm(mac2d.divergence((u, v), ch.spacing, periodic))
    assert np.linalg.norm(mac2d.divergence((pu, pv), ch.spacing, periodic)) <= 1e-8 * div0
    assert not pv[:, 0].any() and not pv[:, -1].any()
    (qu, qv), _, _ = mac2d.project((pu, pv), ch.spacing, mg, dt=1e-3)
    np.testing.assert_allclose(qu, pu, atol=1e-9)
    np.testing.assert_allclose(qv, pv, atol=1e-9)
    # a discrete gradient field is removed entirely (tests/test_mac.py:103-112 in 2D)
    phi = field(ch)
    if periodic:
        phi -= phi.mean()
    gx, gy = mac2d.gradient(phi, ch.spacing, periodic)
    (ru, rv), p, _ = mac2d.project((gx, gy), ch.spacing, mg, dt=1e-3)
    assert np.abs(ru).max() < 1e-7 and np.abs(rv).max() < 1e-7
    np.testing.assert_allclose(p, phi * 1e3, rtol=1e-7, atol=1e-7)
    with pytest.raises(ValueError, match='normal wall'):
        v[3, 0] = 1.
        mac2d.project((u, v), ch.spacing, mg, dt=1e-3)


def wall_points(ch, count, seed):
    gen = np.random.default_rng(seed)
    lo, hi = (0., ch.lengths[0]) if ch.periodic else (2.2 * ch.spacing[0], ch.lengths[0] - 2.2 * ch.spacing[0])
    x = gen.uniform(lo, hi, count)
    # half the points within two cells of a wall, some beyond it (allowed to -hy / Ly+hy)
    y = np.where(gen.random(count) < .5, gen.uniform(-.9, 2., count) * ch.spacing[1],
                 gen.uniform(0, ch.lengths[1], count))
    y = np.where(gen.random(count) < .3, ch.lengths[1] - y + 0 * y, y)
    return np.stack([x, y], 1)


@pytest.mark.parametrize('periodic', MODES)
def test_prepare_matches_brute_force_mirror_operator(periodic):
    ch = mac2d.Channel((24, 12), (3., 1.), periodic=periodic)
    pts = wall_points(ch, 300, 6)
    assert mac2d.support_ok(pts, ch)
    ids, w = mac2d.prepare(pts, ch)
    dense = mac2d.dense_transfer(pts, ch)
    assert any((wc < 0).any() for wc in w)
    for c in range(2):
        H = np.zeros_like(dense[c])
        np.add.at(H, (np.repeat(np.arange(len(pts)), 16), ids[c].reshape(-1)), w[c].reshape(-1))
        np.testing.assert_allclose(H, dense[c], rtol=0, atol=1e-14)


@pytest.mark.parametrize('periodic', MODES)
def test_transfer_force_torque_power_and_adjointness(periodic):
    # tests/test_valve_mac.py:83-118 with point forces given directly
    ch = mac2d.Channel((64, 32), (8., 4.), periodic=periodic)
    gen = np.random.default_rng(7)
    pts = np.stack([gen.uniform(2., 6., 500), gen.uniform(1., 3., 500)], 1)
    force = np.sin(3 * pts)
    ids, w = mac2d.prepare(pts, ch)
    density = mac2d.spread(force, ids, w, ch)
    np.testing.assert_allclose([d.sum() * ch.volume for d in density], force.sum(0), rtol=1e-10)
    torque = ch.volume * ((ch.coordinates(1)[..., 0] * density[1]).sum()
                          - (ch.coordinates(0)[..., 1] * density[0]).sum())
    np.testing.assert_allclose(torque, (pts[:, 0] * force[:, 1] - pts[:, 1] * force[:, 0]).sum(),
                               rtol=1e-9)
    velocity = tuple(np.sin((c + 1) * ch.coordinates(c).sum(-1)) for c in range(2))
    U = mac2d.gather(velocity, ids, w)
    np.testing.assert_allclose((U * force).sum(),
                               ch.volume * sum((u * f).sum() for u, f in zip(velocity, density)),
                               rtol=1e-12)
    U = mac2d.gather(tuple(np.full(ch.face_shape(c), c + 1.) for c in range(2)), ids, w)
    np.testing.assert_allclose(U, np.broadcast_to([1., 2.], U.shape), rtol=1e-13)
    # wall reflection: points near the walls, no-slip velocity, adjointness
    pts = wall_points(ch, 400, 8)
    ids, w = mac2d.prepare(pts, ch)
    force = np.sin(5 * pts)
    density = mac2d.spread(force, ids, w, ch)
    v = velocity[1].copy()
    v[:, 0] = v[:, -1] = 0
    U = mac2d.gather((velocity[0], v), ids, w)
    np.testing.assert_allclose((U * force).sum(),
                               ch.volume * ((velocity[0] * density[0]).sum() + (v * density[1]).sum()),
                               rtol=1e-12)


def test_support_rejection():
    ch = mac2d.Channel((32, 16), (4., 1.))
    hx, hy = ch.spacing
    ok = np.array([[2 * hx, .5], [4 - 2 * hx, -hy], [1., 1 + hy]])
    assert mac2d.support_ok(ok, ch)
    for bad in ([1.9 * hx, .5], [1., -1.01 * hy], [1., np.nan]):
        assert not mac2d.support_ok(np.array([bad]), ch)
    chp = mac2d.Channel((32, 16), (4., 1.), periodic=True)
    assert mac2d.support_ok(np.array([[0., .5], [3.99, .5]]), chp)


def test_manufactured_channel_solution_converges():
    # smooth exact pressure with p = 0 half a cell beyond the outlet and
    # zero normal derivative on the inlet and walls: second-order error
    errs = []
    for n in (16, 32):
        ch = mac2d.Channel((2 * n, n), (2., 1.))
        x = ch.coordinates()
        exact = np.cos(pi * x[..., 0] / 4) * np.cos(pi * x[..., 1])
        # continuous -Laplacian of the exact field
        rhs = (pi ** 2 / 16 + pi ** 2) * exact
        mg = mac2d.ChannelMultigrid(ch.shape, ch.spacing, MGOptions(rtol=1e-12))
        p, _ = mg.solve(rhs)
        errs.append(np.abs(p - exact).max())
    assert errs[1] < errs[0] / 3.5
