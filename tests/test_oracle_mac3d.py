"""Pin the 3D MAC / multigrid oracle (oracle/mac3d.py) without a GPU.

Reference KATs restated: scalar-loop Neumann Laplacian and A=-DG
(tests/test_mac.py:43-61), A.const=0 exactly (tests/test_mac_multigrid_fused
.py:40-42), manufactured cosine solve and projection (tests/test_mac.py:
65-83,103-112). Independent mathematics: dense-matrix Jacobi / Gauss-Seidel,
an axis-by-axis trilinear interpolation with PyTorch's align_corners=False
source-index rule, FFT diagonalisation of the periodic operator and direct
dense solves.
"""
from math import pi
import numpy as np
import pytest

from oracle import mac3d
from oracle.mac3d import MGOptions, Multigrid


def cell_coords(shape, lengths):
    h = [L / n for L, n in zip(lengths, shape)]
    axes = [(np.arange(n) + .5) * hh for n, hh in zip(shape, h)]
    return np.stack(np.meshgrid(*axes, indexing='ij'), -1), tuple(h)


@pytest.mark.parametrize('periodic', [False, True])
def test_laplacian_matches_scalar_loop(periodic):
    shape, lengths = (8, 6, 10), (2., 3., 4.)
    xyz, h = cell_coords(shape, lengths)
    p = np.sin(xyz[..., 0] + .7 * xyz[..., 1] - .3 * xyz[..., 2])
    expected = np.zeros_like(p)
    for idx in np.ndindex(p.shape):
        for axis, hh in enumerate(h):
            for side in (-1, 1):
                nb = list(idx)
                nb[axis] += side
                if periodic:
                    nb[axis] %= shape[axis]
                elif not 0 <= nb[axis] < shape[axis]:
                    continue
                expected[idx] += (p[idx] - p[tuple(nb)]) / hh ** 2
    np.testing.assert_allclose(mac3d.negative_laplacian(p, h, periodic), expected, atol=2e-13, rtol=0)
    # A = -D G and summation by parts
    np.testing.assert_allclose(mac3d.negative_laplacian(p, h, periodic),
                               -mac3d.divergence(mac3d.gradient(p, h, periodic), h, periodic),
                               atol=1e-12, rtol=1e-12)
    const = np.full(shape, 7.)
    assert np.array_equal(mac3d.negative_laplacian(const, h, periodic), np.zeros(shape))


def test_diagonal_is_matrix_diagonal():
    shape, h = (4, 6, 8), (.5, .3, .2)
    for periodic in (False, True):
        A = mac3d.dense_matrix(shape, h, periodic)
        np.testing.assert_allclose(mac3d.diagonal(shape, h, periodic).reshape(-1), np.diag(A), rtol=1e-15)


def trilinear_align_corners_false(coarse, fine_shape, periodic):
    """Independent axis-by-axis linear interpolation: source index
    s=(i+0.5)/2-0.5; Neumann clamps s to [0,n-1], periodic wraps."""
    out = coarse
    for axis, n_f in enumerate(fine_shape):
        n_c = out.shape[axis]
        s = (np.arange(n_f) + .5) / 2 - .5
        if periodic:
            i0 = np.floor(s).astype(int)
            lam = s - i0
            a, b = np.take(out, i0 % n_c, axis), np.take(out, (i0 + 1) % n_c, axis)
        else:
            s = np.clip(s, 0, n_c - 1)
            i0 = np.floor(s).astype(int)
            i1 = np.minimum(i0 + 1, n_c - 1)
            lam = s - i0
            a, b = np.take(out, i0, axis), np.take(out, i1, axis)
        shape = [1, 1, 1]
        shape[axis] = -1
        lam = lam.reshape(shape)
        out = a * (1 - lam) + b * lam
    return out


@pytest.mark.parametrize('periodic', [False, True])
def test_prolongation_and_restriction(periodic):
    gen = np.random.default_rng(2)
    coarse = gen.normal(size=(3, 5, 7))
    fine_shape = (6, 10, 14)
    np.testing.assert_allclose(mac3d.prolong(coarse, fine_shape, periodic),
                               trilinear_align_corners_false(coarse, fine_shape, periodic),
                               rtol=0, atol=2e-15)
    fine = gen.normal(size=fine_shape)
    mean8 = fine.reshape(3, 2, 5, 2, 7, 2).mean((1, 3, 5))
    np.testing.assert_allclose(mac3d.restrict(fine), mean8, rtol=0, atol=4e-16)


def test_hierarchy_rule():
    assert mac3d.hierarchy((64,) * 3, (1 / 64,) * 3)[0][-1] == (4, 4, 4)
    assert len(mac3d.hierarchy((64,) * 3, (1 / 64,) * 3)[0]) == 5
    assert len(mac3d.hierarchy((128,) * 3, (1 / 128,) * 3)[0]) == 6
    assert mac3d.hierarchy((48,) * 3, (1 / 48,) * 3)[0][-1] == (3, 3, 3)
    assert mac3d.hierarchy((6, 10, 14), (1., 1., 1.))[0] == [(6, 10, 14), (3, 5, 7)]
    with pytest.raises(ValueError):
        mac3d.hierarchy((18, 18, 18), (1., 1., 1.))  # coarsest 9^3 = 729 > 512


@pytest.mark.parametrize('periodic', [False, True])
def test_sweeps_match_dense_matrix_restatement(periodic):
    shape, h = (4, 6, 8), (.5, .4, .3)
    mg = Multigrid(shape, h, MGOptions(smooth=1), periodic)
    A = mac3d.dense_matrix(shape, h, periodic)
    gen = np.random.default_rng(7)
    p, b = gen.normal(size=shape), gen.normal(size=shape)
    pv, bv = p.reshape(-1), b.reshape(-1)
    d = np.diag(A)
    jac = pv + (2 / 3) * (bv - A @ pv) / d
    np.testing.assert_allclose(mg.jacobi(0, p, b).reshape(-1), jac, rtol=1e-12, atol=1e-12)
    # Gauss-Seidel, all red cells in index order, then all black cells.
    gs = pv.copy()
    parity = np.indices(shape).sum(0).reshape(-1) % 2
    for color in (0, 1):
        for i in np.nonzero(parity == color)[0]:
            gs[i] = (bv[i] - (A[i] @ gs - A[i, i] * gs[i])) / A[i, i]
    np.testing.assert_allclose(mg.rbgs(0, p, b).reshape(-1), gs, rtol=1e-12, atol=1e-12)
    # Over-relaxed: x_i <- (1 - w) x_i + w * (Gauss-Seidel value), cell by cell
    # in the same order (textbook SOR on the dense matrix).
    w = 1.3
    sor = pv.copy()
    for color in (0, 1):
        for i in np.nonzero(parity == color)[0]:
            sor[i] = (1 - w) * sor[i] + w * (bv[i] - (A[i] @ sor - A[i, i] * sor[i])) / A[i, i]
    relaxed = Multigrid(shape, h, MGOptions(smooth=1, smoother='rbgs', relax=w), periodic)
    np.testing.assert_allclose(relaxed.rbgs(0, p, b).reshape(-1), sor, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize('periodic', [False, True])
def test_over_relaxed_v_cycle_converges_faster_to_the_same_solution(periodic):
    """relax 1.3 keeps the V-cycle's fixed point (the dense direct solution)
    and lowers its cycle count at rtol 1e-10 (10 -> 6 at 32^3 here)."""
    shape = (32, 32, 32)
    h = tuple(1 / n for n in shape)
    rhs = np.random.default_rng(0).normal(size=shape)
    rhs -= rhs.mean()
    runs = {}
    for w in (1., 1.3):
        mg = Multigrid(shape, h, MGOptions(smooth=2, smoother='rbgs', relax=w), periodic)
        runs[w] = mg.solve(rhs)
    (p1, i1), (p2, i2) = runs[1.], runs[1.3]
    assert i2['cycles'] < i1['cycles']
    np.testing.assert_allclose(p2, p1, rtol=0, atol=1e-8 * np.abs(p1).max())


@pytest.mark.parametrize('periodic', [False, True])
def test_coarse_inverse_solves_mean_free_systems(periodic):
    mg = Multigrid((8, 8, 8), (.25, .25, .25), periodic=periodic)
    A = mac3d.dense_matrix(mg.shapes[-1], mg.spacings[-1], periodic)
    b = np.random.default_rng(0).normal(size=A.shape[0])
    b -= b.mean()
    np.testing.assert_allclose(A @ (mg.coarse_inverse @ b), b, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize('smoother', ['jacobi', 'rbgs'])
def test_manufactured_cosine_solve(smoother):
    # tests/test_mac.py:65-75
    shape, lengths = (16, 16, 16), (2., 2., 2.)
    xyz, h = cell_coords(shape, lengths)
    exact = np.cos(pi * xyz[..., 0] / 2) * np.cos(2 * pi * xyz[..., 1] / 2) * np.cos(pi * xyz[..., 2] / 2)
    rhs = mac3d.negative_laplacian(exact, h)
    mg = Multigrid(shape, h, MGOptions(smoother=smoother))
    p, info = mg.solve(rhs)
    np.testing.assert_allclose(p, exact, rtol=2e-8, atol=2e-9)
    assert info['residual_norm'] <= info['tolerance'] and info['cycles'] < 80
    with pytest.raises(ValueError, match='incompatible'):
        mg.solve(np.ones_like(rhs))
    with pytest.raises(ValueError, match='invalid pressure RHS'):
        mg.solve(np.full_like(rhs, np.nan))


@pytest.mark.parametrize('shape', [(4, 4, 4), (8, 12, 16)])
def test_solve_matches_dense_direct_solve(shape):
    h = (2 / shape[0], 3 / shape[1], 4 / shape[2])
    gen = np.random.default_rng(4)
    rhs = gen.normal(size=shape)
    rhs -= rhs.mean()
    mg = Multigrid(shape, h, MGOptions(smooth=3))
    p, _ = mg.solve(rhs)
    A = mac3d.dense_matrix(shape, h)
    direct = np.linalg.lstsq(A, rhs.reshape(-1), rcond=None)[0]
    direct -= direct.mean()
    np.testing.assert_allclose(p.reshape(-1), direct, rtol=1e-8, atol=1e-8 * np.abs(direct).max())


@pytest.mark.parametrize('smoother', ['jacobi', 'rbgs'])
def test_periodic_solve_matches_fft(smoother):
    shape, lengths = (16, 16, 16), (1., 1., 1.)
    _, h = cell_coords(shape, lengths)
    rhs = np.random.default_rng(9).normal(size=shape)
    rhs -= rhs.mean()
    mg = Multigrid(shape, h, MGOptions(smooth=2, smoother=smoother), periodic=True)
    p, info = mg.solve(rhs)
    exact = mac3d.fft_poisson_periodic(rhs, h)
    scale = np.abs(exact).max()
    np.testing.assert_allclose(p, exact, rtol=0, atol=1e-8 * scale)
    assert info['cycles'] <= 30


@pytest.mark.parametrize('periodic', [False, True])
def test_projection_divergence_idempotence_and_gradient_force(periodic):
    shape, lengths = (16, 16, 16), (2., 3., 4.)
    xyz, h = cell_coords(shape, lengths)
    mg = Multigrid(shape, h, periodic=periodic)
    faces = [tuple(n + (a == c and not periodic) for a, n in enumerate(shape)) for c in range(3)]
    gen = np.random.default_rng(8)
    u = [gen.normal(size=s) for s in faces]
    if not periodic:
        for c in range(3):
            u[c][mac3d.slab(c, 0, 1)] = 0
            u[c][mac3d.slab(c, -1, None)] = 0
    vel, p, info = mac3d.project(tuple(u), h, mg, dt=1e-4)
    div = mac3d.divergence(vel, h, periodic)
    assert np.abs(div).max() < 1e-8 * np.abs(mac3d.divergence(tuple(u), h, periodic)).max()
    vel2, _, _ = mac3d.project(vel, h, mg, dt=1e-4)
    for a, b in zip(vel, vel2):
        np.testing.assert_allclose(a, b, rtol=1e-7, atol=1e-8)
    # gradient field: projection removes it (tests/test_mac.py:103-112)
    phi = np.cos(pi * xyz[..., 0] / 2) * np.cos(pi * xyz[..., 2] / 2)
    grad = mac3d.gradient(phi, h, periodic)
    vel3, _, _ = mac3d.project(grad, h, mg, dt=1.)
    assert max(np.abs(v).max() for v in vel3) < 1e-8


def face_coordinates(shape, lengths, c):
    """Closed-box face centres of component c (n+1 along c, cell centres across)."""
    h = [L / n for L, n in zip(lengths, shape)]
    axes = [np.arange(shape[a] + 1) * h[a] if a == c else (np.arange(shape[a]) + .5) * h[a] for a in range(3)]
    return np.stack(np.meshgrid(*axes, indexing='ij'), -1)


def test_predictor_known_answers():
    """tests/test_mac.py:87-100: the sine eigenmode of the velocity Laplacian
    (eigenvalue sum -4 sin^2(pi/2n)/h^2) and the centered momentum flux of the
    linear field u_c = x_c (4 u in the interior)."""
    shape, lengths = (8, 8, 8), (1., 1., 1.)
    h = [L / n for L, n in zip(lengths, shape)]
    eig = sum(-4 * np.sin(pi / (2 * n)) ** 2 / hh ** 2 for n, hh in zip(shape, h))
    for c in range(3):
        xyz = face_coordinates(shape, lengths, c)
        u = np.sin(pi * xyz).prod(-1)
        fields = [np.zeros(face_coordinates(shape, lengths, a).shape[:3]) for a in range(3)]
        fields[c] = u
        u = mac3d.zero_normal(fields)[c]
        np.testing.assert_allclose(mac3d.velocity_laplacian(u, c, h), eig * u, rtol=1e-11, atol=2e-12)
    velocity = tuple(face_coordinates(shape, lengths, c)[..., c] for c in range(3))
    for u, a in zip(velocity, mac3d.convection(velocity, h)):
        np.testing.assert_allclose(a[2:-2, 2:-2, 2:-2], 4 * u[2:-2, 2:-2, 2:-2], atol=1e-12, rtol=1e-12)


def test_gradient_body_force_projects_to_rest():
    """tests/test_mac.py:103-113: a gradient body force leaves zero velocity
    and its potential as the pressure; too large a viscous number raises."""
    shape, lengths = (16, 16, 16), (2., 2., 2.)
    h = tuple(L / n for L, n in zip(lengths, shape))
    x = [(np.arange(n) + .5) * hh for n, hh in zip(shape, h)]
    X, Y, Z = np.meshgrid(*x, indexing='ij')
    p = np.cos(pi * X / 2) * np.cos(pi * Z / 2)
    mg = mac3d.Multigrid(shape, h, mac3d.MGOptions())
    zeros = tuple(np.zeros(face_coordinates(shape, lengths, c).shape[:3]) for c in range(3))
    u, q, _ = mac3d.flow_step(zeros, mac3d.gradient(p, h), h, mg, dt=1e-4)
    np.testing.assert_allclose(q - q.mean(), p - p.mean(), rtol=1e-8, atol=1e-9)
    assert max(np.abs(v).max() for v in u) < 1e-11
    with pytest.raises(ValueError, match='guard'):
        mac3d.flow_step(zeros, mac3d.gradient(p, h), h, mg, dt=1.)
