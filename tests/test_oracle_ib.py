"""Pin the FP64 transfer oracle (oracle/ib3d.py) without a GPU.

Checks are the reference's own known-answer tests (tests/test_ib.py:22-57,
tests/test_mac.py:17-39,149-175, tests/test_cuda_ib_demos.py:19-33 restated)
and independent mathematics: a dense brute-force H built from distances to
every face (no floor/base/index logic), conservation, adjointness and
moment identities.
"""
import numpy as np
import pytest

from oracle import ib3d
from oracle.ib3d import Grid3, peskin4


def test_peskin_moments_and_known_values():
    # tests/test_ib.py:22-32
    s = np.linspace(0, 1, 51)[:, None]
    nodes = np.arange(-2, 4, dtype=float)[None, :]
    w = peskin4(s - nodes)
    np.testing.assert_allclose(w.sum(1), 1., atol=2e-15, rtol=0)
    np.testing.assert_allclose((w * (nodes - s)).sum(1), 0., atol=2e-15, rtol=0)
    np.testing.assert_allclose((w * w).sum(1), 3 / 8, atol=2e-15, rtol=0)
    r = np.array([-3., -2., -1., 0., 1., 2., 3.])
    np.testing.assert_allclose(peskin4(r), [0., 0., .25, .5, .25, 0., 0.], atol=1e-15, rtol=0)
    assert (w >= 0).all()


def test_shared_radical_equals_two_branch_kernel():
    # The four weights of offsets 0..3 share sqrt(1+4f-4f^2); this is the
    # algebra the HIP kernels use (mac/_triton_prepare.py:35-49 documents it).
    f = np.linspace(0, 1, 1001, endpoint=False)
    R = np.sqrt(1 + 4 * f - 4 * f * f)
    shared = np.stack([(3 - 2 * f - R), (3 - 2 * f + R), (1 + 2 * f + R), (1 + 2 * f - R)], 1) / 8
    direct = peskin4(f[:, None] + 1 - np.arange(4)[None, :])
    np.testing.assert_allclose(shared, direct, rtol=0, atol=3e-16)


def shifted_grid(periodic=False):
    return Grid3((16, 20, 24), (4., 6., 8.), (-2., -3., -4.), periodic)


def random_points(gen, grid, n, margin=2.5):
    lo = np.asarray(grid.origin) + margin * np.asarray(grid.spacing)
    hi = np.asarray(grid.origin) + np.asarray(grid.lengths) - margin * np.asarray(grid.spacing)
    return gen.uniform(lo, hi, size=(n, 3))


@pytest.mark.parametrize('periodic', [False, True])
def test_compact_stencil_matches_dense_distance_operator(periodic):
    gen = np.random.default_rng(5)
    grid = shifted_grid(periodic)
    x = random_points(gen, grid, 7)
    if periodic:  # wrap across the box faces too
        x[0] = np.asarray(grid.origin) + 0.3 * np.asarray(grid.spacing)
        x[1] = np.asarray(grid.origin) + np.asarray(grid.lengths) - 0.2 * np.asarray(grid.spacing)
    base, phi = ib3d.prepare_compact(x, grid)
    for c in range(3):
        ids, w = ib3d.expanded_component(base, phi, grid, c)
        sparse = np.zeros((len(x), int(np.prod(grid.face_shape(c)))))
        for p in range(len(x)):
            np.add.at(sparse[p], ids[p], w[p])
        dense = ib3d.dense_operator(x, grid, c)
        np.testing.assert_allclose(sparse, dense, rtol=0, atol=1e-14)
        np.testing.assert_allclose(w.sum(1), 1., rtol=0, atol=1e-14)


def test_bases_are_floor_of_scaled_minus_one_on_face_lattices():
    # tests/test_mac.py:17-39: ids exact, z fastest, weights sum to one.
    grid = shifted_grid()
    gen = np.random.default_rng(1)
    x = random_points(gen, grid, 200)
    base, phi = ib3d.prepare_compact(x, grid)
    for c in range(3):
        o = np.asarray(grid.face_origin(c))
        h = np.asarray(grid.spacing)
        for p in range(0, 200, 37):
            for a in range(3):
                s = (x[p, a] - o[a]) / h[a]
                # the node base is the largest lattice index with s - base >= 1
                assert s - base[c, p, a] >= 1 and s - base[c, p, a] < 2
    assert base.dtype == np.int64


def test_spread_gather_conservation_adjointness_and_reproduction():
    grid = Grid3((12, 14, 16), (3., 3.5, 4.), (.5, -1., 2.))
    gen = np.random.default_rng(3)
    x = random_points(gen, grid, 300)
    F = gen.normal(size=(300, 3))
    base, phi = ib3d.prepare_compact(x, grid)
    f = ib3d.spread_grid(F, base, phi, grid)
    # force conservation
    np.testing.assert_allclose([fc.sum() * grid.volume for fc in f], F.sum(0), rtol=1e-12, atol=1e-12)
    # torque about the origin (tests/test_mac.py:149-175)
    torque = np.zeros(3)
    for c, fc in enumerate(f):
        X = grid.coordinates(c)
        vec = np.zeros(fc.shape + (3,))
        vec[..., c] = fc
        torque += np.cross(X, vec).sum((0, 1, 2)) * grid.volume
    np.testing.assert_allclose(torque, np.cross(x, F).sum(0), rtol=1e-11, atol=1e-11)
    # adjointness V <u, f> = <U, F>
    u = tuple(np.sin((c + 1) * grid.coordinates(c).sum(-1)) for c in range(3))
    U = ib3d.gather_grid(u, base, phi, grid)
    lhs = sum((uc * fc).sum() for uc, fc in zip(u, f)) * grid.volume
    np.testing.assert_allclose(lhs, (U * F).sum(), rtol=1e-12)
    # constant and affine fields are reproduced at the points
    A = np.array([[1., .3, -.2], [.7, -.8, .2], [0., .1, .5]])
    b = np.array([.2, -.4, .7])
    lin = tuple(grid.coordinates(c) @ A[c] + b[c] for c in range(3))
    np.testing.assert_allclose(ib3d.gather_grid(lin, base, phi, grid), x @ A.T + b, rtol=1e-13, atol=1e-13)


def test_spread_and_gather_equal_dense_operator():
    grid = Grid3((8, 9, 10), (1., 1.2, 1.4), periodic=True)
    gen = np.random.default_rng(11)
    x = gen.uniform(0, 1, size=(25, 3)) * np.asarray(grid.lengths)
    F = gen.normal(size=(25, 3))
    base, phi = ib3d.prepare_compact(x, grid)
    f = ib3d.spread_grid(F, base, phi, grid)
    u = tuple(gen.normal(size=grid.face_shape(c)) for c in range(3))
    U = ib3d.gather_grid(u, base, phi, grid)
    for c in range(3):
        H = ib3d.dense_operator(x, grid, c)
        np.testing.assert_allclose(f[c].reshape(-1), H.T @ F[:, c] / grid.volume, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(U[:, c], H @ u[c].reshape(-1), rtol=1e-12, atol=1e-12)


def test_support_predicate():
    grid = Grid3((10, 10, 10), (1., 1., 1.))
    assert ib3d.support_ok(np.array([[.25, .5, .5]]), grid)
    assert not ib3d.support_ok(np.array([[.19, .5, .5]]), grid)
    assert not ib3d.support_ok(np.array([[.5, .8, .5]]), grid)   # s = 8 = n-2
    assert not ib3d.support_ok(np.array([[np.nan, .5, .5]]), grid)
    assert ib3d.support_ok(np.array([[-5., .5, .5]]), Grid3((10, 10, 10), (1., 1., 1.), periodic=True))


def nodal_setup():
    # tests/test_ib.py:14-19
    shape, spacing, origin = (11, 10, 9), (.2, .3, .4), (-.7, 1.1, -.3)
    idx = np.array([[2.17, 3.41, 2.53], [5.32, 4.29, 3.78], [6., 2., 5.], [2.17, 3.41, 2.53]])
    x = idx * np.asarray(spacing) + np.asarray(origin)
    k, j, i = np.meshgrid(np.arange(shape[2]), np.arange(shape[1]), np.arange(shape[0]), indexing='ij')
    X = np.stack((i, j, k), -1).reshape(-1, 3) * np.asarray(spacing) + np.asarray(origin)
    ids, w = ib3d.nodal_prepare(x, shape, spacing, origin)
    return shape, spacing, x, X, ids, w


def test_nodal_lattice_constant_affine_and_conservation():
    # tests/test_ib.py:34-57
    shape, spacing, x, X, ids, w = nodal_setup()
    A = np.array([[1., .3, -.2], [.7, -.8, .2], [0., .1, .5]])
    b = np.array([.2, -.4, .7])
    np.testing.assert_allclose(ib3d.indexed_gather(np.broadcast_to(b, X.shape), ids, w),
                               np.broadcast_to(b, x.shape), atol=1e-14, rtol=1e-13)
    np.testing.assert_allclose(ib3d.indexed_gather(X @ A.T + b, ids, w), x @ A.T + b, atol=1e-14, rtol=1e-13)
    vol = np.prod(spacing)
    u = np.sin(np.arange(X.size, dtype=float)).reshape(X.shape)
    g = np.cos(np.arange(x.size, dtype=float)).reshape(x.shape)
    f = ib3d.indexed_spread(g, ids, w, X.shape[0], vol)
    np.testing.assert_allclose(f.sum(0) * vol, g.sum(0), atol=1e-13, rtol=1e-12)
    np.testing.assert_allclose(np.cross(X, f).sum(0) * vol, np.cross(x, g).sum(0), atol=1e-13, rtol=1e-12)
    np.testing.assert_allclose((u * f).sum() * vol, (ib3d.indexed_gather(u, ids, w) * g).sum(), atol=1e-13, rtol=1e-12)
    with pytest.raises(ValueError, match='support'):
        ib3d.nodal_prepare(x + 100, shape, spacing, (-.7, 1.1, -.3))


@pytest.mark.parametrize('neighbors,components', [(16, 1), (64, 3), (37, 2)])
def test_indexed_signed_duplicate_weights_match_brute_force(neighbors, components):
    # tests/test_cuda_ib_demos.py:19-33 restated against np.add.at
    gen = np.random.default_rng(31)
    ids = gen.integers(0, 53, size=(19, neighbors))
    w = gen.normal(size=ids.shape)
    u = gen.normal(size=(53, components))
    f = gen.normal(size=(19, components))
    v = .017
    expected = np.zeros((19, components))
    for p in range(19):
        for k in range(neighbors):
            expected[p] += u[ids[p, k]] * w[p, k]
    np.testing.assert_allclose(ib3d.indexed_gather(u, ids, w), expected, rtol=1e-12, atol=1e-12)
    ref = np.zeros_like(u)
    np.add.at(ref, ids.reshape(-1), (f[:, None] * w[..., None] / v).reshape(-1, components))
    np.testing.assert_allclose(ib3d.indexed_spread(f, ids, w, 53, v), ref, rtol=1e-12, atol=1e-11)
