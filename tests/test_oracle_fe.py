"""The FE interaction-point oracle (oracle/fe.py) pinned by closed forms:
exact P1 mass matrices, exact evaluation of affine fields, adjointness of
the assembly, M = B^T W B, force conservation, and the PCG against dense
solves and its reference error paths (fluid/solvers.py:39-122)."""
import numpy as np
import pytest

from oracle import fe


def sphere():
    nodes, tris = fe.icosphere(2, .3, (.1, -.2, .05))
    return fe.p1_surface(nodes, tris), nodes


def solid():
    nodes, tets = fe.box_tets(3, (-.4, -.3, -.5), (.5, .4, .3))
    return fe.p1_solid(nodes, tets), nodes


def closed_form_mass(g, nodes, dim):
    """P1 mass: |T|/12 (1 + delta_ab) for triangles, |T|/20 (1 + delta_ab) for tetrahedra."""
    M = np.zeros((g.node_count, g.node_count))
    A = dim + 1
    for cell in g.cells:
        v = nodes[cell]
        if dim == 2:
            size = .5 * np.linalg.norm(np.cross(v[1] - v[0], v[2] - v[0]))
            local = size / 12 * (np.ones((A, A)) + np.eye(A))
        else:
            size = abs(np.dot(v[1] - v[0], np.cross(v[2] - v[0], v[3] - v[0]))) / 6
            local = size / 20 * (np.ones((A, A)) + np.eye(A))
        for a in range(A):
            for b in range(A):
                M[cell[a], cell[b]] += local[a, b]
    return M


@pytest.mark.parametrize('which', ['sphere', 'solid'])
def test_mass_matrix_closed_form(which):
    g, nodes = sphere() if which == 'sphere' else solid()
    M = fe.mass_matrix(g)
    expected = closed_form_mass(g, nodes, 2 if which == 'sphere' else 3)
    np.testing.assert_allclose(M, expected, rtol=1e-13, atol=1e-16)
    np.testing.assert_allclose(fe.mass_diagonal(g)[:, 0], np.diag(expected), rtol=1e-13)
    assert np.allclose(M, M.T) and np.linalg.eigvalsh(M).min() > 0


@pytest.mark.parametrize('which', ['sphere', 'solid'])
def test_evaluate_assemble_adjoint_and_mass(which):
    g, nodes = sphere() if which == 'sphere' else solid()
    rng = np.random.default_rng(3)
    # affine fields are reproduced exactly at the physical quadrature points
    Bm = rng.normal(size=(3, 3))
    a = rng.normal(size=3)
    pts = fe.evaluate(g, nodes)
    np.testing.assert_allclose(fe.evaluate(g, nodes @ Bm.T + a), pts @ Bm.T + a, rtol=1e-13, atol=1e-13)
    v = rng.normal(size=(g.node_count, 3))
    U = rng.normal(size=(g.cells.shape[0] * g.values.shape[0], 3))
    lhs = (fe.assemble(g, U) * v).sum()
    rhs = (U * fe.weighted(g, v)).sum()
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    np.testing.assert_allclose(fe.assemble(g, fe.evaluate(g, v)), fe.mass_matrix(g) @ v, rtol=1e-11, atol=1e-14)
    # total weighted force equals the total nodal force after the mass solve
    b = rng.normal(size=(g.node_count, 3))
    C = np.linalg.solve(fe.mass_matrix(g), b)
    np.testing.assert_allclose(fe.weighted(g, C).sum(0), b.sum(0), rtol=1e-11, atol=1e-12)


def test_pcg_matches_dense_solve_and_error_paths():
    g, nodes = solid()
    M = fe.mass_matrix(g)
    diag = fe.mass_diagonal(g)
    rhs = np.sin(5 * nodes)
    opts = fe.SolverOptions(rtol=1e-11, atol=1e-13, recompute_every=12, check_every=4, max_iterations=500)
    x, info = fe.pcg(lambda v: M @ v, rhs, diag, .1 * np.cos(nodes), opts)
    exact = np.linalg.solve(M, rhs)
    assert info.residual_norm <= info.tolerance and 0 < info.iterations < 500
    assert np.linalg.norm(rhs - M @ x) <= info.tolerance
    np.testing.assert_allclose(x, exact, rtol=1e-8, atol=1e-9 * np.abs(exact).max())
    zero, zinfo = fe.pcg(lambda v: M @ v, np.zeros_like(rhs), diag, None, opts)
    assert zinfo.iterations == 0 and not zero.any()
    with pytest.raises(RuntimeError, match='breakdown'):
        fe.pcg(lambda v: -(M @ v), rhs, diag, None, opts)
    with pytest.raises(RuntimeError, match='converge'):
        fe.pcg(lambda v: M @ v, rhs, diag, None, fe.SolverOptions(max_iterations=1, rtol=1e-14))


def test_p2_rule_and_mass():
    """The 64-point rule integrates monomials of degree <= 5 exactly
    (x^a y^b z^c over the unit tetrahedron = a! b! c! / (a+b+c+3)!), and the
    P2 element mass from it equals an independent degree-11 integration and
    the classical V/420 table; entries sum to the volume."""
    from math import factorial
    pts, w = fe.tet_rule(4)
    for a in range(6):
        for b in range(6 - a):
            for c in range(6 - a - b):
                exact = factorial(a) * factorial(b) * factorial(c) / factorial(a + b + c + 3)
                got = (w * pts[:, 0] ** a * pts[:, 1] ** b * pts[:, 2] ** c).sum()
                assert abs(got - exact) <= 1e-15, (a, b, c)
    N4 = fe.p2_basis(pts)
    p7, w7 = fe.tet_rule(7)
    N7 = fe.p2_basis(p7)
    M4 = np.einsum('q,qa,qb->ab', w, N4, N4)
    M7 = np.einsum('q,qa,qb->ab', w7, N7, N7)
    np.testing.assert_allclose(M4, M7, rtol=1e-13, atol=1e-16)
    V = 1 / 6
    table = np.zeros((10, 10))
    edges = fe.EDGES
    for i in range(4):
        for j in range(4):
            table[i, j] = 6 if i == j else 1
        for e, (p, q) in enumerate(edges):
            table[i, 4 + e] = table[4 + e, i] = -4 if i in (p, q) else -6
    for e, (p, q) in enumerate(edges):
        for f, (r, s) in enumerate(edges):
            shared = len({p, q} & {r, s})
            table[4 + e, 4 + f] = 32 if e == f else (16 if shared == 1 else 8)
    np.testing.assert_allclose(M4, V / 420 * table, rtol=1e-13, atol=1e-17)
    nodes, tets = fe.box_tets(2, (-.4, -.3, -.5), (.5, .4, .3))
    g, nodes2 = fe.p2_solid(nodes, tets)
    assert g.cells.shape[1] == 10 and g.values.shape == (64, 10)
    assert abs(fe.mass_matrix(g).sum() - .9 * .7 * .8) < 1e-13
    assert np.linalg.eigvalsh(fe.mass_matrix(g)).min() > 0
    # P2 reproduces quadratic fields exactly at the physical quadrature points
    q = lambda X: np.stack([X[:, 0] ** 2, X[:, 1] * X[:, 2], X[:, 0] - X[:, 2] ** 2], 1)
    pts_phys = fe.evaluate(g, nodes2)
    np.testing.assert_allclose(fe.evaluate(g, q(nodes2)), q(pts_phys), rtol=1e-12, atol=1e-14)


def test_p2_shared_tables_pinned():
    """The P2 prepare's shared tables (p2_transfer.cu:12-44): interaction
    points equal the affine image of the reference points (straight P2
    tetrahedra), the expanded tables equal the per-component compact stencil
    of oracle/ib3d.py (base bit-exact, phi to rounding: the shared radical
    against the two-branch Peskin formula), the weights are the Peskin
    known answers at f = 0 and f = 1/2, and the support predicate agrees."""
    from oracle import ib3d
    nodes, tets = fe.box_tets(2, (-.4, -.3, -.5), (.5, .4, .3))
    g, nodes2 = fe.p2_solid(nodes, tets)
    x = nodes2 + .01 * np.sin(3 * nodes2)
    pts = fe.p2_points(g, x)
    np.testing.assert_allclose(pts, fe.evaluate(g, x), rtol=0, atol=2e-16)
    # straight element: x_q = v0 + J xi_q
    ref, _ = fe.tet_rule(4)
    v = nodes[tets]
    affine = v[:, None, 0] + np.einsum('eij,qj->eqi', np.stack([v[:, 1] - v[:, 0], v[:, 2] - v[:, 0],
                                                                 v[:, 3] - v[:, 0]], -1), ref)
    np.testing.assert_allclose(fe.p2_points(g, nodes2), affine.reshape(-1, 3), rtol=0, atol=1e-15)
    grid = ib3d.Grid3((16, 20, 24), (4., 6., 8.), (-2., -3., -4.))
    base, phi, valid = fe.p2_shared_tables(pts, grid)
    assert valid.all() and ib3d.support_ok(pts, grid)
    cb, cp = fe.shared_expanded(base, phi)
    rb, rp = ib3d.prepare_compact(pts, grid)
    np.testing.assert_array_equal(cb, rb)
    np.testing.assert_allclose(cp, rp, rtol=0, atol=2e-15)
    # planted mistake: swapping the lattices breaks the expansion
    sb, sp = fe.shared_expanded(base[::-1], phi[::-1])
    assert not np.array_equal(sb, rb)
    # known answers (ib.py:57-67): f = 0 -> (1/4, 1/2, 1/4, 0);
    # f = 1/2 -> ((2-r)/8, (2+r)/8, (2+r)/8, (2-r)/8) with r = sqrt(2)
    dyadic = ib3d.Grid3((16, 16, 32), (4., 2., 4.), (-2., -1., -2.))
    h = np.asarray(dyadic.spacing)
    knots = np.asarray(dyadic.origin) + h * np.array([[5., 6., 7.], [5.5, 6.5, 7.5]])
    kb, kp, kv = fe.p2_shared_tables(knots, dyadic)
    np.testing.assert_array_equal(kb[0, 0], [4, 5, 6])
    np.testing.assert_allclose(kp[0, 0], np.tile([.25, .5, .25, 0.], (3, 1)), atol=1e-16)
    np.testing.assert_allclose(kp[0, 1], np.tile(np.array([2 - 2 ** .5, 2 + 2 ** .5, 2 + 2 ** .5, 2 - 2 ** .5]) / 8, (3, 1)),
                               atol=1e-16)
    np.testing.assert_allclose(kp[1, 1], np.tile([.25, .5, .25, 0.], (3, 1)), atol=1e-16)
    # support: outside, on the lower limit (valid), on the upper limit (invalid)
    h = np.asarray(grid.spacing)
    probe = np.asarray(grid.origin) + h * np.array([[1.99, 5, 5], [2., 5., 5.], [5., 18., 5.], [5., 5., 21.999]])
    assert fe.p2_shared_tables(probe, grid)[2].tolist() == [False, True, False, True]
    for p in probe:
        assert fe.p2_shared_tables(p[None], grid)[2][0] == ib3d.support_ok(p[None], grid)


def test_assembled_operator_oracle_pinned():
    """The table restatement: the SplitMix64 finaliser equals the published
    generator's first output for seed 0 (0xe220a8397b1dcdaf, the mix of the
    golden-ratio increment), forty colliders of one bucket repeated seven
    times sum to 0.875 exactly and a full table flags overflow
    (tests/test_csr_hash_assembly.py:22-40). The dense B_c reproduces the
    pinned spread and gather of the oracle: B^T (W-weighted coefficient) / V
    and B u."""
    from oracle import assembled, ib3d
    assert assembled.splitmix_finalise(0x9e3779b97f4a7c15) == 0xe220a8397b1dcdaf
    colliders = [k for k in range(10000) if assembled.bucket(k, 128) == 3][:40]
    assert len(colliders) == 40
    keys = np.array(colliders * 7 + [987654])
    values = np.full(len(keys), .125)
    values[-1] = 0
    tk, tv, flag = assembled.table_accumulate(keys, values, 128)
    k, v = assembled.sorted_entries(tk, tv)
    assert flag == 0 and k.tolist() == sorted(colliders) and (v == .875).all()
    assert assembled.table_accumulate(np.arange(20), np.ones(20), 8)[2] == 1
    g, X = solid()
    grid = ib3d.Grid3((16, 20, 24), (4., 6., 8.), (-2., -3., -4.))
    x = X + .01 * np.sin(3 * X)
    pts = fe.evaluate(g, x)
    C = np.cos(2 * X)
    base, phi = ib3d.prepare_compact(pts, grid)
    spread = ib3d.spread_grid(fe.weighted(g, C), base, phi, grid)
    u = [np.sin(grid.coordinates(c).sum(-1) + c) for c in range(3)]
    rhs = fe.assemble(g, ib3d.gather_grid(u, base, phi, grid))
    for c in range(3):
        B = assembled.dense_operator(g, pts, grid, c)
        np.testing.assert_allclose((B.T @ C[:, c]).reshape(spread[c].shape) / grid.volume, spread[c],
                                   rtol=0, atol=1e-13 * np.abs(spread[c]).max())
        np.testing.assert_allclose(B @ u[c].reshape(-1), rhs[:, c], rtol=0, atol=1e-13 * np.abs(rhs).max())
