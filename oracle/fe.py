"""CPU restatement of the FE interaction-point transfer (test infrastructure).

Follows the reference's quadrature IB equations (mac/transfer.py:1-5):
M F = b, f = h^-3 K^T W B F, M U = B^T W K u, with the consistent FE mass
M = sum_e sum_q W_eq N_qa N_qb (mac/transfer.py:32-41) solved by the Jacobi
PCG of fluid/solvers.py:39-122. K is the Peskin-4 MAC stencil of
oracle/ib3d.py evaluated at the interaction points B x.

Geometry: ``cells`` (E, A) node ids, ``values`` N (Q, A) basis values at the
reference quadrature points, ``weights`` W (E, Q) physical quadrature weights,
``node_count``. Used only by tests, the smoke check and the benchmark's CPU
baseline.
"""
from dataclasses import dataclass
from math import isfinite

import numpy as np


@dataclass(frozen=True)
class Geometry:
    cells: np.ndarray
    values: np.ndarray
    weights: np.ndarray
    node_count: int


@dataclass(frozen=True)
class SolverOptions:
    """fluid/solvers.py:13-27 (defaults of FETransfer, mac/transfer.py:29-30)."""
    rtol: float = 1e-12
    atol: float = 1e-13
    max_iterations: int = 500
    recompute_every: int = 40
    check_every: int = 4


@dataclass(frozen=True)
class SolveInfo:
    iterations: int
    residual_norm: float
    rhs_norm: float
    tolerance: float


def mass_matrix(g):
    """Dense consistent mass (mac/transfer.py:32-38): local_eab = sum_q W_eq N_qa N_qb,
    scattered to (cells[e,a], cells[e,b]) with duplicates summed."""
    local = np.einsum('eq,qa,qb->eab', g.weights, g.values, g.values)
    M = np.zeros((g.node_count, g.node_count))
    A = g.cells.shape[1]
    rows = np.repeat(g.cells[:, :, None], A, axis=2).reshape(-1)
    cols = np.repeat(g.cells[:, None, :], A, axis=1).reshape(-1)
    np.add.at(M, (rows, cols), local.reshape(-1))
    return M


def mass_diagonal(g):
    """Jacobi diagonal (mac/transfer.py:39-40), expanded to three columns."""
    local = np.einsum('eq,qa,qa->ea', g.weights, g.values, g.values)
    d = np.zeros(g.node_count)
    np.add.at(d, g.cells.reshape(-1), local.reshape(-1))
    return np.repeat(d[:, None], 3, axis=1)


def evaluate(g, nodal):
    """B: interaction-point values sum_a N_qa v[cells[e,a]] (mac/transfer.py:57-58)."""
    return np.einsum('qa,eai->eqi', g.values, nodal[g.cells]).reshape(-1, 3)


def weighted(g, nodal):
    """W B: evaluated values times the quadrature weight (mac/transfer.py:66-67)."""
    return (evaluate(g, nodal).reshape(*g.weights.shape, 3) * g.weights[..., None]).reshape(-1, 3)


def assemble(g, point_values):
    """B^T W: sum_q N_qa W_eq U_eq scattered to the nodes (mac/transfer.py:126-129)."""
    local = np.einsum('qa,eq,eqi->eai', g.values, g.weights, point_values.reshape(*g.weights.shape, 3))
    out = np.zeros((g.node_count, 3))
    np.add.at(out, g.cells.reshape(-1), local.reshape(-1, 3))
    return out


class PCGError(RuntimeError):
    pass


def pcg(action, rhs, diagonal, initial=None, options=SolverOptions(), history=None):
    """Jacobi PCG with the reference's recurrence, check/restart schedule and
    true-residual acceptance (fluid/solvers.py:39-122, no fixed DOFs).
    ``history`` (a list) receives the recurrence norm of every iteration."""
    norm = lambda v: float(np.linalg.norm(v))
    rhs_norm = norm(rhs)
    tol = max(options.atol, options.rtol * rhs_norm)
    correction = np.zeros_like(rhs) if initial is None else initial.copy()
    residual = rhs - action(correction)
    if norm(residual) <= tol:
        return correction, SolveInfo(0, norm(residual), rhs_norm, tol)
    z = residual / diagonal
    direction = z.copy()
    rz = float((residual * z).sum())
    active, broken = True, False
    true_norm = float('nan')
    for iteration in range(1, options.max_iterations + 1):
        Ad = action(direction)
        curvature = float((direction * Ad).sum())
        valid = isfinite(curvature) and curvature > 0 and isfinite(rz) and rz > 0
        broken = broken or (active and not valid)
        working = active and valid and not broken
        alpha = rz / curvature if working else 0.0
        correction = correction + alpha * direction
        residual = residual - alpha * Ad
        recurrence_norm = norm(residual)
        if history is not None:
            history.append(recurrence_norm)
        broken = broken or not isfinite(recurrence_norm)
        active = working and recurrence_norm > tol
        restart = iteration % options.recompute_every == 0
        check_now = restart or iteration % options.check_every == 0 or iteration == options.max_iterations
        converged = False
        if check_now:
            if broken:
                raise PCGError('PCG breakdown: operator/preconditioner must be positive definite and finite')
            converged = recurrence_norm <= tol
        if restart or converged or iteration == options.max_iterations:
            residual = rhs - action(correction)
            true_norm = norm(residual)
            if not isfinite(true_norm):
                raise PCGError('nonfinite true PCG residual')
            if true_norm <= tol:
                return correction, SolveInfo(iteration, true_norm, rhs_norm, tol)
            restart = True
            active = True
        z = residual / diagonal
        next_rz = float((residual * z).sum())
        if restart:
            direction = z
        else:
            beta = next_rz / (rz if active else 1.0)
            direction = z + beta * direction if active else np.zeros_like(z)
        rz = next_rz
    raise PCGError(f'PCG did not converge after {options.max_iterations} iterations: '
                   f'true residual {true_norm:.6e} > tolerance {tol:.6e}')


def interaction_points(g, x):
    """Interaction points B x (mac/transfer.py:69-71)."""
    return evaluate(g, x)


def spread(g, grid, nodal_force, options=SolverOptions()):
    """f = V^-1 K^T W B M^-1 b (mac/transfer.py:107-113, 115-122)."""
    M = mass_matrix(g)
    coefficient, info = pcg(lambda v: M @ v, nodal_force, mass_diagonal(g), options=options)
    return coefficient, weighted(g, coefficient), info


def interpolate_rhs(g, point_velocity):
    """B^T W K u: the right-hand side of the velocity mass solve (mac/transfer.py:131-138)."""
    return assemble(g, point_velocity)


# --- test geometries ---------------------------------------------------------

TRIANGLE_RULE = (np.array([[2 / 3, 1 / 6, 1 / 6], [1 / 6, 2 / 3, 1 / 6], [1 / 6, 1 / 6, 2 / 3]]),
                 np.full(3, 1 / 3))                 # degree 2, barycentric points, weights per unit area
_A, _B = 0.5854101966249685, 0.1381966011250105
TETRA_RULE = (np.array([[_A, _B, _B, _B], [_B, _A, _B, _B], [_B, _B, _A, _B], [_B, _B, _B, _A]]),
              np.full(4, 1 / 4))                    # degree 2, weights per unit volume


def p1_surface(nodes, triangles, rule=TRIANGLE_RULE):
    """P1 triangles: N = barycentric coordinates of the rule's points, W = area x weight."""
    bary, w = rule
    v = nodes[triangles]
    area = .5 * np.linalg.norm(np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]), axis=1)
    return Geometry(triangles.astype(np.int64), bary.copy(), area[:, None] * w[None], len(nodes))


def p1_solid(nodes, tets, rule=TETRA_RULE):
    """P1 tetrahedra: N = barycentric coordinates, W = volume x weight."""
    bary, w = rule
    v = nodes[tets]
    vol = np.abs(np.einsum('ei,ei->e', v[:, 1] - v[:, 0], np.cross(v[:, 2] - v[:, 0], v[:, 3] - v[:, 0]))) / 6
    return Geometry(tets.astype(np.int64), bary.copy(), vol[:, None] * w[None], len(nodes))


def icosphere(level, radius=1.0, centre=(0., 0., 0.)):
    """Subdivided icosahedron: (nodes, triangles)."""
    t = (1 + 5 ** .5) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    nodes = [np.array(p, float) / np.linalg.norm(p) for p in v]
    faces = f
    for _ in range(level):
        cache, new = {}, []

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = nodes[a] + nodes[b]
                nodes.append(m / np.linalg.norm(m))
                cache[key] = len(nodes) - 1
            return cache[key]
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            new += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = new
    return np.asarray(nodes) * radius + np.asarray(centre), np.asarray(faces, dtype=np.int64)


def box_tets(n, lo, hi):
    """Structured n^3 cube split into 6 tetrahedra per cell: (nodes, tets)."""
    axes = [np.linspace(lo[a], hi[a], n + 1) for a in range(3)]
    X, Y, Z = np.meshgrid(*axes, indexing='ij')
    nodes = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    idx = lambda i, j, k: (i * (n + 1) + j) * (n + 1) + k
    tets = []
    for i in range(n):
        for j in range(n):
            for k in range(n):
                c = [idx(i + a, j + b, k + d) for a in (0, 1) for b in (0, 1) for d in (0, 1)]
                # Kuhn split along the main diagonal c[0] -> c[7]
                for p in ((1, 3), (1, 5), (2, 3), (2, 6), (4, 5), (4, 6)):
                    tets.append((c[0], c[p[0]], c[p[0] | p[1]], c[7]))
    return nodes, np.asarray(tets, dtype=np.int64)


# P2 tetrahedra (tetrahedron.py:9-56 node order: 4 vertices, then the midpoints
# of EDGES = (0,1),(0,2),(0,3),(1,2),(1,3),(2,3))
EDGES = ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))


def tet_rule(n):
    """Collapsed (Duffy) Gauss-Legendre rule on the unit tetrahedron with n^3
    positive weights; exact for total degree <= 2n - 3."""
    u, wu = np.polynomial.legendre.leggauss(n)
    u, wu = (u + 1) / 2, wu / 2
    pts, wts = [], []
    for a, wa in zip(u, wu):
        for b, wb in zip(u, wu):
            for c, wc in zip(u, wu):
                pts.append((a, b * (1 - a), c * (1 - a) * (1 - b)))
                wts.append(wa * wb * wc * (1 - a) ** 2 * (1 - b))
    return np.asarray(pts), np.asarray(wts)


def p2_basis(points):
    """N[q, a] of the ten P2 functions at reference points (tetrahedron.py:19-30)."""
    L = np.concatenate([1 - points.sum(1, keepdims=True), points], 1)
    i, j = np.array(EDGES).T
    return np.concatenate([L * (2 * L - 1), 4 * L[:, i] * L[:, j]], 1)


def p2_promote(nodes, tets):
    """One shared midpoint per unique edge (tetrahedron.py:45-56)."""
    pairs = np.sort(tets[:, np.array(EDGES)].reshape(-1, 2), axis=1)
    unique, inverse = np.unique(pairs, axis=0, return_inverse=True)
    mid = nodes[unique].mean(1)
    return np.concatenate([nodes, mid]), np.concatenate([tets, len(nodes) + inverse.reshape(-1, 6)], 1)


def p2_solid(nodes, tets, n=4):
    """P2 geometry of a P1 tet mesh: N (n^3, 10) from tet_rule(n), W = 6 |T| w."""
    nodes2, cells = p2_promote(nodes, tets)
    pts, w = tet_rule(n)
    v = nodes[tets]
    vol = np.abs(np.einsum('ei,ei->e', v[:, 1] - v[:, 0], np.cross(v[:, 2] - v[:, 0], v[:, 3] - v[:, 0]))) / 6
    return Geometry(cells.astype(np.int64), p2_basis(pts), 6 * vol[:, None] * w[None], len(nodes2)), nodes2


# --- P2 fused cell-block transfer (mac/p2_cuda_transfer.py) ------------------
def p2_points(g, x):
    """x_q = sum_a N_qa x[cells[e,a]], accumulated a = 0..9 in order from zero
    with separate multiply and add (p2_transfer.cu:21-28)."""
    pos = np.zeros((g.cells.shape[0], g.values.shape[0], 3))
    for a in range(g.cells.shape[1]):
        pos = pos + g.values[None, :, a, None] * x[g.cells[:, a]][:, None, :]
    return pos.reshape(-1, 3)


def p2_shared_tables(points, grid):
    """Shared face/centre tables of the P2 prepare (p2_transfer.cu:12-44,
    layout mac/shared_stencil.py:8-12): base (2,P,3), phi (2,P,3,4) with
    lattice 0 starting at the origin (faces) and lattice 1 at origin + h/2
    (centres); f = s - lower - 1 and the four weights (3-2f-R)/8,
    (3-2f+R)/8, (1+2f+R)/8, (1+2f-R)/8 with R = sqrt(1+4f-4f^2). Returns
    also the support predicate 2 <= (x-origin)/h < n-2 per point
    (mac/compact_transfer.py:87,94-96); an invalid axis is tabulated at s = 0
    as the reference does."""
    h = np.asarray(grid.spacing)
    origin = np.asarray(grid.origin, dtype=np.float64)
    starts = np.stack([origin, origin + .5 * h])
    physical = (points - origin) / h
    valid = np.isfinite(physical) & (physical >= 2) & (physical < np.asarray(grid.shape) - 2)
    base = np.empty((2, len(points), 3), dtype=np.int64)
    phi = np.empty((2, len(points), 3, 4))
    for lattice in range(2):
        scaled = np.where(valid, (points - starts[lattice]) / h, 0.)
        lower = np.floor(scaled - 1.)
        f = scaled - lower - 1.
        R = np.sqrt(1. + 4. * f - 4. * (f * f))
        base[lattice] = lower.astype(np.int64)
        phi[lattice] = np.stack([(3 - 2 * f - R) / 8, (3 - 2 * f + R) / 8,
                                 (1 + 2 * f + R) / 8, (1 + 2 * f - R) / 8], -1)
    return base, phi, valid.all(1)


def shared_expanded(base, phi):
    """Per-component compact tables of the shared ones (mac/shared_stencil.py:13-19):
    component c takes lattice 0 along axis c and lattice 1 elsewhere."""
    cb = np.stack([np.stack([base[int(a != c), :, a] for a in range(3)], -1) for c in range(3)])
    cp = np.stack([np.stack([phi[int(a != c), :, a, :] for a in range(3)], 1) for c in range(3)])
    return cb, cp
