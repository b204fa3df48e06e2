"""Seeded synthetic point clouds and forces for the hot-path workloads.

All draws use ``numpy.random.Generator(PCG64(seed))``; positions and forces
are FP64. Shapes:

* ``shell(a, count, thickness, ...)``: points on an ellipsoid with semi-axes
  ``a`` centred at ``centre``, uniform in area (sphere directions mapped to
  the ellipsoid, accepted with probability proportional to the area
  element sqrt((bc x)^2+(ac y)^2+(ab z)^2)), offset along the outward normal
  by U(-thickness/2, thickness/2).
* ``patch``: the same construction restricted to the part of the shell
  within ``radius`` of the surface point at the +x tip.
* ``curves2d``: 2D sinusoidal curves y = 0.5 +- 0.1 sin(2 pi x / Lx) with
  jitter of +-0.05 h.
"""
import numpy as np

GEOMETRY_SEED = 20261202


def rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def _sphere(gen, n):
    v = gen.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def _area_weight(u, a):
    b = np.array([a[1] * a[2], a[0] * a[2], a[0] * a[1]])
    return np.sqrt(((u * b) ** 2).sum(1)) / b.max()


def _place(gen, u, a, centre, thickness):
    surf = u * a
    normal = surf / a ** 2
    normal /= np.linalg.norm(normal, axis=1, keepdims=True)
    offset = gen.uniform(-thickness / 2, thickness / 2, size=(len(u), 1))
    return centre + surf + offset * normal, surf


def shell(gen, a, count, thickness=0.02, centre=(.5, .5, .5), truncate_x=None):
    """``count`` shell points; ``truncate_x`` keeps surface points with
    x - centre_x <= truncate_x (a truncated shell)."""
    a, centre = np.asarray(a, float), np.asarray(centre, float)
    out = []
    have = 0
    while have < count:
        u = _sphere(gen, max(2 * (count - have), 1024))
        keep = gen.uniform(size=len(u)) < _area_weight(u, a)
        if truncate_x is not None:
            keep &= (u[:, 0] * a[0]) <= truncate_x
        u = u[keep]
        pts, _ = _place(gen, u, a, centre, thickness)
        out.append(pts)
        have += len(pts)
    return np.concatenate(out)[:count]


def patch(gen, a, count, radius=0.05, thickness=0.02, centre=(.5, .5, .5)):
    """``count`` points of the shell within ``radius`` of its +x tip."""
    a, centre = np.asarray(a, float), np.asarray(centre, float)
    tip = np.array([a[0], 0., 0.])
    cap = min(np.pi, 2.5 * radius / min(a[1], a[2]))
    out = []
    have = 0
    while have < count:
        n = max(4 * (count - have), 4096)
        # uniform directions in the cap of half-angle `cap` around +x
        cos_t = gen.uniform(np.cos(cap), 1., size=n)
        phi = gen.uniform(0., 2 * np.pi, size=n)
        sin_t = np.sqrt(1 - cos_t ** 2)
        u = np.stack([cos_t, sin_t * np.cos(phi), sin_t * np.sin(phi)], 1)
        keep = gen.uniform(size=n) < _area_weight(u, a)
        keep &= np.linalg.norm(u * a - tip, axis=1) <= radius
        u = u[keep]
        pts, _ = _place(gen, u, a, centre, thickness)
        out.append(pts)
        have += len(pts)
    return np.concatenate(out)[:count]


def ellipsoid_area(a):
    """Knud Thomsen's approximation (p=1.6075), for splitting counts."""
    p = 1.6075
    x, y, z = (float(v) ** p for v in a)
    return 4 * np.pi * ((x * y + x * z + y * z) / 3) ** (1 / p)


def normal_forces(gen, count, dim=3):
    return gen.normal(size=(count, dim))


def lognormal_forces(gen, count):
    magnitude = gen.lognormal(0., 1., size=(count, 1))
    return magnitude * _sphere(gen, count)


def truncated_shell_64(points=200_000, seed=GEOMETRY_SEED):
    """64^3 unit cube: truncated shell (0.3,0.2,0.2), thickness 0.02,
    truncation x - 0.5 <= 0.6*0.3 (build's choice), N(0,1) forces."""
    gen = rng(seed)
    x = shell(gen, (.3, .2, .2), points, .02, truncate_x=.6 * .3)
    return x, normal_forces(gen, points)


def clustered_shells_128(points=1_500_000, seed=GEOMETRY_SEED + 1):
    """128^3 unit cube: 90% on shells (0.35,0.25,0.25) and (0.3,0.2,0.2)
    split by area, 10% clustered within 0.05 of the outer +x tip;
    forces LogNormal(0,1) magnitude, uniform direction."""
    gen = rng(seed)
    clustered = points // 10
    rest = points - clustered
    outer, inner = (.35, .25, .25), (.3, .2, .2)
    ao, ai = ellipsoid_area(outer), ellipsoid_area(inner)
    n_outer = int(round(rest * ao / (ao + ai)))
    x = np.concatenate([shell(gen, outer, n_outer), shell(gen, inner, rest - n_outer),
                        patch(gen, outer, clustered)])
    return x, lognormal_forces(gen, points)


def curves2d(gen, lengths=(4., 1.), shape=(256, 64), per_curve=2048):
    """Channel curves: x_i = Lx (i+1/2)/N, y = 0.5 + k 0.1 sin(2 pi x / Lx),
    k = +-1, plus U(-0.05h, 0.05h) jitter on both coordinates."""
    h = lengths[0] / shape[0]
    i = np.arange(per_curve)
    xs = lengths[0] * (i + .5) / per_curve
    pts = []
    for k in (1, -1):
        y = .5 + k * .1 * np.sin(2 * np.pi * xs / lengths[0])
        pts.append(np.stack([xs, y], 1))
    pts = np.concatenate(pts)
    return pts + gen.uniform(-.05 * h, .05 * h, size=pts.shape)


# ---------------------------------------------------------------------------
# many-instance workloads (batched runs and the request stream)

GRIDS_3D = ((32, 32, 32), (48, 48, 48), (64, 64, 64), (96, 96, 96))
GRIDS_2D = ((128, 32), (256, 64), (512, 128))


def random_shell(gen, count):
    """A random centred ellipsoidal shell (semi-axes U(0.15, 0.35) each,
    thickness 0.02) in the unit cube."""
    return shell(gen, gen.uniform(.15, .35, size=3), count)


def random_curves(gen, count, shape, lengths=(4., 1.)):
    """Two sinusoidal curves y = 0.5 +- A sin(2 pi x / Lx + phi),
    A ~ U(0.05, 0.15), phi ~ U(0, 2 pi), count points in total, with
    U(-0.05 h, 0.05 h) jitter (h = Lx / nx)."""
    h = lengths[0] / shape[0]
    per = count // 2
    amp, phase = gen.uniform(.05, .15), gen.uniform(0, 2 * np.pi)
    pts = []
    for k, n in ((1, per), (-1, count - per)):
        xs = lengths[0] * (np.arange(n) + .5) / n
        pts.append(np.stack([xs, .5 + k * amp * np.sin(2 * np.pi * xs / lengths[0] + phase)], 1))
    pts = np.concatenate(pts)
    return pts + gen.uniform(-.05 * h, .05 * h, size=pts.shape)


def instance_spec(gen, kind):
    """(grid shape, point count) of one instance: 'batch' draws the grid
    uniformly from the 3D and 2D sets and points log-uniform in [1e4, 3e5];
    'serve-2d' / 'serve-3d' / 'serve-128' draw the request-stream classes
    (2D 128x32..512x128 with 2k-20k points; 3D 32^3-64^3 with 20k-200k;
    128^3 with 0.5M-1.5M)."""
    def loguniform(lo, hi):
        return int(round(np.exp(gen.uniform(np.log(lo), np.log(hi)))))
    if kind == 'batch':
        shapes = GRIDS_3D + GRIDS_2D
        shape = shapes[gen.integers(len(shapes))]
        return shape, loguniform(1e4, 3e5)
    if kind == 'serve-2d':
        return GRIDS_2D[gen.integers(3)], loguniform(2e3, 2e4)
    if kind == 'serve-3d':
        return GRIDS_3D[gen.integers(3)], loguniform(2e4, 2e5)
    if kind == 'serve-128':
        return (128, 128, 128), loguniform(5e5, 1.5e6)
    raise ValueError(kind)


def instance_points(gen, shape, count):
    """Positions and N(0,1) forces of an instance (shell in 3D, curves in 2D)."""
    if len(shape) == 3:
        x = random_shell(gen, count)
    else:
        x = random_curves(gen, count, shape)
    return x, gen.normal(size=x.shape)


def request_stream(seed, rate=200., duration=60., per_instance=10, live=20,
                   mix=((.6, 'serve-2d'), (.3, 'serve-3d'), (.1, 'serve-128'))):
    """Poisson arrivals at ``rate``/s over ``duration`` s. Arrival k belongs
    to instance (k mod live) + live * (k // (live * per_instance)): ``live``
    instances receive requests in turn, each ``per_instance`` requests.
    Returns (arrival times, instance ids, {instance id: (shape, count, seed)})."""
    gen = rng(seed)
    times = np.cumsum(gen.exponential(1 / rate, size=int(rate * duration * 1.5) + 16))
    times = times[times < duration]
    k = np.arange(len(times))
    inst = (k % live) + live * (k // (live * per_instance))
    specs = {}
    probs = np.array([p for p, _ in mix])
    kinds = [kd for _, kd in mix]
    for i in np.unique(inst):
        kind = kinds[gen.choice(len(kinds), p=probs / probs.sum())]
        shape, count = instance_spec(gen, kind)
        specs[int(i)] = (shape, count, int(gen.integers(2 ** 31)))
    return times, inst, specs


def fe_sphere(level, radius=.3, centre=(.5, .5, .5)):
    """P1 triangulated sphere (subdivided icosahedron, vectorised midpoint
    subdivision) with the degree-2 three-point triangle rule: returns
    (nodes (V, 3), cells (E, 3) int64, values N (3, 3), weights W (E, 3))."""
    t = (1 + 5 ** .5) / 2
    v = np.array([(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
                  (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)], dtype=np.float64)
    f = np.array([(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
                  (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
                  (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)], dtype=np.int64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    for _ in range(level):
        edges = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        key = np.sort(edges, axis=1)
        uniq, inverse = np.unique(key, axis=0, return_inverse=True)
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        ids = len(v) + inverse.reshape(3, -1)
        ab, bc, ca = ids[0], ids[1], ids[2]
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        f = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1), np.stack([c, ca, bc], 1),
                            np.stack([ab, bc, ca], 1)])
        v = np.concatenate([v, mid])
    nodes = v * radius + np.asarray(centre)
    values = np.array([[2 / 3, 1 / 6, 1 / 6], [1 / 6, 2 / 3, 1 / 6], [1 / 6, 1 / 6, 2 / 3]])
    p = nodes[f]
    area = .5 * np.linalg.norm(np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), axis=1)
    return nodes, f, values, np.repeat(area[:, None] / 3, 3, axis=1)


# P2 tetrahedra: node order = 4 vertices, then the midpoints of these edges
# (tetrahedron.py:9-56)
P2_EDGES = ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))


def collapsed_gauss_tet(n):
    """n^3-point collapsed (Duffy) Gauss-Legendre rule on the unit
    tetrahedron, positive weights summing to 1/6; exact for total degree
    <= 2n - 3 (n = 4: the 64-point rule, exact for the P2 mass)."""
    u, wu = np.polynomial.legendre.leggauss(n)
    u, wu = (u + 1) / 2, wu / 2
    a, b, c = np.meshgrid(u, u, u, indexing='ij')
    wa, wb, wc = np.meshgrid(wu, wu, wu, indexing='ij')
    pts = np.stack([a, b * (1 - a), c * (1 - a) * (1 - b)], -1).reshape(-1, 3)
    w = (wa * wb * wc * (1 - a) ** 2 * (1 - b)).reshape(-1)
    return pts, w


def p2_values(points):
    """The ten P2 basis functions at reference points (tetrahedron.py:19-30)."""
    L = np.concatenate([1 - points.sum(1, keepdims=True), points], 1)
    i, j = np.array(P2_EDGES).T
    return np.concatenate([L * (2 * L - 1), 4 * L[:, i] * L[:, j]], 1)


def lv_p1(rule=4, **shape):
    """The idealised LV wall of ``lv_p2`` as straight P1 tetrahedra (vertices
    only) with the ``rule``^3-point collapsed Gauss rule and the barycentric
    basis. Returns (nodes (V,3), cells (E,4) int64, values N (Q,4), weights W (E,Q))."""
    nodes, cells, _, weights = lv_p2(rule=rule, **shape)
    tets = cells[:, :4]
    used = np.unique(tets)
    remap = np.full(len(nodes), -1, dtype=np.int64)
    remap[used] = np.arange(len(used))
    pts, _ = collapsed_gauss_tet(rule)
    values = np.concatenate([1 - pts.sum(1, keepdims=True), pts], 1)
    return nodes[used], remap[tets], values, weights


def lv_p2(nt=3, ntheta=48, nphi=70, rule=4, inner=(.36, .24), outer=(.42, .30), centre=(.5, .5, .5),
          theta=(.15, .6 * np.pi)):
    """Idealised left-ventricle wall of straight P2 tetrahedra: the truncated
    prolate shell between the ellipsoids of semi-axes ``inner`` and
    ``outer`` (long axis along x, apex towards -x), polar angle in
    ``theta`` (an apical hole of 0.15 rad, truncated beyond the equator),
    ``nt`` x ``ntheta`` x ``nphi`` hexahedra split into six tetrahedra along
    their main diagonal, each promoted to ten nodes with shared edge
    midpoints, and the ``rule``^3-point collapsed Gauss rule. The default
    has 60,480 elements and 3,870,720 interaction points.
    Returns (nodes (V,3), cells (E,10) int64, values N (Q,10), weights W (E,Q))."""
    i, j, k = np.meshgrid(np.arange(nt + 1), np.arange(ntheta + 1), np.arange(nphi), indexing='ij')
    s = i / nt
    th = theta[0] + (theta[1] - theta[0]) * j / ntheta
    ph = 2 * np.pi * k / nphi
    a = inner[0] + s * (outer[0] - inner[0])
    b = inner[1] + s * (outer[1] - inner[1])
    vertices = np.stack([centre[0] - a * np.cos(th), centre[1] + b * np.sin(th) * np.cos(ph),
                         centre[2] + b * np.sin(th) * np.sin(ph)], -1).reshape(-1, 3)
    hi, hj, hk = (g.reshape(-1) for g in np.meshgrid(np.arange(nt), np.arange(ntheta), np.arange(nphi),
                                                        indexing='ij'))
    node = lambda di, dj, dk: ((hi + di) * (ntheta + 1) + hj + dj) * nphi + (hk + dk) % nphi
    corner = [node(c & 1, (c >> 1) & 1, (c >> 2) & 1) for c in range(8)]
    tets = []
    for p in ((1, 2, 4), (1, 4, 2), (2, 1, 4), (2, 4, 1), (4, 1, 2), (4, 2, 1)):
        tets.append(np.stack([corner[0], corner[p[0]], corner[p[0] | p[1]], corner[7]], 1))
    tets = np.concatenate(tets).astype(np.int64)
    pairs = np.sort(tets[:, np.array(P2_EDGES)].reshape(-1, 2), axis=1)
    unique, inverse = np.unique(pairs, axis=0, return_inverse=True)
    nodes = np.concatenate([vertices, vertices[unique].mean(1)])
    cells = np.concatenate([tets, len(vertices) + inverse.reshape(-1, 6)], 1)
    pts, w = collapsed_gauss_tet(rule)
    v = vertices[tets]
    volume = np.abs(np.einsum('ei,ei->e', v[:, 1] - v[:, 0], np.cross(v[:, 2] - v[:, 0], v[:, 3] - v[:, 0]))) / 6
    return nodes, cells, p2_values(pts), 6 * volume[:, None] * w[None]
