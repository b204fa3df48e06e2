"""Oracle: Peskin-4 transfer between Lagrangian points and 3D MAC faces (FP64).

Test infrastructure only (see ``oracle/__init__.py``).

Reference (``src/afsi_torch``):
  ib.py:57-67                    peskin4 (two branches, clamped radicands)
  ib.py:78-105,114-144           nodal-lattice prepare/interpolate/spread
  mac/grid.py:27-40              spacing, face_shape, face_origin
  mac/compact_transfer.py:21-26  _prepare: base=floor(scaled-1), phi
  mac/compact_transfer.py:163-201  expanded 64-link ids/weights, chunked
                                   index_add_ spread and gather (4096 points)
  mac/compact_transfer.py:94-96,115-117  support predicate 2 <= (x-o)/h < n-2
The periodic lattice has no reference counterpart: node indices wrap mod n and
every face array has n entries per axis.
"""
from dataclasses import dataclass
import numpy as np

CHUNK = 4096  # mac/compact_transfer.py:180-184


@dataclass(frozen=True)
class Grid3:
    """MAC grid (mac/grid.py:13-40). ``periodic`` selects the build-own
    periodic lattice; otherwise the reference's closed Neumann box."""
    shape: tuple
    lengths: tuple
    origin: tuple = (0., 0., 0.)
    periodic: bool = False

    @property
    def spacing(self):
        return tuple(L / n for L, n in zip(self.lengths, self.shape))

    @property
    def volume(self):
        hx, hy, hz = self.spacing
        return hx * hy * hz

    def face_shape(self, c):
        if self.periodic:
            return tuple(self.shape)
        return tuple(n + (a == c) for a, n in enumerate(self.shape))

    def face_origin(self, c):
        return tuple(o + (0. if a == c else .5 * h)
                     for a, (o, h) in enumerate(zip(self.origin, self.spacing)))

    def cell_count(self):
        nx, ny, nz = self.shape
        return nx * ny * nz

    def coordinates(self, c=None):
        shape = self.shape if c is None else self.face_shape(c)
        origin = (tuple(o + .5 * h for o, h in zip(self.origin, self.spacing))
                  if c is None else self.face_origin(c))
        axes = [o + h * np.arange(n, dtype=np.float64)
                for n, h, o in zip(shape, self.spacing, origin)]
        return np.stack(np.meshgrid(*axes, indexing='ij'), -1)


def peskin4(distance):
    """phi(r) of ib.py:57-67, same operation order (no fused multiply-add)."""
    r = np.abs(distance)
    a = np.minimum(r, 1.)
    b = np.clip(r, 1., 2.)
    inner = (3. - 2. * a + np.sqrt(1. + 4. * a - 4. * (a * a))) / 8.
    outer = (5. - 2. * b - np.sqrt(-7. + 12. * b - 4. * (b * b))) / 8.
    return np.where(r < 1., inner, np.where(r < 2., outer, 0.))


def support_ok(points, grid):
    """mac/compact_transfer.py:94-96,115-117 (and mac/transfer.py:71-78): every
    point finite with 2 <= (x-origin)/h < n-2 on each axis. Always true for
    finite points on the periodic lattice."""
    points = np.asarray(points, dtype=np.float64)
    if not np.isfinite(points).all():
        return False
    if grid.periodic:
        return True
    scaled = (points - np.asarray(grid.origin)) / np.asarray(grid.spacing)
    limits = np.asarray(grid.shape, dtype=np.float64) - 2
    return bool((scaled >= 2).all() and (scaled < limits).all())


def prepare_compact(points, grid):
    """mac/compact_transfer.py:21-26: base (3,P,3) int64, phi (3,P,3,4).

    Component c uses origin ``face_origin(c)``; base=floor((x-o_c)/h-1)."""
    points = np.asarray(points, dtype=np.float64)
    origins = np.asarray([grid.face_origin(c) for c in range(3)])
    spacing = np.asarray(grid.spacing)
    scaled = (points[None] - origins[:, None]) / spacing
    base = np.floor(scaled - 1).astype(np.int64)
    nodes = base[..., None] + np.arange(4)
    phi = peskin4(scaled[..., None] - nodes.astype(np.float64))
    return base, phi


def expanded_component(base, phi, grid, c):
    """mac/compact_transfer.py:163-170 (+ periodic wrap): ids and weights
    (P,64), z fastest, weight=(phi_x*phi_y)*phi_z."""
    nodes = base[c][..., None] + np.arange(4)  # (P,3,4)
    shape = grid.face_shape(c)
    if grid.periodic:
        nodes = nodes % np.asarray(shape)[None, :, None]
    ids = ((nodes[:, 0, :, None, None] * shape[1] + nodes[:, 1, None, :, None])
           * shape[2] + nodes[:, 2, None, None, :])
    ph = phi[c]
    w = (ph[:, 0, :, None, None] * ph[:, 1, None, :, None]) * ph[:, 2, None, None, :]
    n = base.shape[1]
    return ids.reshape(n, 64), w.reshape(n, 64)


def spread_grid(force, base, phi, grid):
    """mac/compact_transfer.py:172-185: f_c[id] += (w*F_c)/V, chunked by
    4096 points; returns three fresh face arrays."""
    force = np.asarray(force, dtype=np.float64)
    fields = [np.zeros(grid.face_shape(c)) for c in range(3)]
    count = force.shape[0]
    vol = grid.volume
    for start in range(0, count, CHUNK):
        stop = min(start + CHUNK, count)
        sub_b, sub_p = base[:, start:stop], phi[:, start:stop]
        for c, out in enumerate(fields):
            ids, w = expanded_component(sub_b, sub_p, grid, c)
            vals = (w * force[start:stop, c, None]) / vol
            flat = out.reshape(-1)
            flat += np.bincount(ids.reshape(-1), weights=vals.reshape(-1),
                                minlength=flat.size)
    return tuple(fields)


def gather_grid(velocity, base, phi, grid):
    """mac/compact_transfer.py:187-201: U_c[p]=sum_64 u_c[id]*w."""
    count = base.shape[1]
    out = np.empty((count, 3))
    for start in range(0, count, CHUNK):
        stop = min(start + CHUNK, count)
        sub_b, sub_p = base[:, start:stop], phi[:, start:stop]
        for c, u in enumerate(velocity):
            ids, w = expanded_component(sub_b, sub_p, grid, c)
            out[start:stop, c] = (u.reshape(-1)[ids] * w).sum(-1)
    return out


def dense_operator(points, grid, c):
    """Brute-force dense H_c (P x faces): H[p,f] = prod_a phi((x_a-X_a)/h_a)
    summed over periodic images. Independent of base/floor logic; used only
    to pin the stencil formulation on tiny inputs."""
    points = np.asarray(points, dtype=np.float64)
    X = grid.coordinates(c).reshape(-1, 3)
    h = np.asarray(grid.spacing)
    H = np.ones((points.shape[0], X.shape[0]))
    for a in range(3):
        d = (points[:, None, a] - X[None, :, a]) / h[a]
        if grid.periodic:
            n = grid.shape[a]
            total = np.zeros_like(d)
            for shift in (-1, 0, 1):
                total += peskin4(d + shift * n)
            H *= total
        else:
            H *= peskin4(d)
    return H


# ---------------------------------------------------------------------------
# Nodal lattice (ib.py), the generic (P,K) indexed backend of mac/cuda_ib.py

def nodal_prepare(x, shape, spacing, origin):
    """ib.py:78-105: 64 neighbours on an x-fast nodal lattice."""
    x = np.asarray(x, dtype=np.float64)
    scaled = (x - np.asarray(origin)) / np.asarray(spacing)
    limits = np.asarray(shape, dtype=np.float64)
    if (not np.isfinite(scaled).all() or (scaled < 1).any()
            or (scaled >= limits - 2).any()):
        raise ValueError('four-point IB support leaves grid; enlarge the fluid box')
    base = np.floor(scaled - 1).astype(np.int64)
    axis = np.arange(4)
    offsets = np.stack(np.meshgrid(axis, axis, axis, indexing='ij'), -1).reshape(64, 3)
    nodes = base[:, None, :] + offsets[None]
    weights = peskin4(scaled[:, None, :] - nodes.astype(np.float64)).prod(-1)
    nx, ny, _ = shape
    indices = nodes[..., 0] + nx * (nodes[..., 1] + ny * nodes[..., 2])
    return indices, weights


def indexed_gather(field, indices, weights):
    """mac/csrc/transfer.cu:10-29 gather: out[p,c]=sum_k field[ids[p,k],c]*w."""
    return (field[indices] * weights[..., None]).sum(1)


def indexed_spread(field, indices, weights, size, volume=1.):
    """mac/csrc/transfer.cu:10-29 spread: out[j,c]+=(f[p,c]*w)/volume."""
    comps = field.shape[1]
    out = np.zeros((size, comps))
    for c in range(comps):
        vals = (field[:, None, c] * weights) / volume
        out[:, c] = np.bincount(indices.reshape(-1), weights=vals.reshape(-1),
                                minlength=size)
    return out
