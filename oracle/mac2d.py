"""Oracle: 2D staggered channel (inlet, no-slip walls, p=0 outlet), FP64.

Test infrastructure only (see ``oracle/__init__.py``).

Reference (``src/afsi_torch``):
  mac2d/grid.py:7-33         ChannelGrid spacing/face_shape/coordinates
  mac2d/grid.py:46-71        divergence, gradient (outlet gx[-1]=-2p/hx),
                             A with the outlet term +2p/hx^2
  mac2d/multigrid.py:10-13   smooth_block (Jacobi 2/3)
  mac2d/multigrid.py:17-45   hierarchy, diagonal (+2/hx^2 at the outlet row),
                             exact coarse inverse (no nullspace)
  mac2d/multigrid.py:65-81   _cycle (no mean removal; ghost prolongation)
  mac2d/multigrid.py:83-109  solve (tolerance, check_every, messages)
  mac2d/_triton_mg.py:7-75   per-axis operator order, restriction order,
                             ghost-based bilinear prolongation (even ghosts at
                             Neumann sides, odd ghost at the outlet)
  mac2d/flow.py:92-103       projection
  mac2d/transfer.py:61-127   16-link reflected stencil, spread and gather

Build-own modes (no reference counterpart):
  periodic=True   periodic in x (u has nx faces), no-slip walls in y; the
                  operator is singular, so the cycle removes means the way the
                  3D closed box does (mac/multigrid.py:107-123) and the
                  coarsest level uses inv(A + 11^T/n);
  smoother='rbgs' red-black Gauss-Seidel, cells with (i+j) even first.
"""
from dataclasses import dataclass
from math import isfinite, prod
import numpy as np
from .ib3d import peskin4
from .mac3d import MGOptions


@dataclass(frozen=True)
class Channel:
    """mac2d/grid.py:7-33 (+ build-own ``periodic``: periodic in x)."""
    shape: tuple = (256, 64)
    lengths: tuple = (8., 1.61)
    periodic: bool = False

    @property
    def spacing(self):
        return tuple(L / n for L, n in zip(self.lengths, self.shape))

    @property
    def volume(self):
        hx, hy = self.spacing
        return hx * hy

    def face_shape(self, c):
        if self.periodic and c == 0:
            return tuple(self.shape)
        return tuple(n + (a == c) for a, n in enumerate(self.shape))

    def coordinates(self, c=None):
        shape = self.shape if c is None else self.face_shape(c)
        axes = [h * (np.arange(n, dtype=np.float64) + (0 if c == a else .5))
                for a, (n, h) in enumerate(zip(shape, self.spacing))]
        return np.stack(np.meshgrid(*axes, indexing='ij'), -1)


def negative_laplacian(p, spacing, periodic=False):
    """mac2d/grid.py:61-71; periodic: wrapped x differences, same order
    (-(right diff) then +(left diff) per axis), no outlet term."""
    hx, hy = spacing
    out = np.zeros_like(p)
    if periodic:
        dx = (np.roll(p, -1, axis=0) - p) / hx ** 2
        out -= dx
        out += np.roll(dx, 1, axis=0)
    else:
        dx = (p[1:] - p[:-1]) / hx ** 2
        out[:-1] -= dx
        out[1:] += dx
    dy = (p[:, 1:] - p[:, :-1]) / hy ** 2
    out[:, :-1] -= dy
    out[:, 1:] += dy
    if not periodic:
        out[-1] += 2 * p[-1] / hx ** 2
    return out


def divergence(velocity, spacing, periodic=False):
    """mac2d/grid.py:46-48 (periodic: u[i+1 mod nx] - u[i])."""
    u, v = velocity
    du = (np.roll(u, -1, axis=0) - u) if periodic else np.diff(u, axis=0)
    return du / spacing[0] + np.diff(v, axis=1) / spacing[1]


def gradient(p, spacing, periodic=False):
    """mac2d/grid.py:51-58 (periodic: gx[i] = (p[i] - p[i-1 mod nx])/hx)."""
    hx, hy = spacing
    if periodic:
        gx = (p - np.roll(p, 1, axis=0)) / hx
    else:
        gx = np.zeros((p.shape[0] + 1, p.shape[1]))
        gx[1:-1] = (p[1:] - p[:-1]) / hx
        gx[-1] = -2 * p[-1] / hx
    gy = np.zeros((p.shape[0], p.shape[1] + 1))
    gy[:, 1:-1] = (p[:, 1:] - p[:, :-1]) / hy
    return gx, gy


def diagonal(shape, spacing, periodic=False):
    """mac2d/multigrid.py:27-33 (the in-loop diag construction)."""
    hx, hy = spacing
    diag = np.zeros(shape)
    if periodic:
        diag += 1 / hx ** 2
        diag += 1 / hx ** 2
    else:
        diag[:-1] += 1 / hx ** 2
        diag[1:] += 1 / hx ** 2
    diag[:, :-1] += 1 / hy ** 2
    diag[:, 1:] += 1 / hy ** 2
    if not periodic:
        diag[-1] += 2 / hx ** 2
    return diag


def restrict(fine):
    """((r00+r01)+r10)+r11, *0.25 (mac2d/_triton_mg.py:40-51)."""
    return (((fine[0::2, 0::2] + fine[0::2, 1::2]) + fine[1::2, 0::2])
            + fine[1::2, 1::2]) * .25


def prolong(coarse, fine_shape, periodic=False):
    """mac2d/_triton_mg.py:54-75: ghost values: even (clamped) on the
    inlet/walls, odd (negated) beyond the outlet row; periodic: wrapped x."""
    NX, NY = coarse.shape

    def value(x, y):
        yy = np.clip(y, 0, NY - 1)
        if periodic:
            return coarse[np.mod(x, NX), yy]
        sign = np.where(x >= NX, -1., 1.)
        xx = np.clip(x, 0, NX - 1)
        return sign * coarse[xx, yy]
    x = np.arange(fine_shape[0])[:, None]
    y = np.arange(fine_shape[1])[None, :]
    ax, ay = x // 2 - 1 + x % 2, y // 2 - 1 + y % 2
    wx = np.where(x % 2 == 0, .75, .25)
    wy = np.where(y % 2 == 0, .75, .25)
    c00, c01 = value(ax, ay), value(ax, ay + 1)
    c10, c11 = value(ax + 1, ay), value(ax + 1, ay + 1)
    return (c00 * (1 - wy) + c01 * wy) * (1 - wx) + (c10 * (1 - wy) + c11 * wy) * wx


def hierarchy(shape, spacing):
    """mac2d/multigrid.py:21-40."""
    shapes, spacings = [], []
    shape, spacing = tuple(shape), tuple(spacing)
    while True:
        shapes.append(shape)
        spacings.append(spacing)
        if min(shape) <= 4 or any(n % 2 for n in shape):
            break
        shape = tuple(n // 2 for n in shape)
        spacing = tuple(2 * h for h in spacing)
    if prod(shape) > 512:
        raise ValueError('channel mesh must coarsen to <=512 cells')
    return shapes, spacings


def dense_matrix(shape, spacing, periodic=False):
    count = prod(shape)
    eye = np.eye(count)
    return np.stack([negative_laplacian(v.reshape(shape), spacing, periodic).reshape(-1)
                     for v in eye], 1)


class ChannelMultigrid:
    """mac2d/multigrid.py:16-109 (reference backend), plus the build-own
    periodic-x mode and red-black Gauss-Seidel smoother."""

    def __init__(self, shape, spacing, options=None, periodic=False):
        self.options = MGOptions() if options is None else options
        self.periodic = periodic
        self.shapes, self.spacings = hierarchy(shape, spacing)
        self.diagonals = [diagonal(s, h, periodic) for s, h in zip(self.shapes, self.spacings)]
        A = dense_matrix(self.shapes[-1], self.spacings[-1], periodic)
        if periodic:
            A = A + np.ones_like(A) / A.shape[0]
        self.inverse = np.linalg.inv(A)

    def apply(self, level, p):
        return negative_laplacian(p, self.spacings[level], self.periodic)

    def jacobi(self, level, p, rhs):
        """mac2d/multigrid.py:10-13 (one iteration)."""
        return p + (2 / 3) * (rhs - self.apply(level, p)) / self.diagonals[level]

    def rbgs(self, level, p, rhs):
        """One red-black sweep: p += r/diag on (i+j) even cells, then odd."""
        parity = np.add.outer(np.arange(p.shape[0]), np.arange(p.shape[1])) % 2
        p = p.copy()
        for color in (0, 1):
            r = rhs - self.apply(level, p)
            mask = parity == color
            p[mask] = p[mask] + r[mask] / self.diagonals[level][mask]
        return p

    def smooth(self, level, p, rhs):
        sweep = self.rbgs if self.options.smoother == 'rbgs' else self.jacobi
        for _ in range(self.options.smooth):
            p = sweep(level, p, rhs)
        return p - p.mean() if self.periodic else p

    def cycle(self, level, p, rhs):
        """mac2d/multigrid.py:65-81 (periodic: means removed as in
        mac/multigrid.py:107-123)."""
        if level == len(self.shapes) - 1:
            solution = (self.inverse @ rhs.reshape(-1)).reshape(rhs.shape)
            return solution - solution.mean() if self.periodic else solution
        p = self.smooth(level, p, rhs)
        residual = rhs - self.apply(level, p)
        coarse = restrict(residual)
        if self.periodic:
            coarse = coarse - coarse.mean()
        correction = self.cycle(level + 1, np.zeros_like(coarse), coarse)
        p = p + prolong(correction, p.shape, self.periodic)
        return self.smooth(level, p, rhs)

    def residual_norm(self, rhs, p):
        return float(np.linalg.norm((rhs - self.apply(0, p)).reshape(-1)))

    def solve(self, rhs, initial=None, history=None):
        """mac2d/multigrid.py:83-109; periodic: the rhs must be compatible
        (mean within 1e-10 of its magnitude) and its mean is removed."""
        if rhs.shape != self.shapes[0]:
            raise ValueError('invalid channel pressure RHS')
        p = np.zeros_like(rhs) if initial is None else np.array(initial, dtype=np.float64)
        if not np.isfinite(rhs).all() or not np.isfinite(p).all():
            raise ValueError('finite pressure RHS/initial guess required')
        if self.periodic:
            if abs(rhs.mean()) > 1e-10 * max(1., np.abs(rhs).max()):
                raise ValueError('periodic channel pressure RHS is incompatible')
            rhs = rhs - rhs.mean()
            p = p - p.mean()
        rhs_norm = float(np.linalg.norm(rhs.reshape(-1)))
        residual = self.residual_norm(rhs, p)
        if not isfinite(rhs_norm) or not isfinite(residual):
            raise RuntimeError('nonfinite channel pressure residual or RHS norm')
        tol = max(self.options.atol, self.options.rtol * rhs_norm)
        if history is not None:
            history.append(residual)
        if residual <= tol:
            return p, dict(cycles=0, residual_norm=residual, tolerance=tol)
        for cycle in range(1, self.options.max_cycles + 1):
            p = self.cycle(0, p, rhs)
            if history is not None:
                history.append(self.residual_norm(rhs, p))
            if cycle % self.options.check_every == 0 or cycle == self.options.max_cycles:
                residual = self.residual_norm(rhs, p)
                if not isfinite(residual):
                    raise RuntimeError('nonfinite channel pressure residual')
                if residual <= tol:
                    return p, dict(cycles=cycle, residual_norm=residual, tolerance=tol)
        raise RuntimeError(f'channel MG did not converge: {residual:g} > {tol:g}')


def project(tentative, spacing, mg, *, dt, rho=1., initial=None):
    """mac2d/flow.py:92-103 (reference branch)."""
    if any(not np.isfinite(v).all() for v in tentative):
        raise ValueError('finite tentative velocity required')
    if np.count_nonzero(tentative[1][:, 0]) or np.count_nonzero(tentative[1][:, -1]):
        raise ValueError('normal wall velocities must be zero')
    periodic = mg.periodic
    p, info = mg.solve(-rho / dt * divergence(tentative, spacing, periodic), initial)
    gx, gy = gradient(p, spacing, periodic)
    u, v = tentative
    return (u - dt / rho * gx, v - dt / rho * gy), p, info


# ---------------------------------------------------------------------------
# wall-reflected 16-link transfer

def support_ok(points, channel):
    """mac2d/transfer.py:88-95 (periodic: the x range is not restricted)."""
    hx, hy = channel.spacing
    ok = np.isfinite(points).all() and (points[:, 1] >= -hy).all() and \
        (points[:, 1] <= channel.lengths[1] + hy).all()
    if not channel.periodic:
        ok = ok and (points[:, 0] >= 2 * hx).all() and \
            (points[:, 0] <= channel.lengths[0] - 2 * hx).all()
    return bool(ok)


def prepare(points, channel):
    """mac2d/transfer.py:61-81: per component (P,16) ids and signed weights.
    Rows outside the walls reflect: u: j<0 -> -j-1, j>=ny -> 2ny-j-1;
    v: j<0 -> -j, j>ny -> 2ny-j; reflected weights are negated.
    Periodic: x node indices wrap modulo nx."""
    points = np.asarray(points, dtype=np.float64)
    h = np.asarray(channel.spacing)
    nx, ny = channel.shape
    offsets = np.stack(np.meshgrid(np.arange(4), np.arange(4), indexing='ij'), -1).reshape(16, 2)
    indices, weights = [], []
    for c in range(2):
        origin = h * np.asarray([0 if c == 0 else .5, 0 if c == 1 else .5])
        scaled = (points - origin) / h
        nodes = np.floor(scaled - 1).astype(np.int64)[:, None] + offsets[None]
        w = peskin4(scaled[:, None] - nodes.astype(np.float64)).prod(-1)
        j = nodes[..., 1]
        if c == 0:
            outside = (j < 0) | (j >= ny)
            reflected = np.where(j < 0, -j - 1, np.where(j >= ny, 2 * ny - j - 1, j))
        else:
            outside = (j < 0) | (j > ny)
            reflected = np.where(j < 0, -j, np.where(j > ny, 2 * ny - j, j))
        w = np.where(outside, -w, w)
        i = np.mod(nodes[..., 0], nx) if channel.periodic else nodes[..., 0]
        ids = i * channel.face_shape(c)[1] + reflected
        indices.append(ids)
        weights.append(w)
    return tuple(indices), tuple(weights)


def spread(force, indices, weights, channel):
    """mac2d/transfer.py:108-115 (point force given directly)."""
    out = []
    for c in range(2):
        size = prod(channel.face_shape(c))
        vals = (force[:, c, None] * weights[c]) / channel.volume
        out.append(np.bincount(indices[c].reshape(-1), weights=vals.reshape(-1),
                               minlength=size).reshape(channel.face_shape(c)))
    return tuple(out)


def gather(velocity, indices, weights):
    """mac2d/transfer.py:123-127."""
    return np.stack([(u.reshape(-1)[ids] * w).sum(-1)
                     for u, ids, w in zip(velocity, indices, weights)], -1)


def dense_transfer(points, channel):
    """Brute-force (P, faces) operator per component from distances to every
    face and its wall mirror images (odd reflection), with periodic x images;
    independent of the floor/offset/index logic of ``prepare``."""
    h = np.asarray(channel.spacing)
    Ly = channel.lengths[1]
    out = []
    for c in range(2):
        coords = channel.coordinates(c).reshape(-1, 2)
        H = np.zeros((len(points), len(coords)))
        x_shifts = (-channel.lengths[0], 0., channel.lengths[0]) if channel.periodic else (0.,)
        # mirror images of a face across y=0 and y=Ly carry a minus sign
        for sx in x_shifts:
            for img, sign in ((lambda y: y, 1.), (lambda y: -y, -1.), (lambda y: 2 * Ly - y, -1.)):
                fx = coords[:, 0][None] + sx
                fy = img(coords[:, 1])[None]
                if c == 1:
                    # v faces on the walls are their own mirror images; the
                    # reference reflects only rows strictly outside (j<0, j>ny)
                    on_wall = (np.abs(coords[:, 1]) < 1e-12 * Ly) | \
                        (np.abs(coords[:, 1] - Ly) < 1e-12 * Ly)
                    if sign < 0:
                        fy = np.where(on_wall[None], np.inf, fy)
                rx = (points[:, 0:1] - fx) / h[0]
                ry = (points[:, 1:2] - fy) / h[1]
                H += sign * peskin4(rx) * peskin4(ry)
        out.append(H)
    return out
