"""Oracle: 2D staggered channel (inlet, no-slip walls, p=0 outlet), FP64.

Test infrastructure only (see ``oracle/__init__.py``).

Reference (``src/afsi_torch``):
  mac2d/grid.py:7-33         ChannelGrid spacing/face_shape/coordinates
  mac2d/grid.py:46-71        divergence, gradient (outlet gx[-1]=-2p/hx),
                             A with the outlet term +2p/hx^2
  mac2d/multigrid.py:16-45   hierarchy, diagonal (+2/hx^2 at the outlet row),
                             exact coarse inverse (no nullspace)
  mac2d/multigrid.py:127-130 smooth_block (Jacobi 2/3)
  mac2d/multigrid.py:186-228 _cycle (no mean removal), solve
  mac2d/_triton_mg.py:7-75   per-axis operator order, restriction order,
                             ghost-based bilinear prolongation (even ghosts at
                             Neumann sides, odd ghost at the outlet)
  mac2d/flow.py:72-86        projection
  mac2d/transfer.py:61-127   16-link reflected stencil, spread and gather
"""
from dataclasses import dataclass
from math import isfinite, prod
import numpy as np
from .ib3d import peskin4
from .mac3d import MGOptions


@dataclass(frozen=True)
class Channel:
    """mac2d/grid.py:7-33."""
    shape: tuple = (256, 64)
    lengths: tuple = (8., 1.61)

    @property
    def spacing(self):
        return tuple(L / n for L, n in zip(self.lengths, self.shape))

    @property
    def volume(self):
        hx, hy = self.spacing
        return hx * hy

    def face_shape(self, c):
        return tuple(n + (a == c) for a, n in enumerate(self.shape))

    def coordinates(self, c=None):
        shape = self.shape if c is None else self.face_shape(c)
        axes = [h * (np.arange(n, dtype=np.float64) + (0 if c == a else .5))
                for a, (n, h) in enumerate(zip(shape, self.spacing))]
        return np.stack(np.meshgrid(*axes, indexing='ij'), -1)


def negative_laplacian(p, spacing):
    """mac2d/grid.py:61-71."""
    hx, hy = spacing
    out = np.zeros_like(p)
    dx = (p[1:] - p[:-1]) / hx ** 2
    dy = (p[:, 1:] - p[:, :-1]) / hy ** 2
    out[:-1] -= dx
    out[1:] += dx
    out[:, :-1] -= dy
    out[:, 1:] += dy
    out[-1] += 2 * p[-1] / hx ** 2
    return out


def divergence(velocity, spacing):
    """mac2d/grid.py:46-48."""
    u, v = velocity
    return np.diff(u, axis=0) / spacing[0] + np.diff(v, axis=1) / spacing[1]


def gradient(p, spacing):
    """mac2d/grid.py:51-58."""
    hx, hy = spacing
    gx = np.zeros((p.shape[0] + 1, p.shape[1]))
    gy = np.zeros((p.shape[0], p.shape[1] + 1))
    gx[1:-1] = (p[1:] - p[:-1]) / hx
    gx[-1] = -2 * p[-1] / hx
    gy[:, 1:-1] = (p[:, 1:] - p[:, :-1]) / hy
    return gx, gy


def diagonal(shape, spacing):
    """mac2d/multigrid.py:28-33 (the in-loop diag construction)."""
    hx, hy = spacing
    diag = np.zeros(shape)
    diag[:-1] += 1 / hx ** 2
    diag[1:] += 1 / hx ** 2
    diag[:, :-1] += 1 / hy ** 2
    diag[:, 1:] += 1 / hy ** 2
    diag[-1] += 2 / hx ** 2
    return diag


def restrict(fine):
    """((r00+r01)+r10)+r11, *0.25 (mac2d/_triton_mg.py:393-408)."""
    return (((fine[0::2, 0::2] + fine[0::2, 1::2]) + fine[1::2, 0::2])
            + fine[1::2, 1::2]) * .25


def prolong(coarse, fine_shape):
    """mac2d/_triton_mg.py:411-432: ghost values: even (clamped) on the
    inlet/walls, odd (negated) beyond the outlet row."""
    NX, NY = coarse.shape

    def value(x, y):
        sign = np.where(x >= NX, -1., 1.)
        xx = np.clip(x, 0, NX - 1)
        yy = np.clip(y, 0, NY - 1)
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
    """mac2d/multigrid.py:21-36."""
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


class ChannelMultigrid:
    """mac2d/multigrid.py:16-228 (reference backend)."""

    def __init__(self, shape, spacing, options=None):
        self.options = MGOptions() if options is None else options
        self.shapes, self.spacings = hierarchy(shape, spacing)
        self.diagonals = [diagonal(s, h) for s, h in zip(self.shapes, self.spacings)]
        count = prod(self.shapes[-1])
        eye = np.eye(count)
        A = np.stack([negative_laplacian(v.reshape(self.shapes[-1]), self.spacings[-1]).reshape(-1)
                      for v in eye], 1)
        self.inverse = np.linalg.inv(A)

    def smooth(self, level, p, rhs):
        for _ in range(self.options.smooth):
            p = p + (2 / 3) * (rhs - negative_laplacian(p, self.spacings[level])) / self.diagonals[level]
        return p

    def cycle(self, level, p, rhs):
        if level == len(self.shapes) - 1:
            return (self.inverse @ rhs.reshape(-1)).reshape(rhs.shape)
        p = self.smooth(level, p, rhs)
        residual = rhs - negative_laplacian(p, self.spacings[level])
        coarse = restrict(residual)
        correction = self.cycle(level + 1, np.zeros_like(coarse), coarse)
        p = p + prolong(correction, p.shape)
        return self.smooth(level, p, rhs)

    def residual_norm(self, rhs, p):
        return float(np.linalg.norm((rhs - negative_laplacian(p, self.spacings[0])).reshape(-1)))

    def solve(self, rhs, initial=None):
        if rhs.shape != self.shapes[0]:
            raise ValueError('invalid channel pressure RHS')
        p = np.zeros_like(rhs) if initial is None else np.array(initial, dtype=np.float64)
        if not np.isfinite(rhs).all() or not np.isfinite(p).all():
            raise ValueError('finite pressure RHS/initial guess required')
        rhs_norm = float(np.linalg.norm(rhs.reshape(-1)))
        residual = self.residual_norm(rhs, p)
        if not isfinite(rhs_norm) or not isfinite(residual):
            raise RuntimeError('nonfinite channel pressure residual or RHS norm')
        tol = max(self.options.atol, self.options.rtol * rhs_norm)
        history = [residual]
        if residual <= tol:
            return p, dict(cycles=0, residual_norm=residual, tolerance=tol, history=history)
        for cycle in range(1, self.options.max_cycles + 1):
            p = self.cycle(0, p, rhs)
            history.append(self.residual_norm(rhs, p))
            if cycle % self.options.check_every == 0 or cycle == self.options.max_cycles:
                residual = history[-1]
                if not isfinite(residual):
                    raise RuntimeError('nonfinite channel pressure residual')
                if residual <= tol:
                    return p, dict(cycles=cycle, residual_norm=residual, tolerance=tol, history=history)
        raise RuntimeError(f'channel MG did not converge: {residual:g} > {tol:g}')


def project(tentative, spacing, mg, *, dt, rho=1., initial=None):
    """mac2d/flow.py:210-225 (reference branch)."""
    if any(not np.isfinite(v).all() for v in tentative):
        raise ValueError('finite tentative velocity required')
    if np.count_nonzero(tentative[1][:, 0]) or np.count_nonzero(tentative[1][:, -1]):
        raise ValueError('normal wall velocities must be zero')
    p, info = mg.solve(-rho / dt * divergence(tentative, spacing), initial)
    gx, gy = gradient(p, spacing)
    u, v = tentative
    return (u - dt / rho * gx, v - dt / rho * gy), p, info


# ---------------------------------------------------------------------------
# wall-reflected 16-link transfer

def support_ok(points, channel):
    """mac2d/transfer.py:88-95."""
    hx, hy = channel.spacing
    return bool(np.isfinite(points).all() and (points[:, 0] >= 2 * hx).all()
                and (points[:, 0] <= channel.lengths[0] - 2 * hx).all()
                and (points[:, 1] >= -hy).all() and (points[:, 1] <= channel.lengths[1] + hy).all())


def prepare(points, channel):
    """mac2d/transfer.py:61-81: per component (P,16) ids and signed weights.
    Rows outside the walls reflect: u: j<0 -> -j-1, j>=ny -> 2ny-j-1;
    v: j<0 -> -j, j>ny -> 2ny-j; reflected weights are negated."""
    points = np.asarray(points, dtype=np.float64)
    h = np.asarray(channel.spacing)
    ny = channel.shape[1]
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
        ids = nodes[..., 0] * channel.face_shape(c)[1] + reflected
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
