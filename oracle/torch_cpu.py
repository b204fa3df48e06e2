"""The reference's CPU tensor path in PyTorch (FP64), timed as the CPU baseline.

Test infrastructure only (see ``oracle/__init__.py``): ``bench.py``'s
``cpu_baseline`` leg times it on the host's cores, ``tests/test_oracle_torch_cpu.py``
pins it against the NumPy oracle. It follows the reference's torch branches
op for op, so its cost is the cost of the path a CPU caller runs:

  mac/compact_transfer.py:21-26    _prepare: scaled=(x-o)/h, base=floor(scaled-1),
                                   phi = peskin4(scaled - (base + 0..3))
  ib.py:57-67                      peskin4 (two clamped branches)
  mac/compact_transfer.py:163-185  spread: chunks of 4096 points, 64 expanded
                                   links (z fastest), index_add_ of (w*F_c)/V
  mac/compact_transfer.py:187-201  gather: chunks, u_c.flatten()[ids]*w summed
  mac/multigrid.py:107-123         V-cycle: Jacobi(2/3) sweeps, mean removal,
                                   residual, 2x2x2 average, coarse solve with
                                   inv(A + 11^T/n), trilinear prolongation

The periodic lattice (wrapped node indices, rolled Laplacian, wrapped
trilinear prolongation) is the build-own periodic variant of those formulas,
as in ``oracle.ib3d`` / ``oracle.mac3d``.
"""
import torch
import torch.nn.functional as F

CHUNK = 4096


def peskin4(r):
    """ib.py:57-67."""
    r = r.abs()
    a = torch.clamp(r, max=1.)
    b = torch.clamp(r, min=1., max=2.)
    inner = (3. - 2. * a + torch.sqrt(1. + 4. * a - 4. * a * a)) / 8.
    outer = (5. - 2. * b - torch.sqrt(-7. + 12. * b - 4. * b * b)) / 8.
    return torch.where(r < 1., inner, torch.where(r < 2., outer, torch.zeros_like(r)))


def prepare(points, shape, spacing):
    """mac/compact_transfer.py:21-26 for the three face lattices of a grid
    with origin 0: base (3,P,3) int64, phi (3,P,3,4)."""
    h = torch.tensor(spacing, dtype=torch.float64)
    origins = torch.stack([torch.tensor([0. if a == c else .5 * spacing[a] for a in range(3)],
                                        dtype=torch.float64) for c in range(3)])
    scaled = (points[None] - origins[:, None]) / h
    base = torch.floor(scaled - 1.).to(torch.int64)
    nodes = base[..., None] + torch.arange(4)
    return base, peskin4(scaled[..., None] - nodes.to(torch.float64))


def _links(base, phi, c, shape):
    """mac/compact_transfer.py:163-170 (+ periodic wrap): (P,64) ids, weights."""
    nodes = (base[c][..., None] + torch.arange(4)) % torch.tensor(shape)[None, :, None]
    ids = ((nodes[:, 0, :, None, None] * shape[1] + nodes[:, 1, None, :, None]) * shape[2]
           + nodes[:, 2, None, None, :])
    ph = phi[c]
    w = (ph[:, 0, :, None, None] * ph[:, 1, None, :, None]) * ph[:, 2, None, None, :]
    return ids.reshape(-1, 64), w.reshape(-1, 64)


def spread(force, base, phi, shape, volume):
    """mac/compact_transfer.py:172-185 on a periodic grid (faces n^3 each)."""
    out = [torch.zeros(shape, dtype=torch.float64).reshape(-1) for _ in range(3)]
    for start in range(0, force.shape[0], CHUNK):
        stop = min(start + CHUNK, force.shape[0])
        for c in range(3):
            ids, w = _links(base[:, start:stop], phi[:, start:stop], c, shape)
            out[c].index_add_(0, ids.reshape(-1), ((w * force[start:stop, c, None]) / volume).reshape(-1))
    return tuple(o.reshape(shape) for o in out)


def gather(velocity, base, phi, shape):
    """mac/compact_transfer.py:187-201."""
    out = torch.empty((base.shape[1], 3), dtype=torch.float64)
    for start in range(0, base.shape[1], CHUNK):
        stop = min(start + CHUNK, base.shape[1])
        for c in range(3):
            ids, w = _links(base[:, start:stop], phi[:, start:stop], c, shape)
            out[start:stop, c] = (velocity[c].reshape(-1)[ids] * w).sum(-1)
    return out


def apply(p, spacing):
    """Periodic 7-point -Laplacian (mac/grid.py:86-93 with wrapped neighbours)."""
    out = torch.zeros_like(p)
    for a, h in enumerate(spacing):
        out += (2. * p - torch.roll(p, 1, a) - torch.roll(p, -1, a)) / (h * h)
    return out


def prolong(coarse):
    """Trilinear prolongation of a periodic cell field (align_corners=False
    weights 3/4, 1/4 along each axis, neighbours wrapped)."""
    fine = coarse
    for a in range(3):
        lo = .75 * fine + .25 * torch.roll(fine, 1, a)
        hi = .75 * fine + .25 * torch.roll(fine, -1, a)
        fine = torch.stack((lo, hi), a + 1).flatten(a, a + 1)
    return fine


class JacobiMultigrid:
    """mac/multigrid.py:24-123: Jacobi(2/3) V(nu,nu), periodic lattice."""

    def __init__(self, shape, spacing, smooth=4):
        self.smooth_count = smooth
        self.shapes, self.spacings = [tuple(shape)], [tuple(spacing)]
        while min(self.shapes[-1]) > 4 and not any(n % 2 for n in self.shapes[-1]):
            self.shapes.append(tuple(n // 2 for n in self.shapes[-1]))
            self.spacings.append(tuple(2 * h for h in self.spacings[-1]))
        self.diagonals = [sum(2. / (h * h) for h in s) for s in self.spacings]
        n = 1
        for k in self.shapes[-1]:
            n *= k
        eye = torch.eye(n, dtype=torch.float64).reshape((n,) + self.shapes[-1])
        A = torch.stack([apply(e, self.spacings[-1]).reshape(-1) for e in eye])
        self.coarse_inverse = torch.linalg.inv(A + torch.ones_like(A) / n)

    def smooth(self, level, p, rhs):
        """mac/multigrid.py:107-110."""
        for _ in range(self.smooth_count):
            p = p + (2. / 3.) * (rhs - apply(p, self.spacings[level])) / self.diagonals[level]
        return p - p.mean()

    def cycle(self, level, p, rhs):
        """mac/multigrid.py:112-123."""
        if level == len(self.shapes) - 1:
            s = (self.coarse_inverse @ rhs.reshape(-1)).reshape(rhs.shape)
            return s - s.mean()
        p = self.smooth(level, p, rhs)
        r = rhs - apply(p, self.spacings[level])
        coarse = F.avg_pool3d(r[None, None], 2)[0, 0]
        coarse = coarse - coarse.mean()
        p = p + prolong(self.cycle(level + 1, torch.zeros_like(coarse), coarse))
        return self.smooth(level, p, rhs)

    def solve(self, rhs, initial=None, rtol=1e-10, atol=1e-12, check_every=2, max_cycles=200):
        """mac/multigrid.py:125-172: stop when ||b - A p|| <= max(atol,
        rtol ||b||), checked every ``check_every`` cycles."""
        rhs = rhs - rhs.mean()
        p = torch.zeros_like(rhs) if initial is None else initial - initial.mean()
        tol = max(atol, rtol * float(torch.linalg.vector_norm(rhs)))
        for cycle in range(1, max_cycles + 1):
            p = self.cycle(0, p, rhs)
            if cycle % check_every == 0:
                if float(torch.linalg.vector_norm(rhs - apply(p, self.spacings[0]))) <= tol:
                    return p, cycle
        raise RuntimeError('pressure multigrid did not converge')


def divergence(u, spacing):
    """mac/grid.py:70-71, periodic: sum_c (u_c[i+1] - u_c[i]) / h_c."""
    return sum((torch.roll(uc, -1, c) - uc) / h for c, (uc, h) in enumerate(zip(u, spacing)))


def gradient(p, spacing):
    """mac/grid.py:74-83, periodic: g_c[i] = (p[i] - p[i-1]) / h_c."""
    return tuple((p - torch.roll(p, 1, c)) / h for c, h in enumerate(spacing))


def ib_step(x, force, u, p_prev, shape, spacing, mg, *, dt, rho=1.):
    """One step in the order of mac/coupling.py:74-85 with the force term as
    the predictor: prepare, spread, u* = u + (dt/rho) f, p = MG(-(rho/dt)
    D u*, p_prev) (mac/flow.py:101-120), u = u* - (dt/rho) G p, gather,
    x + dt U. Returns (x_new, u, p, U, cycles)."""
    volume = spacing[0] * spacing[1] * spacing[2]
    base, phi = prepare(x, shape, spacing)
    f = spread(force, base, phi, shape, volume)
    star = tuple(uc + dt / rho * fc for uc, fc in zip(u, f))
    p, cycles = mg.solve(-rho / dt * divergence(star, spacing), p_prev)
    velocity = tuple(s - dt / rho * g for s, g in zip(star, gradient(p, spacing)))
    U = gather(velocity, base, phi, shape)
    return x + dt * U, velocity, p, U, cycles
