"""Oracle: 3D MAC operators, geometric multigrid and projection (FP64).

Test infrastructure only (see ``oracle/__init__.py``).

Reference (``src/afsi_torch``):
  mac/grid.py:70-93          divergence, Neumann gradient, A=-D G
  mac/multigrid.py:9-21      MGOptions (rtol 1e-10, atol 1e-12, smooth 4, check 2)
  mac/multigrid.py:33-59     hierarchy, diagonal, coarse inverse inv(A+11^T/n)
  mac/multigrid.py:107-123   _smooth (Jacobi 2/3, then mean removal), _cycle
  mac/multigrid.py:125-172   solve (compatibility, mean removal, checks)
  mac/_triton_mg.py:12-118   per-axis operator order, 8-child restriction
                             order, explicit clamped trilinear prolongation
  mac/flow.py:101-120        projection rhs=-(rho/dt) D u*, u=u*-(dt/rho) G p
Periodic BCs and the red-black Gauss-Seidel smoother are build-own modes; the
tests pin them by FFT diagonalisation and a dense-matrix restatement.
"""
from dataclasses import dataclass
from math import isfinite, prod
import numpy as np


def slab(axis, start=None, stop=None):
    key = [slice(None)] * 3
    key[axis] = slice(start, stop)
    return tuple(key)


@dataclass(frozen=True)
class MGOptions:
    """mac/multigrid.py:9-21."""
    rtol: float = 1e-10
    atol: float = 1e-12
    max_cycles: int = 100
    smooth: int = 4
    check_every: int = 2
    smoother: str = 'jacobi'   # build-own: 'rbgs'
    relax: float = 1.0         # build-own: red-black over-relaxation weight


# ---------------------------------------------------------------------------
# operators

def negative_laplacian(p, spacing, periodic=False):
    """mac/grid.py:86-93 (Neumann); periodic: wrapped 7-point stencil with
    the same per-axis order (-(right diff) then +(left diff))."""
    out = np.zeros_like(p)
    for c, h in enumerate(spacing):
        h2 = h * h
        if periodic:
            d = (np.roll(p, -1, axis=c) - p) / h2
            out -= d
            out += np.roll(d, 1, axis=c)
        else:
            d = np.diff(p, axis=c) / h2
            out[slab(c, None, -1)] -= d
            out[slab(c, 1, None)] += d
    return out


def diagonal(shape, spacing, periodic=False):
    """mac/multigrid.py:38-41: per-axis 1/h^2 added for each existing
    neighbour, in the order x-low-side, x-high-side, y..., z..."""
    diag = np.zeros(shape)
    for c, h in enumerate(spacing):
        w = 1 / h ** 2
        if periodic:
            diag += w
            diag += w
        else:
            diag[slab(c, None, -1)] += w
            diag[slab(c, 1, None)] += w
    return diag


def divergence(velocity, spacing, periodic=False):
    """mac/grid.py:70-71: sum_c diff(u_c)/h_c (periodic: u_c[i+1]-u_c[i]
    with wrap, faces sized n)."""
    total = 0
    for c, (u, h) in enumerate(zip(velocity, spacing)):
        d = (np.roll(u, -1, axis=c) - u) if periodic else np.diff(u, axis=c)
        total = total + d / h
    return total


def gradient(p, spacing, periodic=False):
    """mac/grid.py:74-83: boundary-normal faces exactly zero (Neumann);
    periodic: g_c[i]=(p[i]-p[i-1])/h with wrap."""
    result = []
    for c, h in enumerate(spacing):
        if periodic:
            result.append((p - np.roll(p, 1, axis=c)) / h)
        else:
            shape = list(p.shape)
            shape[c] += 1
            g = np.zeros(shape)
            g[slab(c, 1, -1)] = np.diff(p, axis=c) / h
            result.append(g)
    return tuple(result)


def restrict(fine):
    """8-child average; child sum order (dx,dy,dz), dz fastest, then *0.125
    (mac/_triton_mg.py:52-68; equals avg_pool3d(2,2))."""
    value = np.zeros(tuple(n // 2 for n in fine.shape))
    for a in range(2):
        for b in range(2):
            for c in range(2):
                value = value + fine[a::2, b::2, c::2]
    return value * .125


def prolong(coarse, fine_shape, periodic=False):
    """Cell-centred trilinear prolongation (mac/_triton_mg.py:89-118): fine
    index i maps to coarse i/2-1/4; ax=i//2-1+i%2, weight of the upper
    coarse node .75 (even i) or .25 (odd i). Neumann clamps both ends
    (== F.interpolate(trilinear, align_corners=False)); periodic wraps."""
    idx, wts = [], []
    for n, nc in zip(fine_shape, coarse.shape):
        i = np.arange(n)
        ax = i // 2 - 1 + i % 2
        if periodic:
            i0, i1 = ax % nc, (ax + 1) % nc
        else:
            i0, i1 = np.maximum(ax, 0), np.minimum(ax + 1, nc - 1)
        idx.append((i0, i1))
        wts.append(np.where(i % 2 == 0, .75, .25))
    (x0, x1), (y0, y1), (z0, z1) = idx
    wx = wts[0][:, None, None]
    wy = wts[1][None, :, None]
    wz = wts[2][None, None, :]

    def g(xi, yi, zi):
        return coarse[np.ix_(xi, yi, zi)]
    c00 = g(x0, y0, z0) * (1. - wz) + g(x0, y0, z1) * wz
    c01 = g(x0, y1, z0) * (1. - wz) + g(x0, y1, z1) * wz
    c10 = g(x1, y0, z0) * (1. - wz) + g(x1, y0, z1) * wz
    c11 = g(x1, y1, z0) * (1. - wz) + g(x1, y1, z1) * wz
    return (c00 * (1. - wy) + c01 * wy) * (1. - wx) + (c10 * (1. - wy) + c11 * wy) * wx


def dense_matrix(shape, spacing, periodic=False):
    """Assembled A (count x count) by applying the operator to unit vectors."""
    count = prod(shape)
    A = np.empty((count, count))
    e = np.zeros(count)
    for j in range(count):
        e[j] = 1.
        A[:, j] = negative_laplacian(e.reshape(shape), spacing, periodic).reshape(-1)
        e[j] = 0.
    return A


# ---------------------------------------------------------------------------
# multigrid

def hierarchy(shape, spacing):
    """mac/multigrid.py:33-48: halve every axis until min<=4 or any odd."""
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
        raise ValueError('MAC dimensions must coarsen to <=512 cells; use powers of two')
    return shapes, spacings


class Multigrid:
    """mac/multigrid.py:24-172 (torch backend), plus build-own periodic BCs
    and red-black Gauss-Seidel smoothing."""

    def __init__(self, shape, spacing, options=None, periodic=False):
        self.options = MGOptions() if options is None else options
        self.periodic = periodic
        self.shapes, self.spacings = hierarchy(shape, spacing)
        self.diagonals = [diagonal(s, h, periodic) for s, h in zip(self.shapes, self.spacings)]
        count = prod(self.shapes[-1])
        A = dense_matrix(self.shapes[-1], self.spacings[-1], periodic)
        # mac/multigrid.py:49-59: rank-one term fixes the constant nullspace.
        self.coarse_inverse = np.linalg.inv(A + np.ones_like(A) / count)
        self.history = []

    def apply(self, level, p):
        return negative_laplacian(p, self.spacings[level], self.periodic)

    def jacobi(self, level, p, rhs):
        """One sweep p+(2/3)*(rhs-Ap)/d (mac/multigrid.py:109)."""
        return p + (2 / 3) * (rhs - self.apply(level, p)) / self.diagonals[level]

    def rbgs(self, level, p, rhs):
        """One red-black Gauss-Seidel sweep: cells with (i+j+k) even first;
        over-relaxed (p + relax * r / d) when options.relax != 1."""
        shape = p.shape
        parity = np.add.outer(np.add.outer(np.arange(shape[0]), np.arange(shape[1])),
                              np.arange(shape[2])) % 2
        relax = self.options.relax
        p = p.copy()
        for color in (0, 1):
            r = rhs - self.apply(level, p)
            mask = parity == color
            step = r[mask] / self.diagonals[level][mask]
            p[mask] = p[mask] + (step if relax == 1 else relax * step)
        return p

    def smooth(self, level, p, rhs):
        """mac/multigrid.py:107-110: nu sweeps, then subtract the mean."""
        sweep = self.rbgs if self.options.smoother == 'rbgs' else self.jacobi
        for _ in range(self.options.smooth):
            p = sweep(level, p, rhs)
        return p - p.mean()

    def cycle(self, level, p, rhs):
        """mac/multigrid.py:112-123."""
        if level == len(self.shapes) - 1:
            solution = (self.coarse_inverse @ rhs.reshape(-1)).reshape(rhs.shape)
            return solution - solution.mean()
        p = self.smooth(level, p, rhs)
        residual = rhs - self.apply(level, p)
        coarse = restrict(residual)
        coarse = coarse - coarse.mean()
        correction = self.cycle(level + 1, np.zeros_like(coarse), coarse)
        p = p + prolong(correction, p.shape, self.periodic)
        return self.smooth(level, p, rhs)

    def residual_norm(self, rhs, p):
        return float(np.linalg.norm((rhs - self.apply(0, p)).reshape(-1)))

    def solve(self, rhs, initial=None, rtol=None):
        """mac/multigrid.py:125-172. Returns (p, info); info['history'] holds
        the residual norm after every cycle (cycle 0 = start)."""
        rtol = self.options.rtol if rtol is None else rtol
        if rhs.shape != self.shapes[0]:
            raise ValueError('invalid pressure RHS')
        if not np.isfinite(rhs).all():
            raise ValueError('invalid pressure RHS')
        mean, absolute_mean = rhs.mean(), np.abs(rhs).mean()
        if abs(mean) > 1e-12 + 1e-10 * absolute_mean:
            raise ValueError('incompatible Neumann pressure RHS: net flux must vanish')
        rhs = rhs - rhs.mean()
        p = np.zeros_like(rhs) if initial is None else np.array(initial, dtype=np.float64)
        if not np.isfinite(p).all():
            raise ValueError('invalid initial pressure')
        p = p - p.mean()
        rhs_norm = float(np.linalg.norm(rhs.reshape(-1)))
        residual = self.residual_norm(rhs, p)
        if not isfinite(rhs_norm) or not isfinite(residual):
            raise RuntimeError('nonfinite multigrid residual or RHS norm')
        tolerance = max(self.options.atol, rtol * rhs_norm)
        history = [residual]
        if residual <= tolerance:
            return p, dict(cycles=0, residual_norm=residual, tolerance=tolerance, history=history)
        for cycle in range(1, self.options.max_cycles + 1):
            p = self.cycle(0, p, rhs)
            history.append(self.residual_norm(rhs, p))
            if cycle % self.options.check_every == 0 or cycle == self.options.max_cycles:
                residual = history[-1]
                if not isfinite(residual):
                    raise RuntimeError('nonfinite multigrid residual')
                if residual <= tolerance:
                    return p, dict(cycles=cycle, residual_norm=residual,
                                   tolerance=tolerance, history=history)
        raise RuntimeError(f'pressure multigrid did not converge: {residual:g} > {tolerance:g}')


# ---------------------------------------------------------------------------
# projection

def project(tentative, spacing, mg, *, dt, rho=1., initial=None):
    """mac/flow.py:101-120 (torch branch): checks, p=MG(-(rho/dt) D u*),
    u=u*-(dt/rho) G p."""
    periodic = mg.periodic
    for c, u in enumerate(tentative):
        if not np.isfinite(u).all():
            raise ValueError('nonfinite tentative velocity')
        if not periodic and (np.count_nonzero(u[slab(c, 0, 1)]) or
                             np.count_nonzero(u[slab(c, -1, None)])):
            raise ValueError('closed-box normal velocity must be zero before projection')
    rhs = -rho / dt * divergence(tentative, spacing, periodic)
    p, info = mg.solve(rhs, initial)
    grad = gradient(p, spacing, periodic)
    velocity = tuple(u - dt / rho * g for u, g in zip(tentative, grad))
    return velocity, p, info


def fft_poisson_periodic(rhs, spacing):
    """Exact discrete solution of the periodic 7-point A p = rhs (mean-free),
    by DFT diagonalisation: eigenvalues sum_a (4/h_a^2) sin^2(pi k_a/n_a)."""
    shape = rhs.shape
    lam = np.zeros(shape)
    for a, (n, h) in enumerate(zip(shape, spacing)):
        k = np.arange(n)
        e = 4. / (h * h) * np.sin(np.pi * k / n) ** 2
        lam = lam + e.reshape([-1 if b == a else 1 for b in range(3)])
    R = np.fft.fftn(rhs - rhs.mean())
    lam.flat[0] = 1.
    P = R / lam
    P.flat[0] = 0.
    return np.real(np.fft.ifftn(P))


# ---------------------------------------------------------------------------
# explicit predictor of the closed box (MACFlow.step)

COURANT_LIMIT = .25            # transport.py:9-11
VISCOUS_LIMIT = .25
ADVECTION_DIFFUSION_LIMIT = .25


def zero_normal(velocity):
    """mac/grid.py:62-67: copies with the wall-normal faces set to zero."""
    out = tuple(np.array(u, dtype=np.float64, copy=True) for u in velocity)
    for c, u in enumerate(out):
        u[slab(c, 0, 1)] = 0.
        u[slab(c, -1, None)] = 0.
    return out


def velocity_laplacian(u, component, spacing):
    """mac/grid.py:96-105: odd tangential ghosts (no slip), wall-normal rows 0;
    axes accumulated in order, ((left - 2u) + right) / h^2."""
    out = np.zeros_like(u)
    for axis, h in enumerate(spacing):
        left = np.concatenate((-u[slab(axis, 0, 1)], u[slab(axis, None, -1)]), axis)
        right = np.concatenate((u[slab(axis, 1, None)], -u[slab(axis, -1, None)]), axis)
        out += (left - 2 * u + right) / (h * h)
    out[slab(component, 0, 1)] = 0.
    out[slab(component, -1, None)] = 0.
    return out


def convection(velocity, spacing):
    """mac/grid.py:108-134: centered conservative MAC momentum flux with odd
    ghosts across the walls; wall-normal rows 0."""
    result = []
    for c, u in enumerate(velocity):
        out = np.zeros_like(u)
        for axis, h in enumerate(spacing):
            if axis == c:
                flux = np.square(.5 * (u[slab(c, 1, None)] + u[slab(c, None, -1)]))
                out[slab(c, 1, -1)] += np.diff(flux, axis=c) / h
            else:
                other = velocity[axis]
                padded = np.concatenate((-other[slab(c, 0, 1)], other, -other[slab(c, -1, None)]), c)
                advector = .5 * (padded[slab(c, 1, None)] + padded[slab(c, None, -1)])
                padded_u = np.concatenate((-u[slab(axis, 0, 1)], u, -u[slab(axis, -1, None)]), axis)
                carried = .5 * (padded_u[slab(axis, 1, None)] + padded_u[slab(axis, None, -1)])
                out += np.diff(advector * carried, axis=axis) / h
        out[slab(c, 0, 1)] = 0.
        out[slab(c, -1, None)] = 0.
        result.append(out)
    return tuple(result)


def predict(velocity, density, spacing, *, dt, rho=1., mu=1.):
    """mac/flow.py:59-64: u* = zero_normal(u + dt (-N(u) + (mu/rho) L u + f/rho))
    with u = zero_normal(velocity)."""
    vel = zero_normal(velocity)
    adv = convection(vel, spacing)
    return zero_normal(tuple(u + dt * (-a + mu / rho * velocity_laplacian(u, c, spacing) + f / rho)
                             for c, (u, a, f) in enumerate(zip(vel, adv, density))))


def transport_numbers(velocity, spacing, *, dt, rho=1., mu=1.):
    """mac/flow.py:142-146: Courant sum_c dt max|u_c|/h_c, cell Reynolds
    max_c max|u_c| h_c/(mu/rho), advection-diffusion dt sum_c max|u_c|^2 /
    (2 mu/rho)."""
    speeds = np.array([np.abs(u).max() for u in velocity])
    h = np.asarray(spacing)
    nu = mu / rho
    return (float((dt * speeds / h).sum()), float((speeds * h / nu).max()),
            float(dt * np.square(speeds).sum() / (2 * nu)))


def flow_step(velocity, density, spacing, mg, *, dt, rho=1., mu=1., initial=None):
    """MACFlow.step (mac/flow.py:122-157) on the closed box: predictor, then
    the projection; raises ValueError on the transport guards."""
    courant, reynolds, ad = transport_numbers(velocity, spacing, dt=dt, rho=rho, mu=mu)
    viscous = dt * mu / rho * sum(1 / hh ** 2 for hh in spacing)
    if courant > COURANT_LIMIT or ad > ADVECTION_DIFFUSION_LIMIT or viscous > VISCOUS_LIMIT:
        raise ValueError('MAC centered-advection stability guard exceeded')
    star = predict(velocity, density, spacing, dt=dt, rho=rho, mu=mu)
    u, p, info = project(star, spacing, mg, dt=dt, rho=rho, initial=initial)
    return u, p, dict(info, courant=courant, cell_reynolds=reynolds, advection_diffusion_number=ad)
