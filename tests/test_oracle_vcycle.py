"""Pin the composition of the oracle's V-cycles (oracle/mac3d.Multigrid.cycle,
oracle/mac2d.ChannelMultigrid.cycle) against one cycle formed as a dense
affine map p' = M p + N rhs.

The map is composed here from matrices assembled independently of the oracle:
  * A: scalar loops over cells and their existing neighbours (closed box /
    periodic 7-point, channel 5-point with the Dirichlet outlet ghost
    -p, mac2d/grid.py:61-71);
  * D = diag(A); Jacobi (2/3) and red-black sweeps as matrices;
  * restriction: Kronecker product of 1D pair averages (avg_pool, 2x2x2);
  * prolongation: Kronecker product of 1D cell-centred linear interpolation
    at source index i/2 - 1/4 with even ghosts (closed sides), the odd outlet
    ghost (2D channel, mac2d/multigrid.py:76-81) or a periodic wrap;
  * mean projector Q = I - 11^T/n (mac/multigrid.py:107-123: after every
    smoothing stage, of every restricted residual and of the coarse solution);
  * coarse inverse (A + 11^T/n)^-1 (singular operators) or A^-1 (channel).
The cycle order is mac/multigrid.py:112-123 / mac2d/multigrid.py:69-83:
pre-smooth, residual, restrict, recurse from zero, prolong-add, post-smooth.
Random p and rhs (not mean-free) make every term count. Mistakes planted in
subclasses of the oracle (a dropped post-smooth, no coarse mean removal, a
doubled correction, swapped red/black order, a mirrored prolongation index)
must each fail the comparison.
"""
from itertools import product
from math import prod

import numpy as np
import pytest
from scipy import sparse

from oracle import mac2d, mac3d
from oracle.mac3d import MGOptions

TOL = 1e-12


def assemble_operator(shape, spacing, mode):
    """A by scalar loops. mode: 'neumann' | 'periodic' (3D or 2D-periodic-x
    with walls in y when len(shape)==2) | 'channel' (2D reference)."""
    count = prod(shape)
    A = np.zeros((count, count))
    strides = [prod(shape[a + 1:]) for a in range(len(shape))]
    for idx in np.ndindex(*shape):
        row = sum(i * s for i, s in zip(idx, strides))
        for axis, h in enumerate(spacing):
            w = 1.0 / (h * h)
            wraps = mode == 'periodic' and (len(shape) == 3 or axis == 0)
            for side in (-1, 1):
                nb = list(idx)
                nb[axis] += side
                if wraps:
                    nb[axis] %= shape[axis]
                elif not 0 <= nb[axis] < shape[axis]:
                    if mode == 'channel' and axis == 0 and side == 1:
                        A[row, row] += 2 * w   # ghost -p beyond the outlet
                    continue
                col = sum(i * s for i, s in zip(nb, strides))
                A[row, row] += w
                A[row, col] -= w
    return A


def restriction_1d(n):
    R = np.zeros((n // 2, n))
    for i in range(n // 2):
        R[i, 2 * i] = R[i, 2 * i + 1] = .5
    return R


def prolongation_1d(nc, ghost):
    """(2 nc, nc) linear interpolation at coarse coordinate i/2 - 1/4.
    ghost: 'even' (copy the edge value), 'odd_high' (even below, -edge above),
    'wrap'."""
    P = np.zeros((2 * nc, nc))
    for i in range(2 * nc):
        s = i / 2 - .25
        a = int(np.floor(s))
        t = s - a
        for node, wgt in ((a, 1 - t), (a + 1, t)):
            sign = 1.
            if ghost == 'wrap':
                node %= nc
            elif node < 0:
                node = 0
            elif node >= nc:
                node = nc - 1
                if ghost == 'odd_high':
                    sign = -1.
            P[i, node] += sign * wgt
    return P


def kron(mats):
    out = mats[0]
    for m in mats[1:]:
        out = np.kron(out, m)
    return out


def parity(shape):
    return np.indices(shape).sum(0).reshape(-1) % 2


class DenseCycle:
    """One V(nu,nu) cycle as (M, N) with cycle(p, rhs) = M p + N rhs."""

    def __init__(self, shapes, spacings, mode, smoother, nu, mean_free):
        self.shapes, self.spacings, self.mode = shapes, spacings, mode
        self.smoother, self.nu, self.mean_free = smoother, nu, mean_free
        self.A = [sparse.csr_matrix(assemble_operator(s, h, mode)) for s, h in zip(shapes, spacings)]

    def Q(self, n):
        return np.eye(n) - np.ones((n, n)) / n if self.mean_free else np.eye(n)

    def smooth(self, level):
        A = self.A[level]
        n = A.shape[0]
        dinv = 1. / A.diagonal()
        S, T = np.eye(n), np.zeros((n, n))
        colors = ((2 / 3) * np.ones(n),) if self.smoother == 'jacobi' else \
            tuple((parity(self.shapes[level]) == c).astype(float) for c in (0, 1))
        for _ in range(self.nu):
            for mask in colors:
                G = (mask * dinv)[:, None]
                S, T = S - G * (A @ S), T + G * (np.eye(n) - A @ T)
        q = self.Q(n)
        return q @ S, q @ T

    def transfer(self, level):
        fine, coarse = self.shapes[level], self.shapes[level + 1]
        R = kron([restriction_1d(n) for n in fine])
        ghosts = []
        for axis, nc in enumerate(coarse):
            if self.mode == 'neumann':
                ghosts.append('even')
            elif self.mode == 'periodic':
                ghosts.append('wrap' if len(fine) == 3 or axis == 0 else 'even')
            else:
                ghosts.append('odd_high' if axis == 0 else 'even')
        P = kron([prolongation_1d(nc, g) for nc, g in zip(coarse, ghosts)])
        return R, P

    def maps(self, level=0):
        A = self.A[level]
        n = A.shape[0]
        if level == len(self.shapes) - 1:
            C = A.toarray() + np.ones((n, n)) / n if self.mean_free else A.toarray()
            return np.zeros((n, n)), self.Q(n) @ np.linalg.inv(C)
        S1, T1 = self.smooth(level)
        R, P = self.transfer(level)
        _, Nc = self.maps(level + 1)
        K = P @ Nc @ self.Q(R.shape[0]) @ R
        M2 = S1 - K @ (A @ S1)
        N2 = T1 + K @ (np.eye(n) - A @ T1)
        S2, T2 = self.smooth(level)
        return S2 @ M2, S2 @ N2 + T2


CASES_3D = [((8, 8, 8), (1 / 8, 1 / 8, 1 / 8)),           # 2 levels
            ((8, 12, 16), (.125, 1 / 12, .0625)),         # 2 levels, anisotropic
            ((16, 16, 16), (1 / 16, 1 / 20, 1 / 24))]     # 3 levels


def relerr(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def oracle_3d(shape, spacing, smoother, periodic, nu, cls=mac3d.Multigrid):
    return cls(shape, spacing, MGOptions(smooth=nu, smoother=smoother), periodic=periodic)


def check_3d(mg, shape, spacing, smoother, periodic, nu, seed=0):
    dense = DenseCycle(mg.shapes, mg.spacings, 'periodic' if periodic else 'neumann', smoother, nu, True)
    M, N = dense.maps()
    gen = np.random.default_rng(seed)
    errs = []
    for _ in range(2):
        p, rhs = gen.normal(size=shape), gen.normal(size=shape)
        ref = (M @ p.reshape(-1) + N @ rhs.reshape(-1)).reshape(shape)
        errs.append(relerr(mg.cycle(0, p, rhs), ref))
    return max(errs)


@pytest.mark.parametrize('case', range(len(CASES_3D)))
@pytest.mark.parametrize('smoother', ['jacobi', 'rbgs'])
@pytest.mark.parametrize('periodic', [False, True])
def test_3d_cycle_equals_dense_composition(case, smoother, periodic):
    shape, spacing = CASES_3D[case]
    if case == 2 and (smoother, periodic) not in (('jacobi', False), ('rbgs', True)):
        pytest.skip('3-level case: one Jacobi/closed and one RBGS/periodic composition')
    nu = 4 if smoother == 'jacobi' else 2
    mg = oracle_3d(shape, spacing, smoother, periodic, nu)
    assert len(mg.shapes) == (3 if case == 2 else 2)
    assert check_3d(mg, shape, spacing, smoother, periodic, nu) <= TOL


# --- mistakes planted in the oracle must fail the same comparison ---------

class NoPostSmooth(mac3d.Multigrid):
    def cycle(self, level, p, rhs):
        if level == len(self.shapes) - 1:
            return super().cycle(level, p, rhs)
        p = self.smooth(level, p, rhs)
        coarse = mac3d.restrict(rhs - self.apply(level, p))
        coarse = coarse - coarse.mean()
        return p + mac3d.prolong(self.cycle(level + 1, np.zeros_like(coarse), coarse), p.shape, self.periodic)


class NoCoarseMean(mac3d.Multigrid):
    def cycle(self, level, p, rhs):
        if level == len(self.shapes) - 1:
            return super().cycle(level, p, rhs)
        p = self.smooth(level, p, rhs)
        coarse = mac3d.restrict(rhs - self.apply(level, p))
        p = p + mac3d.prolong(self.cycle(level + 1, np.zeros_like(coarse), coarse), p.shape, self.periodic)
        return self.smooth(level, p, rhs)


class DoubledCorrection(mac3d.Multigrid):
    def cycle(self, level, p, rhs):
        if level == len(self.shapes) - 1:
            return super().cycle(level, p, rhs)
        p = self.smooth(level, p, rhs)
        coarse = mac3d.restrict(rhs - self.apply(level, p))
        coarse = coarse - coarse.mean()
        p = p + 2 * mac3d.prolong(self.cycle(level + 1, np.zeros_like(coarse), coarse), p.shape, self.periodic)
        return self.smooth(level, p, rhs)


class BlackFirst(mac3d.Multigrid):
    def rbgs(self, level, p, rhs):
        par = np.indices(p.shape).sum(0) % 2
        p = p.copy()
        for color in (1, 0):
            r = rhs - self.apply(level, p)
            m = par == color
            p[m] = p[m] + r[m] / self.diagonals[level][m]
        return p


class MirroredProlongation(mac3d.Multigrid):
    """Coarse correction prolonged with its x index reversed."""
    def cycle(self, level, p, rhs):
        if level == len(self.shapes) - 1:
            return super().cycle(level, p, rhs)
        p = self.smooth(level, p, rhs)
        coarse = mac3d.restrict(rhs - self.apply(level, p))
        coarse = coarse - coarse.mean()
        corr = self.cycle(level + 1, np.zeros_like(coarse), coarse)
        p = p + mac3d.prolong(corr[::-1], p.shape, self.periodic)
        return self.smooth(level, p, rhs)


@pytest.mark.parametrize('planted', [NoPostSmooth, NoCoarseMean, DoubledCorrection, BlackFirst,
                                     MirroredProlongation])
def test_planted_cycle_mistakes_are_caught(planted):
    # Without the coarse mean removal a two-level cycle is unchanged (the
    # coarse inverse maps a constant to a constant that the final mean
    # removal takes out), so that mistake is planted in a three-level cycle,
    # where the middle level's Jacobi sweeps see the constant through the
    # non-constant diagonal of the closed box.
    shape, spacing = CASES_3D[2 if planted is NoCoarseMean else 1]
    smoother = 'rbgs' if planted is BlackFirst else 'jacobi'
    periodic = planted not in (MirroredProlongation, NoCoarseMean)
    nu = 2
    good = oracle_3d(shape, spacing, smoother, periodic, nu)
    assert check_3d(good, shape, spacing, smoother, periodic, nu) <= TOL
    bad = oracle_3d(shape, spacing, smoother, periodic, nu, cls=planted)
    assert check_3d(bad, shape, spacing, smoother, periodic, nu) > 1e-6


# --- 2D channel (reference BCs) and the build-own periodic-x mode ---------

CASES_2D = [((16, 8), (.5, .125)),      # 2 levels
            ((32, 16), (.25, 1 / 16))]  # 3 levels


@pytest.mark.parametrize('case', range(len(CASES_2D)))
@pytest.mark.parametrize('smoother', ['jacobi', 'rbgs'])
@pytest.mark.parametrize('periodic', [False, True])
def test_2d_cycle_equals_dense_composition(case, smoother, periodic):
    shape, spacing = CASES_2D[case]
    nu = 4 if smoother == 'jacobi' else 2
    mg = mac2d.ChannelMultigrid(shape, spacing, MGOptions(smooth=nu, smoother=smoother), periodic=periodic)
    assert len(mg.shapes) == case + 2
    dense = DenseCycle(mg.shapes, mg.spacings, 'periodic' if periodic else 'channel', smoother, nu, periodic)
    M, N = dense.maps()
    gen = np.random.default_rng(3)
    for _ in range(2):
        p, rhs = gen.normal(size=shape), gen.normal(size=shape)
        ref = (M @ p.reshape(-1) + N @ rhs.reshape(-1)).reshape(shape)
        assert relerr(mg.cycle(0, p, rhs), ref) <= TOL


class ChannelNoPostSmooth(mac2d.ChannelMultigrid):
    def cycle(self, level, p, rhs):
        if level == len(self.shapes) - 1:
            return super().cycle(level, p, rhs)
        p = self.smooth(level, p, rhs)
        coarse = mac2d.restrict(rhs - self.apply(level, p))
        return p + mac2d.prolong(self.cycle(level + 1, np.zeros_like(coarse), coarse), p.shape, self.periodic)


class ChannelEvenOutlet(mac2d.ChannelMultigrid):
    """Clamped (even) ghost at the outlet instead of the odd one."""
    def cycle(self, level, p, rhs):
        if level == len(self.shapes) - 1:
            return super().cycle(level, p, rhs)
        p = self.smooth(level, p, rhs)
        coarse = mac2d.restrict(rhs - self.apply(level, p))
        corr = self.cycle(level + 1, np.zeros_like(coarse), coarse)
        fine = np.zeros(p.shape)
        for i, j in product(range(p.shape[0]), range(p.shape[1])):
            sx, sy = i / 2 - .25, j / 2 - .25
            ax, ay = int(np.floor(sx)), int(np.floor(sy))
            tx, ty = sx - ax, sy - ay
            for (a, wa), (b, wb) in product(((ax, 1 - tx), (ax + 1, tx)), ((ay, 1 - ty), (ay + 1, ty))):
                fine[i, j] += wa * wb * corr[min(max(a, 0), corr.shape[0] - 1), min(max(b, 0), corr.shape[1] - 1)]
        return self.smooth(level, p + fine, rhs)


@pytest.mark.parametrize('planted', [ChannelNoPostSmooth, ChannelEvenOutlet])
def test_planted_channel_mistakes_are_caught(planted):
    shape, spacing = CASES_2D[1]
    opts = MGOptions(smooth=2, smoother='jacobi')
    mg = planted(shape, spacing, opts)
    dense = DenseCycle(mg.shapes, mg.spacings, 'channel', 'jacobi', 2, False)
    M, N = dense.maps()
    gen = np.random.default_rng(4)
    p, rhs = gen.normal(size=shape), gen.normal(size=shape)
    ref = (M @ p.reshape(-1) + N @ rhs.reshape(-1)).reshape(shape)
    good = mac2d.ChannelMultigrid(shape, spacing, opts).cycle(0, p, rhs)
    assert relerr(good, ref) <= TOL
    assert relerr(mg.cycle(0, p, rhs), ref) > 1e-6
