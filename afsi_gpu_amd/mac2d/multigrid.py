"""Channel pressure multigrid on the GPU (mac2d/multigrid.py:16-134).

Same hierarchy rule, exact coarse inverse (no nullspace for the reference
channel), V(nu,nu) cycle with ghost prolongation and no mean removal, same
solve contract: tolerance max(atol, rtol*||b||), convergence checked every
``check_every`` cycles, messages ('invalid channel pressure RHS', 'finite
pressure RHS/initial guess required', 'channel MG did not converge') and a
cloned result. Every cycle is issued by the native driver
``afsi_mg2d_cycles_*``; ``backend='graph'`` captures each block of
1..check_every cycles plus the residual reduction as a HIP graph.
Build-own: ``options.smoother='rbgs'`` and periodic-x grids (means removed
as in the 3D closed box; the rhs must be compatible).
"""
from math import isfinite, prod
import numpy as np
import torch

from .. import _abi
from ..mac.multigrid import MGOptions

OMEGA = 2 / 3  # mac2d/multigrid.py:12


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


def coarse_operator(shape, spacing, periodic):
    """Dense A of the coarsest level (setup only, host FP64)."""
    nx, ny = shape
    hx, hy = spacing
    ids = np.arange(nx * ny).reshape(shape)
    A = np.zeros((nx * ny, nx * ny))
    for c, w in ((0, 1 / hx ** 2), (1, 1 / hy ** 2)):
        if c == 0 and periodic:
            a, b = ids.reshape(-1), np.roll(ids, -1, axis=0).reshape(-1)
        elif c == 0:
            a, b = ids[:-1].reshape(-1), ids[1:].reshape(-1)
        else:
            a, b = ids[:, :-1].reshape(-1), ids[:, 1:].reshape(-1)
        np.add.at(A, (a, a), w)
        np.add.at(A, (b, b), w)
        np.add.at(A, (a, b), -w)
        np.add.at(A, (b, a), -w)
    if not periodic:
        last = ids[-1].reshape(-1)
        A[last, last] += 2 / hx ** 2
    return A


class ChannelMultigrid:
    def __init__(self, grid, *, device='cuda', dtype=torch.float64, options=None, fused=True,
                 backend='graph'):
        if backend not in ('reference', 'workspace', 'graph'):
            raise ValueError('pressure backend must be reference, workspace or graph')
        if dtype not in (torch.float64, torch.float32):
            raise ValueError('multigrid supports float64/float32')
        self.device = torch.device(device)
        if self.device.type != 'cuda':
            raise ValueError('the HIP channel multigrid requires a GPU device')
        self.grid, self.dtype = grid, dtype
        self.options = options or MGOptions()
        self.periodic = bool(getattr(grid, 'periodic', False))
        self.shapes, self.spacings = hierarchy(grid.shape, grid.spacing)
        self.suffix = 'f64' if dtype == torch.float64 else 'f32'
        kw = dict(device=self.device, dtype=dtype)
        self.p = [torch.zeros(s, **kw) for s in self.shapes]
        self.rhs = [torch.zeros(s, **kw) for s in self.shapes]
        self.tmp = [torch.zeros(s, **kw) for s in self.shapes]
        A = coarse_operator(self.shapes[-1], self.spacings[-1], self.periodic)
        if self.periodic:
            A = A + np.ones_like(A) / A.shape[0]
        self.inverse = torch.tensor(np.linalg.inv(A), **kw)
        lib = _abi.load()
        self.partials = torch.zeros(4 * int(lib.afsi_reduce_partials(prod(self.shapes[0]))),
                                    device=self.device, dtype=torch.float64)
        self.results = torch.zeros((4, 4), device=self.device, dtype=torch.float64)
        self.plan = self._make_plan()
        self.backend = 'hip-channel-graph' if backend == 'graph' else 'hip-channel'
        self.graphs = {}
        if backend == 'graph':
            if self.options.check_every > 16:
                raise ValueError('pressure graph check_every must be <=16')
            self._capture()

    def _make_plan(self):
        pl = _abi.MGPlan2()
        pl.levels = len(self.shapes)
        pl.periodic = int(self.periodic)
        pl.smoother = 1 if self.options.smoother == 'rbgs' else 0
        pl.nu = self.options.smooth
        pl.omega = OMEGA
        for l, (s, h) in enumerate(zip(self.shapes, self.spacings)):
            for a in range(2):
                pl.n[l][a] = s[a]
                pl.h2[l][a] = h[a] ** 2
                pl.inv_h2[l][a] = 1 / h[a] ** 2
            pl.p[l] = self.p[l].data_ptr()
            pl.rhs[l] = self.rhs[l].data_ptr()
            pl.tmp[l] = self.tmp[l].data_ptr()
        pl.coarse_inverse = self.inverse.data_ptr()
        pl.coarse_cells = prod(self.shapes[-1])
        pl.reduce_partials = self.partials.data_ptr()
        pl.reduce_result = self.results[0].data_ptr()
        pl.partials_capacity = self.partials.numel()
        return pl

    def _reduce(self, x, slot):
        _abi.call(f'afsi_reduce_{self.suffix}', _abi.ptr(x), x.numel(), _abi.ptr(self.partials),
                  _abi.ptr(self.results[slot]), _abi.stream())

    def _sub_mean(self, x, slot):
        _abi.call(f'afsi_sub_mean_{self.suffix}', _abi.ptr(x), x.numel(),
                  _abi.ptr(self.results[slot]), _abi.stream())

    def _block(self, count):
        _abi.call(f'afsi_mg2d_cycles_{self.suffix}', self.plan, int(count), _abi.stream())
        _abi.call(f'afsi_mg2d_residual_{self.suffix}', self.plan, _abi.stream())

    def _capture(self):
        self.p[0].zero_()
        self.rhs[0].zero_()
        stream = torch.cuda.Stream(device=self.device)
        stream.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(stream):
            for _ in range(2):
                self._block(self.options.check_every)
        torch.cuda.current_stream().wait_stream(stream)
        for count in range(1, self.options.check_every + 1):
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph, stream=stream):
                self._block(count)
            self.graphs[count] = graph
        torch.cuda.current_stream().wait_stream(stream)
        torch.cuda.synchronize(self.device)

    def advance(self, count):
        """``count`` V-cycles, then the level-0 residual reduction."""
        if self.graphs:
            self.graphs[count].replay()
        else:
            self._block(count)

    def cycle(self, p, rhs):
        """One V-cycle from (p, rhs); returns the workspace iterate."""
        self.p[0].copy_(p)
        self.rhs[0].copy_(rhs)
        _abi.call(f'afsi_mg2d_cycles_{self.suffix}', self.plan, 1, _abi.stream())
        return self.p[0]

    @torch.no_grad()
    def solve(self, rhs, initial=None, *, history=None):
        """mac2d/multigrid.py:83-134 (workspace branch :111-134)."""
        if rhs.shape != self.shapes[0] or rhs.device != self.device or rhs.dtype != self.dtype:
            raise ValueError('invalid channel pressure RHS')
        if initial is not None and (initial.shape != rhs.shape or initial.dtype != rhs.dtype
                                    or initial.device != rhs.device):
            raise ValueError('invalid channel pressure initial guess')
        w_p, w_rhs = self.p[0], self.rhs[0]
        w_rhs.copy_(rhs)
        if initial is None:
            w_p.zero_()
        else:
            w_p.copy_(initial)
        self._reduce(w_rhs, 1)
        self._reduce(w_p, 2)
        if self.periodic:
            self._sub_mean(w_rhs, 1)
            self._sub_mean(w_p, 2)
        self._reduce(w_rhs, 3)
        _abi.call(f'afsi_mg2d_residual_{self.suffix}', self.plan, _abi.stream())
        (s1, _, a1, nf1), (_, _, _, nf2), (_, q3, _, _), (_, q0, _, nf0) = \
            self.results[[1, 2, 3, 0]].tolist()
        if nf1 or nf2:
            raise ValueError('finite pressure RHS/initial guess required')
        count = w_rhs.numel()
        if self.periodic and abs(s1 / count) > 1e-12 + 1e-10 * a1 / count:
            raise ValueError('periodic channel pressure RHS is incompatible: net flux must vanish')
        rhs_norm, residual = q3 ** .5, q0 ** .5
        if not isfinite(rhs_norm) or not isfinite(residual) or nf0:
            raise RuntimeError('nonfinite channel pressure residual or RHS norm')
        tol = max(self.options.atol, self.options.rtol * rhs_norm)
        if history is not None:
            history.append(residual)
        cycle = 0
        while residual > tol and cycle < self.options.max_cycles:
            count = min(self.options.check_every, self.options.max_cycles - cycle)
            self.advance(count)
            cycle += count
            _, q0, _, nf0 = self.results[0].tolist()
            residual = q0 ** .5
            if history is not None:
                history.append(residual)
            if nf0 or not isfinite(residual):
                raise RuntimeError('nonfinite channel pressure residual')
        if residual > tol:
            raise RuntimeError(f'channel MG did not converge: {residual:g} > {tol:g}')
        if self.periodic:
            self._reduce(w_p, 2)
            self._sub_mean(w_p, 2)
        return w_p.clone(), dict(cycles=cycle, residual_norm=residual, tolerance=tol,
                                 backend=self.backend)
