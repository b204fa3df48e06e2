"""Geometric multigrid for the MAC pressure equation on the GPU.

Mirrors ``MGOptions`` and ``GeometricMultigrid`` of mac/multigrid.py:9-172:
same hierarchy rule (halve until min<=4 or any odd, coarsest <=512 cells),
same coarse inverse inv(A + 11^T/n), same V-cycle (smooth, mean removal,
residual restriction, coarse mean removal, recursion, trilinear
prolongation, smooth), same compatibility/finite checks, tolerance
max(atol, rtol*||b||), checks every ``check_every`` cycles, messages
('invalid pressure RHS', 'incompatible', 'did not converge', 'nonfinite')
and ownership (the result is a clone, never workspace storage).

All cycle launches are issued by the native driver ``afsi_mg3d_cycles_*``;
with ``backend='graph'`` each block of 1..check_every cycles plus the
residual reduction is captured once as a HIP graph and replayed. Build-own
options: ``smoother='rbgs'`` (red-black Gauss-Seidel) and periodic grids.
"""
from dataclasses import dataclass
from math import isfinite, prod
import numpy as np
import torch

from .. import _abi
from .cuda_ib import build as _build

OMEGA = 2 / 3  # mac/multigrid.py:109
# Red-black over-relaxation weight of the mixed path's FP32 V-cycles (build-
# own): the V(2,2) cycle's residual reduction falls from about 0.09 to 0.018
# per cycle, 6 cycles instead of 10 to rtol 1e-10 at 128^3.
MIXED_RELAX = 1.3
# FP32 V-cycles per FP64 refinement step of the mixed path: with the relaxed
# cycle's 0.018 reduction, three cycles cut the error by ~6e-6, well above
# FP32 rounding, so a 6-cycle solve needs 2 FP64 residual passes instead of 6
# (128^3 with 1.5M points: 603 -> 634 steps/s).
MIXED_INNER_CYCLES = 3


@dataclass(frozen=True)
class MGOptions:
    rtol: float = 1e-10
    atol: float = 1e-12
    max_cycles: int = 100
    smooth: int = 4
    check_every: int = 2
    smoother: str = 'jacobi'
    relax: float = 1.0  # build-own: red-black over-relaxation weight (1: Gauss-Seidel)
    inner_cycles: int = 1  # build-own: FP32 V-cycles per FP64 refinement step (mixed precision)

    def __post_init__(self):
        if (any(not isfinite(x) or x < 0 for x in (self.rtol, self.atol)) or
                self.rtol + self.atol == 0 or any(type(n) is not int or n < 1 for n in
                (self.max_cycles, self.smooth, self.check_every))):
            raise ValueError('invalid multigrid options')
        if self.smoother not in ('jacobi', 'rbgs'):
            raise ValueError('smoother must be jacobi or rbgs')
        if not (isfinite(self.relax) and 0 < self.relax < 2):
            raise ValueError('relax must lie in (0, 2)')
        if type(self.inner_cycles) is not int or self.inner_cycles < 1:
            raise ValueError('inner_cycles must be a positive integer')


def hierarchy(shape, spacing):
    """mac/multigrid.py:33-48."""
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


def coarse_operator(shape, spacing, periodic):
    """Dense A of the coarsest level (setup only; host FP64)."""
    count = prod(shape)
    ids = np.arange(count).reshape(shape)
    A = np.zeros((count, count))
    for c, h in enumerate(spacing):
        w = 1 / h ** 2
        if periodic:
            a, b = ids.reshape(-1), np.roll(ids, -1, axis=c).reshape(-1)
        else:
            sl_a = [slice(None)] * 3
            sl_b = [slice(None)] * 3
            sl_a[c], sl_b[c] = slice(None, -1), slice(1, None)
            a, b = ids[tuple(sl_a)].reshape(-1), ids[tuple(sl_b)].reshape(-1)
        np.add.at(A, (a, a), w)
        np.add.at(A, (b, b), w)
        np.add.at(A, (a, b), -w)
        np.add.at(A, (b, a), -w)
    return A


class GeometricMultigrid:
    def __init__(self, grid, *, device='cuda', dtype=torch.float64, options=None, backend='graph'):
        if backend not in ('torch', 'fused', 'workspace', 'graph', 'unfused'):
            raise ValueError('pressure backend must be torch, fused, workspace, graph or unfused')
        if dtype not in (torch.float64, torch.float32):
            raise ValueError('multigrid supports float64/float32')
        _build()
        self.grid = grid
        self.options = MGOptions() if options is None else options
        self.dtype, self.device = dtype, torch.device(device)
        if self.device.type != 'cuda':
            raise ValueError('the HIP multigrid requires a GPU device')
        self.periodic = bool(getattr(grid, 'periodic', False))
        self.shapes, self.spacings = hierarchy(grid.shape, grid.spacing)
        self.use_graphs = backend == 'graph'
        self.fused = backend != 'unfused'
        self.backend = ('hip-graph' if self.use_graphs else 'hip') if self.fused else 'hip-unfused'
        self.suffix = 'f64' if dtype == torch.float64 else 'f32'
        kw = dict(device=self.device, dtype=dtype)
        self.p = [torch.zeros(s, **kw) for s in self.shapes]
        self.rhs = [torch.zeros(s, **kw) for s in self.shapes]
        self.tmp = [torch.zeros(s, **kw) for s in self.shapes]
        A = coarse_operator(self.shapes[-1], self.spacings[-1], self.periodic)
        count = prod(self.shapes[-1])
        # mac/multigrid.py:49-59: the rank-one term fixes the constant nullspace.
        self.coarse_inverse = torch.tensor(np.linalg.inv(A + np.ones_like(A) / count), **kw)
        cells0 = prod(self.shapes[0])
        lib = _abi.load()
        capacity = max(4 * int(lib.afsi_reduce_partials(cells0)),
                       int(lib.afsi_mg3d_partials_doubles(_abi.i64x3(self.shapes[0]))))
        self.partials = torch.zeros(capacity, device=self.device, dtype=torch.float64)
        self.results = torch.zeros((8, 4), device=self.device, dtype=torch.float64)
        self.plan = self._make_plan()
        if self.fused and not lib.afsi_mg3d_supported(self.plan):
            self.fused = False
            self.backend = 'hip-unfused'
        self.allocated_bytes = sum(t.numel() * t.element_size()
                                   for group in (self.p, self.rhs, self.tmp) for t in group)
        self.graphs = {}
        if self.use_graphs:
            if self.options.check_every > 16:
                raise ValueError('pressure graph check_every must be <=16')
            self._capture()

    def _make_plan(self):
        pl = _abi.MGPlan3()
        pl.levels = len(self.shapes)
        pl.periodic = int(self.periodic)
        pl.smoother = 1 if self.options.smoother == 'rbgs' else 0
        pl.nu = self.options.smooth
        pl.omega = OMEGA
        pl.relax = self.options.relax
        for l, (s, h) in enumerate(zip(self.shapes, self.spacings)):
            for a in range(3):
                pl.n[l][a] = s[a]
                pl.h2[l][a] = h[a] * h[a]
                pl.inv_h2[l][a] = 1 / h[a] ** 2
            pl.p[l] = self.p[l].data_ptr()
            pl.rhs[l] = self.rhs[l].data_ptr()
            pl.tmp[l] = self.tmp[l].data_ptr()
        pl.coarse_inverse = self.coarse_inverse.data_ptr()
        pl.coarse_cells = prod(self.shapes[-1])
        pl.reduce_partials = self.partials.data_ptr()
        pl.reduce_result = self.results[0].data_ptr()
        pl.partials_capacity = self.partials.numel()
        return pl

    # -- native calls on the current stream ---------------------------------
    def _cycles(self, count):
        kind = '' if self.fused else 'unfused_'
        _abi.call(f'afsi_mg3d_cycles_{kind}{self.suffix}', self.plan, int(count), _abi.stream())

    def _residual(self):
        kind = '' if self.fused else 'unfused_'
        _abi.call(f'afsi_mg3d_residual_{kind}{self.suffix}', self.plan, _abi.stream())

    def _materialize(self):
        """Remove the constant offset the fused driver leaves in p[0]."""
        if self.fused:
            self._reduce(self.p[0], 4)
            self._sub_mean(self.p[0], 4)

    def _reduce(self, x, slot):
        _abi.call(f'afsi_reduce_{self.suffix}', _abi.ptr(x), x.numel(), _abi.ptr(self.partials),
                  _abi.ptr(self.results[slot]), _abi.stream())

    def _sub_mean(self, x, slot):
        _abi.call(f'afsi_sub_mean_{self.suffix}', _abi.ptr(x), x.numel(),
                  _abi.ptr(self.results[slot]), _abi.stream())

    def _block(self, count):
        self._cycles(count)
        self._residual()

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
        """Run ``count`` V-cycles and the level-0 residual reduction."""
        if self.graphs:
            self.graphs[count].replay()
        else:
            self._block(count)

    # -- reference-compatible API ---------------------------------------------
    def cycle(self):
        """One V-cycle on the workspace iterate (MGWorkspace.cycle)."""
        self._cycles(1)
        self._materialize()
        return self.p[0]

    def initialize(self, p, rhs):
        self.p[0].copy_(p)
        self.rhs[0].copy_(rhs)

    def residual_norm(self):
        self._residual()
        return self.results[0, 1].sqrt()

    @torch.no_grad()
    def solve(self, rhs, initial=None, *, rtol=None, history=None, precheck=None, view=False):
        """mac/multigrid.py:125-172 (workspace branch :75-98). ``history``, if
        a list, receives the level-0 residual norm after every check;
        ``precheck``, if given, is called right after the first host
        synchronisation (before any cycle and before the RHS checks), so a
        caller can read device flags it copied to pinned memory earlier on
        the stream without a synchronisation of its own. ``rhs`` or
        ``initial`` that already are the solver's own level-0 buffers are not
        copied; ``view=True`` returns the solver's level-0 iterate itself
        (valid until the next solve) instead of a copy."""
        rtol = self.options.rtol if rtol is None else rtol
        if not isfinite(rtol) or rtol < 0:
            raise ValueError('invalid multigrid solve tolerance')
        if rhs.shape != self.shapes[0] or rhs.device != self.device or rhs.dtype != self.dtype:
            raise ValueError('invalid pressure RHS')
        if initial is not None and (initial.shape != rhs.shape or initial.device != rhs.device
                                    or initial.dtype != rhs.dtype):
            raise ValueError('invalid initial pressure')
        w_p, w_rhs = self.p[0], self.rhs[0]
        if rhs.data_ptr() != w_rhs.data_ptr():
            w_rhs.copy_(rhs)
        if initial is None:
            w_p.zero_()
        elif initial.data_ptr() != w_p.data_ptr():
            w_p.copy_(initial)
        self.launch_start(zero=initial is None)
        cycle, residual, tolerance = self.settle(rtol, history, precheck, zero=initial is None)
        self.launch_finish()
        return (w_p if view else w_p.clone()), dict(cycles=cycle, residual_norm=residual, tolerance=tolerance,
                                                    backend=self.backend)

    # -- the parts of ``solve`` (IBStep captures the launch-only ones) ------
    def launch_start(self, zero=False):
        """_start_metrics (mac/multigrid.py:68-73): mean removal of the RHS and
        of the iterate in the level-0 buffers, the RHS norm and the initial
        residual (device work only). ``zero``: the iterate is zero, so its
        mean and finiteness need no pass."""
        w_p, w_rhs = self.p[0], self.rhs[0]
        self._reduce(w_rhs, 1)
        if not zero:
            self._reduce(w_p, 2)
        self._sub_mean(w_rhs, 1)
        if not zero:
            self._sub_mean(w_p, 2)
        self._reduce(w_rhs, 3)
        self._residual()

    def settle(self, rtol, history=None, precheck=None, zero=False):
        """One host synchronisation for the start metrics, their checks, then
        V-cycles until the tolerance; returns (cycles, residual, tolerance).
        ``zero`` as for ``launch_start``."""
        count = self.rhs[0].numel()
        (s1, _, a1, nf1), (_, _, _, nf2), (_, q3, _, _), (_, q0, _, nf0) = \
            self.results[[1, 2, 3, 0]].tolist()
        if zero:
            nf2 = 0
        if precheck is not None:
            precheck()
        if nf1:
            raise ValueError('invalid pressure RHS')
        mean, absolute_mean = s1 / count, a1 / count
        if abs(mean) > 1e-12 + 1e-10 * absolute_mean:
            raise ValueError('incompatible Neumann pressure RHS: net flux must vanish')
        if nf2:
            raise ValueError('invalid initial pressure')
        rhs_norm, residual = q3 ** .5, q0 ** .5
        if not isfinite(rhs_norm) or not isfinite(residual) or nf0:
            raise RuntimeError('nonfinite multigrid residual or RHS norm')
        tolerance = max(self.options.atol, rtol * rhs_norm)
        cycle = 0
        if history is not None:
            history.append(residual)
        while residual > tolerance and cycle < self.options.max_cycles:
            count = min(self.options.check_every, self.options.max_cycles - cycle)
            self.advance(count)
            cycle += count
            _, q0, _, nf0 = self.results[0].tolist()
            residual = q0 ** .5
            if history is not None:
                history.append(residual)
            if nf0 or not isfinite(residual):
                raise RuntimeError('nonfinite multigrid residual')
        if residual > tolerance:
            raise RuntimeError(f'pressure multigrid did not converge: {residual:g} > {tolerance:g}')
        return cycle, residual, tolerance

    def launch_finish(self):
        """Remove the constant offset the fused driver leaves in p[0]."""
        self._materialize()


class MixedMultigrid:
    """Mixed-precision pressure solve (build-own): FP32 V-cycles inside FP64
    iterative refinement.

    Each outer iteration runs ``options.inner_cycles`` fused FP32 V-cycles on
    the FP32 copy of the FP64 residual (zero initial guess; the error of the
    correction falls by the V-cycle factor per cycle, far above FP32
    rounding for one or two cycles), then one fused pass adds the FP32
    correction to the FP64 iterate (ping-pong buffers) and recomputes
    r = b - A x in FP64; the residual norm is reduced only where the solve
    checks convergence. The residual, the accumulation of x and the removal
    of the constant null space (mean of b and of the result) stay in FP64,
    so the iteration converges to the FP64 tolerance max(atol, rtol*||b||)
    of ``GeometricMultigrid.solve`` (the FP32 cycle only has to reduce the
    error by the V-cycle factor). ``solve`` keeps the signature, checks and
    messages of the FP64 solver.
    """

    def __init__(self, grid, *, device='cuda', options=None, backend='graph'):
        self.options = options or MGOptions(smooth=2, smoother='rbgs')
        self.device = torch.device(device)
        self.grid = grid
        self.inner = GeometricMultigrid(grid, device=device, dtype=torch.float32,
                                        options=self.options, backend='workspace')
        if not self.inner.fused:
            raise ValueError('mixed precision needs the fused multigrid driver')
        self.inner.plan.zero_initial = 1
        self.shapes = self.inner.shapes
        self.periodic = self.inner.periodic
        self.dtype = torch.float64
        shape = self.shapes[0]
        self.xs = [torch.zeros(shape, device=self.device, dtype=torch.float64) for _ in range(2)]
        self.cur = 0
        self.b = torch.zeros(shape, device=self.device, dtype=torch.float64)
        self.partials = torch.zeros_like(self.inner.partials)
        self.results = torch.zeros((8, 4), device=self.device, dtype=torch.float64)
        h = self.inner.spacings[0]
        self._n = _abi.i64x3(shape)
        self._w = _abi.f64x3([1 / a ** 2 for a in h])
        self.backend = 'hip-mixed-graph' if backend == 'graph' else 'hip-mixed'
        self.graphs = {}
        self.last_cycles = None
        if backend == 'graph':
            self._capture()

    @property
    def x(self):
        """The current FP64 iterate."""
        return self.xs[self.cur]

    def _residual(self, slot=0):
        _abi.call('afsi_refine_residual_3d', _abi.ptr(self.x), _abi.ptr(self.b),
                  _abi.ptr(self.inner.rhs[0]), self._n, self._w, int(self.periodic),
                  _abi.ptr(self.partials), self.partials.numel(), _abi.ptr(self.results[slot]),
                  _abi.stream())

    def _iteration(self, start, reduce):
        """FP32 V-cycles, then x_{1-start} = x_start + e and the new residual."""
        self.inner._cycles(self.options.inner_cycles)
        _abi.call('afsi_refine_step_3d', _abi.ptr(self.xs[start]), _abi.ptr(self.inner.p[0]),
                  _abi.ptr(self.b), _abi.ptr(self.xs[1 - start]), _abi.ptr(self.inner.rhs[0]),
                  self._n, self._w, int(self.periodic), _abi.ptr(self.partials),
                  self.partials.numel(), _abi.ptr(self.results[0]) if reduce else None,
                  _abi.stream())

    def _block(self, start, count):
        for i in range(count):
            self._iteration((start + i) % 2, i == count - 1)

    def _capture(self):
        for x in self.xs:
            x.zero_()
        self.b.zero_()
        stream = torch.cuda.Stream(device=self.device)
        stream.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(stream):
            self._residual()
            for start in (0, 1):
                self._block(start, self.options.check_every)
        torch.cuda.current_stream().wait_stream(stream)
        for start in (0, 1):
            for count in range(1, self.options.check_every + 1):
                graph = torch.cuda.CUDAGraph()
                with torch.cuda.graph(graph, stream=stream):
                    self._block(start, count)
                self.graphs[start, count] = graph
        torch.cuda.current_stream().wait_stream(stream)
        torch.cuda.synchronize(self.device)

    def _graph(self, start, count):
        """The captured block of ``count`` iterations from buffer ``start``
        (blocks longer than ``check_every`` are captured on first use)."""
        graph = self.graphs.get((start, count))
        if graph is None:
            stream = torch.cuda.Stream(device=self.device)
            stream.wait_stream(torch.cuda.current_stream())
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph, stream=stream):
                self._block(start, count)
            torch.cuda.current_stream().wait_stream(stream)
            self.graphs[start, count] = graph
        return graph

    def advance(self, count):
        """``count`` outer iterations from the current iterate, then the
        residual norm reduction."""
        if self.graphs:
            self._graph(self.cur, count).replay()
        else:
            self._block(self.cur, count)
        self.cur = (self.cur + count) % 2

    def _reduce(self, t, slot):
        _abi.call('afsi_reduce_f64', _abi.ptr(t), t.numel(), _abi.ptr(self.partials),
                  _abi.ptr(self.results[slot]), _abi.stream())

    def _sub_mean(self, t, slot):
        _abi.call('afsi_sub_mean_f64', _abi.ptr(t), t.numel(), _abi.ptr(self.results[slot]),
                  _abi.stream())

    @torch.no_grad()
    def solve(self, rhs, initial=None, *, rtol=None, history=None, precheck=None, view=False):
        """Same contract as ``GeometricMultigrid.solve`` (``precheck`` and
        ``view`` included; ``b`` and either FP64 iterate count as the
        solver's own buffers); ``info['cycles']`` counts FP32 V-cycles
        (``options.inner_cycles`` per outer iteration; convergence is
        checked after whole outer iterations)."""
        rtol = self.options.rtol if rtol is None else rtol
        if not isfinite(rtol) or rtol < 0:
            raise ValueError('invalid multigrid solve tolerance')
        if rhs.shape != self.shapes[0] or rhs.device != self.device or rhs.dtype != torch.float64:
            raise ValueError('invalid pressure RHS')
        if initial is not None and (initial.shape != rhs.shape or initial.device != rhs.device
                                    or initial.dtype != rhs.dtype):
            raise ValueError('invalid initial pressure')
        if rhs.data_ptr() != self.b.data_ptr():
            self.b.copy_(rhs)
        own = [k for k in (0, 1) if initial is not None and initial.data_ptr() == self.xs[k].data_ptr()]
        self.cur = own[0] if own else 0
        if initial is None:
            self.x.zero_()
        elif not own:
            self.x.copy_(initial)
        lead = self.lead(history)
        self.launch_start(self.cur, lead, zero=initial is None)
        cycle, residual, tolerance = self.settle(self.cur, lead, rtol, history, precheck, zero=initial is None)
        self.launch_finish()
        inner = self.options.inner_cycles
        return (self.x if view else self.x.clone()), dict(cycles=cycle * inner, residual_norm=residual,
                                                          tolerance=tolerance, backend=self.backend)

    # -- the parts of ``solve`` (IBStep captures the launch-only ones) ------
    # The first block runs one outer iteration less than the previous solve
    # needed (consecutive time steps need about as many), then the residual is
    # checked after every iteration. The block's first iteration ("lead") is
    # launched before the readback of the start metrics (it writes only the
    # other FP64 buffer, so it is dropped if the initial iterate already meets
    # the tolerance or the checks raise): two host synchronisations per solve
    # in the steady state.
    def _outer_limit(self):
        return -(-self.options.max_cycles // self.options.inner_cycles)

    def lead(self, history=None):
        """1 if the next solve launches its first outer iteration with the
        start metrics, else 0."""
        predicted = self.last_cycles - 1 if self.last_cycles and history is None else 0
        return 1 if predicted > 0 and self._outer_limit() > 0 else 0

    def launch_start(self, start, lead, captured=False, zero=False):
        """Mean removal and metrics of b and of the iterate xs[start], its
        residual (results slot 5) and, with ``lead``, the first outer
        iteration (device work only; ``captured``: issue the iteration's
        launches directly, for a graph being captured; ``zero``: the iterate
        is zero, so its mean and finiteness need no pass)."""
        self.cur = start
        self._reduce(self.b, 1)
        if not zero:
            self._reduce(self.x, 2)
        self._sub_mean(self.b, 1)
        if not zero:
            self._sub_mean(self.x, 2)
        self._reduce(self.b, 3)
        self._residual(5)
        if lead:
            if captured:
                self._block(start, 1)
            else:
                self.advance(1)
        self.cur = start

    def settle(self, start, lead, rtol, history=None, precheck=None, zero=False):
        """Read the start metrics (one host synchronisation), check them, then
        outer iterations until the tolerance; leaves ``cur`` on the result and
        returns (outer iterations, residual, tolerance). ``zero`` as for
        ``launch_start``."""
        predicted = self.last_cycles - 1 if self.last_cycles and history is None else 0
        most = self._outer_limit()
        count = self.b.numel()
        res = self.results.tolist()
        if zero:
            res[2][3] = 0
        self.cur = start
        if precheck is not None:
            precheck()
        (s1, _, a1, nf1), (_, _, _, nf2), (_, q3, _, _), (_, q0, _, nf0) = res[1], res[2], res[3], res[5]
        if nf1:
            raise ValueError('invalid pressure RHS')
        mean, absolute_mean = s1 / count, a1 / count
        if abs(mean) > 1e-12 + 1e-10 * absolute_mean:
            raise ValueError('incompatible Neumann pressure RHS: net flux must vanish')
        if nf2:
            raise ValueError('invalid initial pressure')
        rhs_norm, residual = q3 ** .5, q0 ** .5
        if not isfinite(rhs_norm) or not isfinite(residual) or nf0:
            raise RuntimeError('nonfinite multigrid residual or RHS norm')
        tolerance = max(self.options.atol, rtol * rhs_norm)
        cycle = 0
        if history is not None:
            history.append(residual)
        if lead and residual > tolerance:
            self.cur = 1 - start
            cycle = 1
            _, q0, _, nf0 = res[0]
            residual = q0 ** .5
            if nf0 or not isfinite(residual):
                raise RuntimeError('nonfinite multigrid residual')
        pending = predicted - lead
        while residual > tolerance and cycle < most:
            if pending > 0:
                step, pending = pending, 0
            else:
                step = 1 if predicted > 0 else self.options.check_every
            n = min(step, most - cycle)
            self.advance(n)
            cycle += n
            _, q0, _, nf0 = self.results[0].tolist()
            residual = q0 ** .5
            if history is not None:
                history.append(residual)
            if nf0 or not isfinite(residual):
                raise RuntimeError('nonfinite multigrid residual')
        if residual > tolerance:
            raise RuntimeError(f'pressure multigrid did not converge: {residual:g} > {tolerance:g}')
        self.last_cycles = cycle
        return cycle, residual, tolerance

    def launch_finish(self):
        """Mean removal of the result xs[cur] (device work only)."""
        self._reduce(self.x, 4)
        self._sub_mean(self.x, 4)
