"""Many independent IB instances stepped together (build-own).

``BatchedMixedMultigrid`` solves the pressure equations of uneven 3D
instances at once through ``afsi_mgb_*``: every multigrid phase (the
red-black sweeps of a level, restriction, prolongation, the LDS-resident
coarse levels, the FP64 refinement step) is one launch over a device work
list of (instance, tile or row) entries, and convergence is tracked per
instance on the device (converged instances skip their work; the host
reads the flags once per ``check_every`` iterations). The solve keeps the
contract of ``MixedMultigrid.solve`` per instance: FP32 V-cycles inside
FP64 iterative refinement, tolerance max(atol, rtol ||b||), the same
compatibility/finite checks and messages, cloned results.

``BatchedTransfer`` bins, spreads and interpolates all instances with
one launch per phase (one key space, one work list of (instance, tile,
chunk)). ``BatchedIBStep`` advances a list of 3D instances by one IB step
each with both.
"""
from concurrent.futures import ThreadPoolExecutor
import ctypes
from ctypes import byref, c_double, c_int64, c_size_t
from math import isfinite
import torch

from . import _abi
from .mac.multigrid import GeometricMultigrid, MGOptions
from .mac.transfer import DEFAULT_TILE


class MixedState:
    """Resident per-instance storage of the batched mixed solve: the FP32
    multigrid hierarchy of its inner V-cycle and its FP64 iterates and
    right-hand side."""

    def __init__(self, grid, *, device='cuda', options=None):
        self.grid = grid
        self.options = options or MGOptions(smooth=2, smoother='rbgs')
        self.device = torch.device(device)
        self.inner = GeometricMultigrid(grid, device=self.device, dtype=torch.float32,
                                        options=self.options, backend='workspace')
        if not self.inner.fused:
            raise ValueError(f'grid {grid.shape} is not supported by the fused driver')
        shape = self.inner.shapes[0]
        self.x = [torch.zeros(shape, device=self.device, dtype=torch.float64) for _ in range(2)]
        self.b = torch.zeros(shape, device=self.device, dtype=torch.float64)

    @property
    def allocated_bytes(self):
        return self.inner.allocated_bytes + 3 * self.b.numel() * 8


class BatchedMixedMultigrid:
    def __init__(self, grids=None, *, device='cuda', options=None, graphs=True, states=None):
        self.options = options or MGOptions(smooth=2, smoother='rbgs')
        if self.options.smoother != 'rbgs':
            raise ValueError('the batched solve uses the red-black smoother')
        self.device = torch.device(device)
        if states is None:
            states = [MixedState(g, device=self.device, options=self.options) for g in (grids or [])]
        self.states = list(states)
        if not self.states:
            raise ValueError('empty batch')
        self.grids = [st.grid for st in self.states]
        self.inner = [st.inner for st in self.states]
        self.x = [st.x for st in self.states]
        self.b = [st.b for st in self.states]
        n = len(self.states)
        self.items = (_abi.MGBItem * n)()
        for i, st in enumerate(self.states):
            self.items[i].plan = st.inner.plan
            self.items[i].x[0] = st.x[0].data_ptr()
            self.items[i].x[1] = st.x[1].data_ptr()
            self.items[i].b = st.b.data_ptr()
        self.active = torch.zeros(n, device=self.device, dtype=torch.int32)
        self.tol = torch.zeros(n, device=self.device, dtype=torch.float64)
        self.results = torch.zeros((n, 16), device=self.device, dtype=torch.float64)
        bt = self._batch(None, 0, None)
        lib = _abi.load()
        nbytes = int(lib.afsi_mgb_workspace_bytes(bt))
        parts = int(lib.afsi_mgb_partials_doubles(bt))
        if nbytes <= 0 or parts <= 0:
            bt.workspace_bytes = 1
            _abi.call('afsi_mgb_prepare', bt, _abi.stream())  # raises with the reason
        self.workspace = torch.empty(nbytes, device=self.device, dtype=torch.uint8)
        self.partials = torch.zeros(parts, device=self.device, dtype=torch.float64)
        self.batch = self._batch(self.workspace, nbytes, self.partials)
        _abi.call('afsi_mgb_prepare', self.batch, _abi.stream())
        self.cur = 0
        self.graphs = {}
        self.backend = 'hip-batched-mixed-graph' if graphs else 'hip-batched-mixed'
        if graphs:
            self._capture()

    def _batch(self, workspace, nbytes, partials):
        bt = _abi.MGBBatch()
        bt.count = len(self.states)
        bt.inner_cycles = self.options.inner_cycles
        bt.items = _abi.cast_items(self.items)
        bt.workspace = workspace.data_ptr() if workspace is not None else None
        bt.workspace_bytes = nbytes
        bt.active = self.active.data_ptr()
        bt.tol = self.tol.data_ptr()
        bt.results = self.results.data_ptr()
        bt.partials = partials.data_ptr() if partials is not None else None
        return bt

    @property
    def allocated_bytes(self):
        return sum(st.allocated_bytes for st in self.states) + self.workspace.numel()

    def _iterate(self, start, count):
        _abi.call('afsi_mgb_iterate', self.batch, int(start), int(count), _abi.stream())

    def _capture(self):
        stream = torch.cuda.Stream(device=self.device)
        stream.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(stream):
            _abi.call('afsi_mgb_begin', self.batch, c_double(self.options.rtol),
                      c_double(self.options.atol), _abi.stream())
            for start in (0, 1):
                self._iterate(start, self.options.check_every)
        torch.cuda.current_stream().wait_stream(stream)
        for start in (0, 1):
            for count in range(1, self.options.check_every + 1):
                graph = torch.cuda.CUDAGraph()
                with torch.cuda.graph(graph, stream=stream):
                    self._iterate(start, count)
                self.graphs[start, count] = graph
        torch.cuda.current_stream().wait_stream(stream)
        torch.cuda.synchronize(self.device)

    def advance(self, count):
        if self.graphs:
            self.graphs[self.cur, count].replay()
        else:
            self._iterate(self.cur, count)
        self.cur = (self.cur + count) % 2

    @torch.no_grad()
    def solve(self, rhs, initial=None):
        """Per instance i: the FP64 solution of A p_i = rhs_i (compatible,
        mean-free) and an info dict like ``MixedMultigrid.solve``'s."""
        n = len(self.grids)
        if len(rhs) != n or (initial is not None and len(initial) != n):
            raise ValueError('one right-hand side (and initial guess) per instance')
        for i in range(n):
            r = rhs[i]
            if r.shape != self.b[i].shape or r.device != self.device or r.dtype != torch.float64:
                raise ValueError('invalid pressure RHS')
            self.b[i].copy_(r)
            if initial is None or initial[i] is None:
                self.x[i][0].zero_()
            else:
                if initial[i].shape != r.shape or initial[i].dtype != r.dtype:
                    raise ValueError('invalid initial pressure')
                self.x[i][0].copy_(initial[i])
        infos = self.solve_resident(0)
        return [self.x[i][self.cur].clone() for i in range(n)], infos

    @torch.no_grad()
    def solve_resident(self, start):
        """The same solve on the solver's own buffers, with no copies: the
        right-hand sides already in ``b[i]`` and the initial guesses in
        ``x[i][start]`` (``start = cur`` continues from the previous
        solution). The solutions are left in ``x[i][cur]``; returns the
        per-instance info dicts."""
        n = len(self.grids)
        self.cur = int(start)
        _abi.call('afsi_mgb_begin_at', self.batch, self.cur, c_double(self.options.rtol),
                  c_double(self.options.atol), _abi.stream())
        res = self.results.tolist()
        tol = self.tol.tolist()
        for i in range(n):
            count = self.b[i].numel()
            s1, a1, nf1 = res[i][4], res[i][6], res[i][7]
            if nf1:
                raise ValueError(f'invalid pressure RHS (instance {i})')
            if abs(s1 / count) > 1e-12 + 1e-10 * a1 / count:
                raise ValueError(f'incompatible Neumann pressure RHS: net flux must vanish (instance {i})')
            if res[i][15]:
                raise ValueError(f'invalid initial pressure (instance {i})')
            if not isfinite(res[i][9]) or not isfinite(res[i][1]) or res[i][3]:
                raise RuntimeError(f'nonfinite multigrid residual or RHS norm (instance {i})')
        residual = [r[1] ** .5 for r in res]
        cycles = [0] * n
        active = self.active.tolist()
        cycle = 0
        inner = self.options.inner_cycles   # FP32 V-cycles per refinement iteration
        most = -(-self.options.max_cycles // inner)
        while any(active) and cycle < most:
            step = min(self.options.check_every, most - cycle)
            self.advance(step)
            cycle += step
            now = self.active.tolist()
            res = self.results.tolist()
            for i in range(n):
                if active[i]:
                    cycles[i] = cycle * inner
                    residual[i] = res[i][1] ** .5
                    if res[i][3] or not isfinite(residual[i]):
                        raise RuntimeError(f'nonfinite multigrid residual (instance {i})')
            active = now
        for i in range(n):
            if residual[i] > tol[i]:
                raise RuntimeError(f'pressure multigrid did not converge (instance {i}): '
                                   f'{residual[i]:g} > {tol[i]:g}')
        _abi.call('afsi_mgb_finish', self.batch, int(self.cur), _abi.stream())
        return [dict(cycles=cycles[i], residual_norm=residual[i], tolerance=tol[i], backend=self.backend)
                for i in range(n)]


class BatchedTransfer:
    """Binning, spread and interpolation of many 3D instances with one
    launch per phase (``afsi_tb_*``). ``bind`` records the per-instance
    tensors (positions, forces, face arrays, point-velocity outputs) and
    uploads the descriptors; call it again when any tensor changes."""

    def __init__(self, grids, counts, *, device='cuda', tile=DEFAULT_TILE):
        self.grids = list(grids)
        self.counts = [int(c) for c in counts]
        self.device = torch.device(device)
        self.tile = tuple(int(t) for t in tile)
        self.total = sum(self.counts)
        self.items = None
        self.bound = None
        self.flags = torch.zeros((), device=self.device, dtype=torch.int32)

    def bind(self, xs, forces, faces, outs):
        n = len(self.grids)
        items = (_abi.TBItem * n)()
        for i in range(n):
            if xs[i].shape != (self.counts[i], 3):
                raise ValueError('positions must match the bound point counts')
            items[i].x = xs[i].data_ptr()
            items[i].force = forces[i].data_ptr()
            for c in range(3):
                items[i].face[c] = faces[i][c].data_ptr()
            items[i].out = outs[i].data_ptr()
            items[i].points = self.counts[i]
            items[i].grid = self.grids[i].abi()
        bt = _abi.TBBatch()
        bt.count = n
        for a in range(3):
            bt.tile[a] = self.tile[a]
        bt.items = ctypes_items(items)
        total, keys, nbytes = c_int64(), c_int64(), c_size_t()
        _abi.call('afsi_tb_sizes', bt, byref(total), byref(keys), byref(nbytes))
        if self.bound is None or self.bound[0] != (keys.value, nbytes.value):
            self.keys = torch.empty(max(total.value, 1), device=self.device, dtype=torch.int32)
            self.perm = torch.empty(max(total.value, 1), device=self.device, dtype=torch.int32)
            self.cell_ptr = torch.empty(keys.value + 1, device=self.device, dtype=torch.int32)
            self.workspace = torch.empty(nbytes.value, device=self.device, dtype=torch.uint8)
        bt.keys_sorted = self.keys.data_ptr()
        bt.perm = self.perm.data_ptr()
        bt.cell_ptr = self.cell_ptr.data_ptr()
        bt.workspace = self.workspace.data_ptr()
        bt.workspace_bytes = self.workspace.numel()
        bt.flags = self.flags.data_ptr()
        _abi.call('afsi_tb_prepare', bt, _abi.stream())
        self.items, self.batch = items, bt
        self.bound = ((keys.value, nbytes.value), tuple(t.data_ptr() for t in xs))
        return self

    def share_bins(self, other):
        """Use ``other``'s sorted keys, perm and cell_ptr (same instances)."""
        self.batch.keys_sorted = other.keys.data_ptr()
        self.batch.perm = other.perm.data_ptr()
        self.batch.cell_ptr = other.cell_ptr.data_ptr()
        self.batch.flags = other.flags.data_ptr()
        self.keys, self.perm, self.cell_ptr = other.keys, other.perm, other.cell_ptr
        return self

    def bin(self):
        _abi.call('afsi_tb_bin', self.batch, _abi.stream())

    def spread(self):
        _abi.call('afsi_tb_spread', self.batch, _abi.stream())

    def interpolate(self):
        _abi.call('afsi_tb_interp', self.batch, _abi.stream())


def ctypes_items(array):
    from ctypes import POINTER, cast
    return cast(array, POINTER(_abi.TBItem))


class BatchedIBStep:
    """One IB step for each of many 3D instances, one launch per phase over
    all of them: binning, spread, the projection's right-hand side (the IB
    predictor folded in), every multigrid phase of the mixed-precision
    pressure solve, the gradient correction, interpolation and the point
    advance. ``grids[i]`` is a ``MACGrid`` and ``counts[i]`` its number of
    points; all grids share the boundary type."""

    def __init__(self, grids, counts, *, dt, rho=1., device='cuda', options=None,
                 tile=DEFAULT_TILE, graphs=True):
        self.grids = list(grids)
        self.counts = [int(c) for c in counts]
        self.dt, self.rho = float(dt), float(rho)
        self.device = torch.device(device)
        self.solver = BatchedMixedMultigrid(self.grids, device=self.device, options=options,
                                            graphs=graphs)
        self.density = [g.zeros(device=self.device) for g in self.grids]
        self.point_velocity = [torch.empty((c, 3), device=self.device, dtype=torch.float64)
                               for c in self.counts]
        self.spreader = BatchedTransfer(self.grids, self.counts, device=self.device, tile=tile)
        self.gatherer = BatchedTransfer(self.grids, self.counts, device=self.device, tile=tile)
        n = len(self.grids)
        self.items = (_abi.IB3BItem * n)()
        self.ib = _abi.IB3BBatch()
        self.ib.count = n
        self.ib.items = ctypes.cast(self.items, ctypes.POINTER(_abi.IB3BItem))
        self.ib_workspace = None
        self._key = None

    def _bind(self, xs, forces, velocities):
        key = tuple(t.data_ptr() for t in (*xs, *forces, *(u for v in velocities for u in v)))
        if key == self._key:
            return
        self.spreader.bind(xs, forces, self.density, self.point_velocity)
        self.gatherer.bind(xs, forces, velocities, self.point_velocity)
        self.gatherer.share_bins(self.spreader)
        for i, g in enumerate(self.grids):
            it = self.items[i]
            for c in range(3):
                it.velocity[c] = velocities[i][c].data_ptr()
                it.force[c] = self.density[i][c].data_ptr()
            it.rhs = self.solver.b[i].data_ptr()
            it.pressure[0] = self.solver.x[i][0].data_ptr()
            it.pressure[1] = self.solver.x[i][1].data_ptr()
            it.x = xs[i].data_ptr()
            it.point_velocity = self.point_velocity[i].data_ptr()
            it.points = self.counts[i]
            it.grid = g.abi()
        nbytes = int(_abi.load().afsi_ib3b_workspace_bytes(ctypes.byref(self.ib)))
        if nbytes == 0:
            _abi.call('afsi_ib3b_prepare', self.ib, _abi.stream())  # raises with the reason
        if self.ib_workspace is None or self.ib_workspace.numel() < nbytes:
            self.ib_workspace = torch.empty(nbytes, device=self.device, dtype=torch.uint8)
        self.ib.workspace = self.ib_workspace.data_ptr()
        self.ib.workspace_bytes = self.ib_workspace.numel()
        _abi.call('afsi_ib3b_prepare', self.ib, _abi.stream())
        self._key = key

    def _initial(self, pressures):
        """Index of the solver buffer pair holding the initial guesses: the
        previous solution when ``pressures`` are this stepper's own views of
        it, else copies of ``pressures`` (zeros for None) in x[i][0]."""
        sol = self.solver
        own = pressures is not None and all(
            p is not None and p.data_ptr() == sol.x[i][sol.cur].data_ptr() for i, p in enumerate(pressures))
        if own:
            return sol.cur
        for i in range(len(self.grids)):
            p = None if pressures is None else pressures[i]
            if p is None:
                sol.x[i][0].zero_()
            else:
                if p.shape != sol.x[i][0].shape or p.dtype != torch.float64:
                    raise ValueError('invalid initial pressure')
                sol.x[i][0].copy_(p)
        return 0

    @torch.no_grad()
    def step(self, xs, forces, velocities, pressures=None, *, advance=True):
        """Returns (velocities, pressures, point_velocities, infos). The
        velocity tensors are updated in place; the pressures are views of
        the solver's buffers, valid until the next step (pass them back as
        ``pressures`` to warm-start it without a copy)."""
        if len(xs) != len(self.grids) or len(forces) != len(self.grids) or len(velocities) != len(self.grids):
            raise ValueError('one point set, force set and velocity per instance')
        a = self.dt / self.rho
        self._bind(xs, forces, velocities)
        start = self._initial(pressures)
        self.spreader.bin()
        self.spreader.spread()
        _abi.call('afsi_ib3b_rhs', self.ib, a, -self.rho / self.dt, _abi.stream())
        infos = self.solver.solve_resident(start)
        cur = self.solver.cur
        _abi.call('afsi_ib3b_correct', self.ib, cur, a, a, _abi.stream())
        self.gatherer.interpolate()
        if advance:
            _abi.call('afsi_ib3b_advance', self.ib, self.dt, _abi.stream())
        ps = [self.solver.x[i][cur] for i in range(len(self.grids))]
        return velocities, ps, self.point_velocity, infos


class ConcurrentGroups:
    """Runs independent groups of a batch (e.g. the 3D ``BatchedIBStep`` and
    the 2D ``BatchedChannelStep`` of one mix) each on its own stream and host
    thread, so that their launches, host convergence checks and kernels
    overlap on the device.

    ``run(*steps)`` calls each step (a callable) with its stream current.
    Each step starts after the work already queued on the caller's current
    stream, and that stream waits for all of them; the results come back in
    order. A step must not capture graphs or synchronise the whole device
    (the batched steps synchronise only their own stream), and the groups
    must not share buffers. Streams keep the default priority."""

    def __init__(self, device='cuda', count=2):
        self.device = torch.device(device)
        if self.device.index is None:  # the worker threads set an indexed device
            self.device = torch.device('cuda', torch.cuda.current_device())
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(count)]
        self.pool = ThreadPoolExecutor(max_workers=count)

    def run(self, *steps):
        if len(steps) > len(self.streams):
            raise ValueError(f'at most {len(self.streams)} concurrent groups')
        origin = torch.cuda.current_stream(self.device)

        def call(step, stream):
            torch.cuda.set_device(self.device)
            stream.wait_stream(origin)
            with torch.cuda.stream(stream):
                return step()

        futures = [self.pool.submit(call, step, stream) for step, stream in zip(steps, self.streams)]
        results = [f.result() for f in futures]
        for stream in self.streams[:len(steps)]:
            origin.wait_stream(stream)
        return results

    def close(self):
        self.pool.shutdown()

