"""One IB time step on the GPU: bin -> spread -> predictor -> projection ->
interpolate -> advance points, all FP64 ("plain" path).

Ordering follows the reference's coupled step (mac/coupling.py:74-85):
spread the lagged point forces, project u* = u + (dt/rho) f with the
previous pressure as the initial guess (mac/flow.py:101-120), interpolate
the projected velocity and move the points x += dt U. Point forces are
given directly (the FE mass solve and predictor physics are off-path).
All work is on the current stream. The only host synchronisations are the
multigrid convergence checks; the binning and spread error flags are copied
to pinned host memory on the stream and read at the solver's first
synchronisation, before any velocity is changed, so a point whose support
reaches a closed wall or is nonfinite raises the reference's ValueError
(mac/compact_transfer.py:94-96,115-117) with the step's state untouched.
"""
import torch

from . import _abi
from .mac.flow import MACFlow, correct, divergence, project_correct, project_rhs
from .mac.multigrid import MIXED_INNER_CYCLES, MIXED_RELAX, MGOptions
from .mac.transfer import DEFAULT_TILE, Binner, PointBins, TiledTransfer, raise_flags


class IBStep:
    def __init__(self, grid, points, *, dt, rho=1., device='cuda', options=None,
                 tile=DEFAULT_TILE, pressure_backend='graph', precision='fp64', resident_pressure=False,
                 graphs=False):
        """``resident_pressure``: the projection's right-hand side is formed
        in the solver's own buffer and the returned pressure is the solver's
        iterate (valid until the next step; passed back as ``pressure`` it
        warm-starts the next solve without a copy).

        ``graphs`` (needs ``resident_pressure``): the step runs as two HIP
        graphs around the solver's host convergence checks, each captured
        once per set of buffers and solver state and replayed afterwards:
        binning, spread, right-hand side and the solver's start (metrics,
        initial residual and the first outer iteration) in one; the solver's
        finish, correction, interpolation and point advance in the other.
        The V-cycle blocks in between are the solver's own graphs. Steps
        given ``timers`` run as plain launches."""
        self.grid, self.dt, self.rho = grid, float(dt), float(rho)
        self.resident = bool(resident_pressure)
        if graphs and not self.resident:
            raise ValueError('graph steps need resident_pressure')
        self.graphs = {} if graphs else None
        self.device = torch.device(device)
        self.count = int(points)
        self.options = options or (MGOptions(smooth=2, smoother='rbgs', relax=MIXED_RELAX,
                                             inner_cycles=MIXED_INNER_CYCLES) if precision == 'mixed'
                                   else MGOptions(smooth=2, smoother='rbgs'))
        self.precision = precision
        self.flow = MACFlow(grid, dt=dt, rho=rho, device=device, options=self.options,
                            pressure_backend=pressure_backend, precision=precision, predictor=False)
        self.binner = Binner(grid, self.count, self.device, tile)
        # one error word for binning and both transfers (cleared by the binner)
        self.transfer = TiledTransfer(grid, self.device, tile, flags=self.binner.flags)
        self.flags_host = torch.zeros((), dtype=torch.int32).pin_memory()
        self.force_density = grid.zeros(device=self.device)
        self.rhs = torch.empty(grid.shape, device=self.device, dtype=torch.float64)
        self.point_velocity = torch.empty((self.count, 3), device=self.device, dtype=torch.float64)

    def _check_flags(self):
        raise_flags(int(self.flags_host))

    def step(self, x, force, velocity, pressure=None, *, advance=True, timers=None):
        """Returns (velocity, pressure, point_velocity, info); ``velocity``
        tensors are updated in place to the projected field; with
        ``advance`` the points ``x`` move by dt*U in place. ``timers`` maps
        phase names ('bin', 'spread', 'project', 'interp') to (start, end)
        events recorded on the current stream around them."""
        if self.graphs is not None and not timers:
            return self._graph_step(x, force, velocity, pressure, advance)
        mark = _Marks(timers)
        mark('bin', 0)
        bins = self.binner(x)
        mark('bin', 1)
        mark('spread', 0)
        f = self.transfer.spread(x, force, bins, out=self.force_density)
        self.flags_host.copy_(self.binner.flags, non_blocking=True)
        mark('spread', 1)
        mark('project', 0)
        a = self.dt / self.rho
        solver = self.flow.pressure_solver
        out = self.rhs
        if self.resident:
            out = solver.b if hasattr(solver, 'b') else solver.rhs[0]
        rhs = project_rhs(velocity, f, a, self.grid, -self.rho / self.dt, out=out)
        self.last_rhs = rhs
        p, info = solver.solve(rhs, pressure, precheck=self._check_flags, view=self.resident)
        project_correct(velocity, f, a, p, self.grid, a)
        mark('project', 1)
        mark('interp', 0)
        U = self.transfer.interpolate(x, velocity, bins, out=self.point_velocity)
        mark('interp', 1)
        if advance:
            _abi.call('afsi_axpy_f64', _abi.ptr(x), _abi.ptr(U), x.numel(), self.dt, _abi.stream())
        return velocity, p, U, dict(pressure=info, bins=bins)


    # -- graph steps ------------------------------------------------------------
    def _capture(self, fn):
        stream = torch.cuda.Stream(device=self.device)
        stream.wait_stream(torch.cuda.current_stream())
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=stream):
            fn()
        torch.cuda.current_stream().wait_stream(stream)
        return graph

    def _launch_start(self, x, force, velocity, zero, start, lead):
        """Binning, spread, right-hand side into the solver's buffer and the
        solver's start (device work only)."""
        solver = self.flow.pressure_solver
        bins = self.binner(x)
        f = self.transfer.spread(x, force, bins, out=self.force_density)
        mixed = hasattr(solver, 'xs')
        out = solver.b if mixed else solver.rhs[0]
        project_rhs(velocity, f, self.dt / self.rho, self.grid, -self.rho / self.dt, out=out)
        if mixed:
            if zero:
                solver.xs[start].zero_()
            solver.launch_start(start, lead, captured=True, zero=zero)
        else:
            if zero:
                solver.p[0].zero_()
            solver.launch_start(zero=zero)

    def _launch_finish(self, x, velocity, pressure, advance):
        """The solver's finish, correction, interpolation and advance."""
        a = self.dt / self.rho
        self.flow.pressure_solver.launch_finish()
        project_correct(velocity, self.force_density, a, pressure, self.grid, a)
        U = self.transfer.interpolate(x, velocity, self._bins, out=self.point_velocity)
        if advance:
            _abi.call('afsi_axpy_f64', _abi.ptr(x), _abi.ptr(U), x.numel(), self.dt, _abi.stream())

    def _graph_step(self, x, force, velocity, pressure, advance):
        if not (x.is_cuda and force.is_cuda and x.device == force.device):
            raise ValueError('positions and forces must be on the step\'s GPU')
        if x.shape != (self.count, 3) or x.dtype != torch.float64 or not x.is_contiguous():
            raise ValueError('positions must match the step (N,3) contiguous float64')
        if force.shape != x.shape or force.dtype != torch.float64 or not force.is_contiguous():
            raise ValueError('forces must match the positions')
        if len(velocity) != 3 or any(u.dtype != torch.float64 or not u.is_contiguous() or u.device != x.device or
                                     tuple(u.shape) != self.grid.face_shape(c) for c, u in enumerate(velocity)):
            raise ValueError('velocity must be the three contiguous float64 face arrays')
        solver = self.flow.pressure_solver
        mixed = hasattr(solver, 'xs')
        buffers = (x.data_ptr(), force.data_ptr(), *(u.data_ptr() for u in velocity))
        own = solver.xs if mixed else (solver.p[0],)
        index = [k for k, t in enumerate(own) if pressure is not None and pressure.data_ptr() == t.data_ptr()]
        if pressure is not None and not index:
            if pressure.shape != own[0].shape or pressure.dtype != torch.float64:
                raise ValueError('invalid initial pressure')
            own[0].copy_(pressure)
        start = index[0] if index else 0
        lead = solver.lead() if mixed else 0
        key = ('start', buffers, start, lead, pressure is None)
        graph = self.graphs.get(key)
        if graph is None:
            graph = self.graphs[key] = self._capture(
                lambda: self._launch_start(x, force, velocity, pressure is None, start, lead))
        self._bins = PointBins(self.binner.tile, self.binner.keys, self.binner.perm, self.binner.cell_ptr,
                               self.binner.flags)
        graph.replay()
        self.last_rhs = solver.b if mixed else solver.rhs[0]
        self.flags_host.copy_(self.binner.flags, non_blocking=True)
        rtol = solver.options.rtol
        zero = pressure is None
        if mixed:
            cycles, residual, tolerance = solver.settle(start, lead, rtol, precheck=self._check_flags, zero=zero)
            cycles *= solver.options.inner_cycles
            end = solver.cur
        else:
            cycles, residual, tolerance = solver.settle(rtol, precheck=self._check_flags, zero=zero)
            end = 0
        p = own[end]
        key = ('finish', buffers, end, bool(advance))
        graph = self.graphs.get(key)
        if graph is None:
            graph = self.graphs[key] = self._capture(lambda: self._launch_finish(x, velocity, p, advance))
        graph.replay()
        info = dict(cycles=cycles, residual_norm=residual, tolerance=tolerance, backend=solver.backend)
        return velocity, p, self.point_velocity, dict(pressure=info, bins=self._bins)


class _Marks:
    def __init__(self, timers):
        self.timers = timers

    def __call__(self, name, which):
        if self.timers and name in self.timers:
            self.timers[name][which].record(torch.cuda.current_stream())


class FEIBStep:
    """One coupled step with the FE interaction-point transfer, in the order
    of MACIBStepper.step (mac/coupling.py:74-123): prepare the stencil at
    the interaction points B x, spread the nodal force through the mass solve
    (M F = b, f = V^-1 K^T W B F), u* = u + (dt/rho) f, project, interpolate
    through the mass solve (M U = B^T W K u), x += dt U with the grid
    displacement guard. Nodal forces are given (the solid model is off the
    path). ``predictor='force'`` uses u* = u + (dt/rho) f, as ``IBStep``;
    ``predictor='full'`` runs ``MACFlow.step`` (transport screen, explicit
    convection and viscosity, projection; closed box), the reference's
    flow step."""

    def __init__(self, grid, geometry, *, dt, rho=1., mu=1., device='cuda', options=None, transfer_options=None,
                 warm_start=False, max_displacement=.25, tile=DEFAULT_TILE, pressure_backend='graph',
                 precision='fp64', predictor='force', fused=False):
        from .mac.fe import FETransfer
        if not max_displacement > 0 or max_displacement != max_displacement:
            raise ValueError('max_displacement must be finite and positive')
        if predictor not in ('force', 'full'):
            raise ValueError("predictor must be 'force' or 'full'")
        self.grid, self.dt, self.rho = grid, float(dt), float(rho)
        self.device = torch.device(device)
        self.max_displacement = float(max_displacement)
        self.options = options or (MGOptions(smooth=2, smoother='rbgs', relax=MIXED_RELAX,
                                             inner_cycles=MIXED_INNER_CYCLES) if precision == 'mixed'
                                   else MGOptions(smooth=2, smoother='rbgs'))
        self.predictor = predictor
        self.flow = MACFlow(grid, dt=dt, rho=rho, mu=mu, device=device, options=self.options,
                            pressure_backend=pressure_backend, precision=precision,
                            predictor=predictor == 'full')
        self.transfer = FETransfer(grid, geometry, options=transfer_options, warm_start=warm_start, tile=tile,
                                   fused=fused)
        self.rhs = torch.empty(grid.shape, device=self.device, dtype=torch.float64)
        self._h = torch.tensor(grid.spacing, device=self.device, dtype=torch.float64)

    def step(self, x, force, velocity, pressure=None, *, diagnostics=False, timers=None):
        """Returns (x_new, velocity, pressure, nodal_velocity, info); the
        velocity tensors are updated in place to the projected field."""
        mark = _Marks(timers)
        mark('bin', 0)
        stencil = self.transfer.prepare(x)
        mark('bin', 1)
        mark('spread', 0)
        f, spread_info = self.transfer.spread(force, stencil)
        mark('spread', 1)
        mark('project', 0)
        if self.predictor == 'full':   # MACFlow.step: transport screen, predictor, projection
            result = self.flow.step(velocity, f, pressure_initial=pressure)
            for uc, vc in zip(velocity, result.velocity):
                uc.copy_(vc)
            p, pinfo = result.pressure, result.diagnostics
        else:
            a = self.dt / self.rho
            for uc, fc in zip(velocity, f):
                _abi.call('afsi_axpy_f64', _abi.ptr(uc), _abi.ptr(fc), uc.numel(), a, _abi.stream())
            rhs = divergence(velocity, self.grid, -self.rho / self.dt, out=self.rhs)
            p, pinfo = self.flow.pressure_solver.solve(rhs, pressure)
            correct(velocity, p, self.grid, a)
        mark('project', 1)
        mark('interp', 0)
        U, interp_info = self.transfer.interpolate(velocity, stencil)
        mark('interp', 1)
        increment = self.dt * U
        fraction = float((increment.abs() / self._h).max())
        if not fraction == fraction or fraction > self.max_displacement:
            raise ValueError('MAC solid displacement exceeds the grid-based limit; reduce dt')
        info = dict(pressure=pinfo, force_mass=spread_info, velocity_mass=interp_info,
                    max_grid_displacement=fraction)
        if diagnostics:
            solid_power = (U * force).sum()
            fluid_power = self.grid.volume * sum((u * fc).sum() for u, fc in zip(velocity, f))
            info.update(solid_power=float(solid_power), fluid_power=float(fluid_power),
                        power_error=abs(float(solid_power - fluid_power)))
        return x + increment, velocity, p, U, info
