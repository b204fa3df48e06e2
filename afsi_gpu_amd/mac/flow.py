"""MAC projection on the GPU: ``MACFlow.project`` of mac/flow.py:101-120.

rhs = -(rho/dt) D u* (HIP divergence kernel), p = multigrid solve with the
previous pressure as the initial guess, u = u* - (dt/rho) G p (HIP gradient
correction; Neumann boundary-normal faces untouched, mac/grid.py:74-83).
``step`` adds the explicit closed-box predictor of mac/flow.py:59-64,122-157
(HIP kernel ``afsi_mac_predict_f64``: centered conservative convection,
no-slip velocity Laplacian, force term) with the transport screen of
transport.py:9-73 in front of the projection.
``precision='mixed'`` solves the pressure with FP32 V-cycles inside FP64
iterative refinement (divergence, gradient and velocities stay FP64).
"""
from dataclasses import dataclass
from math import isfinite
import torch

from .. import _abi
from .multigrid import GeometricMultigrid, MixedMultigrid


COURANT_LIMIT = .25             # transport.py:9-11
VISCOUS_LIMIT = .25
ADVECTION_DIFFUSION_LIMIT = .25
TRANSPORT_POLICY = 'centered-explicit-advection-diffusion-v1'


def transport_policy():
    """transport.py:53-59."""
    return dict(name=TRANSPORT_POLICY, courant_limit=COURANT_LIMIT, viscous_number_limit=VISCOUS_LIMIT,
                advection_diffusion_limit=ADVECTION_DIFFUSION_LIMIT,
                advection_diffusion_formula='dt*sum(component_max_abs_velocity**2)/(2*mu/rho)',
                cell_reynolds_monitor_only=True,
                scope='frozen-coefficient transport screen; not full nonlinear IB-FSI stability or spatial accuracy')


def transport_violations(courant, advection_diffusion_number, viscous_number):
    """transport.py:69-73."""
    return [name for name, value, limit in
            (('courant', courant, COURANT_LIMIT),
             ('advection_diffusion', advection_diffusion_number, ADVECTION_DIFFUSION_LIMIT),
             ('viscous', viscous_number, VISCOUS_LIMIT)) if value > limit]


class MACTransportGuardError(ValueError):
    """A conservative transport check stopped the step before its predictor
    (mac/flow.py:11-24)."""

    def __init__(self, diagnostics):
        self.diagnostics = diagnostics
        super().__init__(
            'MAC centered-advection stability guard exceeded: '
            f'CFL={diagnostics["courant"]:.9g} (limit {COURANT_LIMIT}), '
            f'advection_diffusion={diagnostics["advection_diffusion_number"]:.9g} '
            f'(limit {ADVECTION_DIFFUSION_LIMIT}), '
            f'cell_Re={diagnostics["cell_reynolds"]:.9g} (monitor only); '
            f'triggered={",".join(diagnostics["triggered"])}. '
            'Reduce dt for a time-step violation. '
            'Passing this transport screen does not establish nonlinear FSI stability.')


@dataclass(frozen=True)
class MACFlowResult:
    velocity: tuple
    pressure: torch.Tensor
    diagnostics: dict


def divergence(velocity, grid, scale=1., out=None):
    """scale * sum_c diff(u_c)/h_c (mac/grid.py:70-71), FP64."""
    if out is None:
        out = torch.empty(grid.shape, device=velocity[0].device, dtype=torch.float64)
    _abi.call('afsi_divergence_3d_f64', *(_abi.ptr(u) for u in velocity), _abi.ptr(out),
              _abi.i64x3(grid.shape), _abi.f64x3(grid.spacing), float(scale),
              int(grid.periodic), _abi.stream())
    return out


def project_rhs(velocity, force, a, grid, scale, out):
    """scale * D(u + a f) with u left unchanged (the IB predictor folded into
    the projection's right-hand side)."""
    _abi.call('afsi_project_rhs_3d_f64', *(_abi.ptr(u) for u in velocity), *(_abi.ptr(f) for f in force),
              float(a), _abi.ptr(out), _abi.i64x3(grid.shape), _abi.f64x3(grid.spacing), float(scale),
              int(grid.periodic), _abi.stream())
    return out


def project_correct(velocity, force, a, pressure, grid, coef):
    """In place u_c = (u_c + a f_c) - coef * G_c p."""
    _abi.call('afsi_project_correct_3d_f64', *(_abi.ptr(u) for u in velocity), *(_abi.ptr(f) for f in force),
              float(a), _abi.ptr(pressure), _abi.i64x3(grid.shape), _abi.f64x3(grid.spacing), float(coef),
              int(grid.periodic), _abi.stream())
    return velocity


def correct(velocity, pressure, grid, coef):
    """In place u_c -= coef * G_c p (mac/flow.py:66-67)."""
    _abi.call('afsi_grad_correct_3d_f64', *(_abi.ptr(u) for u in velocity), _abi.ptr(pressure),
              _abi.i64x3(grid.shape), _abi.f64x3(grid.spacing), float(coef), int(grid.periodic),
              _abi.stream())
    return velocity


class MACFlow:
    def __init__(self, grid, *, dt, rho=1., mu=1., device='cuda', dtype=torch.float64,
                 options=None, pressure_backend='graph', precision='fp64', predictor=True):
        """Arguments of mac/flow.py:35-36. Build-own: ``precision`` ('fp64' or
        'mixed' pressure solve) and ``predictor`` (False: projection only, as
        the IB steps here use it; no explicit-viscous check, no ``step``)."""
        if any(not isfinite(v) or v <= 0 for v in (dt, rho, mu)):
            raise ValueError('this explicit-viscous MAC scheme requires positive dt/rho/mu')
        if dtype != torch.float64:
            raise ValueError('the plain projection is FP64')
        self.grid, self.dt, self.rho, self.mu = grid, float(dt), float(rho), float(mu)
        self.viscous_number = dt * mu / rho * sum(1 / h ** 2 for h in grid.spacing)
        self.predictor = bool(predictor)
        if self.predictor and self.viscous_number > VISCOUS_LIMIT:
            raise ValueError('explicit viscous time step too large; reduce dt')
        if precision not in ('fp64', 'mixed'):
            raise ValueError('precision must be fp64 or mixed')
        self.precision = precision
        if precision == 'mixed':
            self.pressure_solver = MixedMultigrid(grid, device=device, options=options,
                                                  backend=pressure_backend)
        else:
            self.pressure_solver = GeometricMultigrid(grid, device=device, dtype=dtype,
                                                      options=options, backend=pressure_backend)
        self._rhs = torch.empty(grid.shape, device=device, dtype=dtype)
        self._speeds = None
        self._star = None

    @torch.no_grad()
    def project(self, tentative, initial=None):
        """mac/flow.py:102-121 (torch branch): finite and closed-wall checks,
        pressure solve, gradient correction into fresh velocity tensors."""
        self.grid.check_velocity(tentative)
        for c, u in enumerate(tentative):
            if not torch.isfinite(u).all():
                raise ValueError('nonfinite tentative velocity')
            if not self.grid.periodic:
                lo = u.narrow(c, 0, 1)
                hi = u.narrow(c, u.shape[c] - 1, 1)
                if lo.count_nonzero() or hi.count_nonzero():
                    raise ValueError('closed-box normal velocity must be zero before projection')
        tentative = tuple(u.contiguous() for u in tentative)
        rhs = divergence(tentative, self.grid, -self.rho / self.dt, out=self._rhs)
        p, info = self.pressure_solver.solve(rhs, initial)
        velocity = correct(tuple(u.clone() for u in tentative), p, self.grid, self.dt / self.rho)
        return MACFlowResult(velocity, p, dict(pressure=info))

    def _transport(self, velocity, density):
        """finite flag and the transport numbers of mac/flow.py:137-146 from one
        device reduction (max |u_c| per component, non-finite count)."""
        if self._speeds is None:
            dev = velocity[0].device
            self._speeds = torch.empty(4, device=dev, dtype=torch.float64)
            self._speeds_ws = torch.empty(int(_abi.load().afsi_mac_speeds_workspace_bytes()), dtype=torch.uint8,
                                          device=dev)
        _abi.call('afsi_mac_speeds_f64', *(_abi.ptr(u) for u in velocity), *(_abi.ptr(f) for f in density),
                  _abi.i64x3(self.grid.shape), _abi.ptr(self._speeds), _abi.ptr(self._speeds_ws),
                  self._speeds_ws.numel(), _abi.stream())
        s0, s1, s2, bad = self._speeds.tolist()
        speeds = (s0, s1, s2)
        h = self.grid.spacing
        nu = self.mu / self.rho
        courant = sum(self.dt * s / hh for s, hh in zip(speeds, h))
        reynolds = max(s * hh / nu for s, hh in zip(speeds, h))
        advection_diffusion = self.dt * sum(s * s for s in speeds) / (2 * nu)
        return bad == 0, speeds, courant, reynolds, advection_diffusion

    @torch.no_grad()
    def step(self, velocity, density, *, pressure_initial=None):
        """mac/flow.py:122-157 (closed box): finiteness and transport screen,
        explicit predictor u* = zero_normal(u + dt(-N(u) + (mu/rho) L u + f/rho)),
        then ``project``."""
        if not self.predictor:
            raise ValueError('this MACFlow was built for projection only (predictor=False)')
        if self.grid.periodic:
            raise ValueError('the explicit MAC predictor is defined for the closed box')
        self.grid.check_velocity(velocity)
        self.grid.check_velocity(density)
        velocity = tuple(u.contiguous() for u in velocity)
        density = tuple(f.contiguous() for f in density)
        if any(u.dtype != torch.float64 for u in (*velocity, *density)):
            raise ValueError('MAC field shape/device/dtype mismatch')
        finite, speeds, courant, reynolds, advection_diffusion = self._transport(velocity, density)
        if not finite:
            raise ValueError('finite MAC velocity and force density required')
        triggered = transport_violations(courant, advection_diffusion, self.viscous_number)
        if triggered:
            raise MACTransportGuardError(dict(
                courant=courant, cell_reynolds=reynolds, advection_diffusion_number=advection_diffusion,
                **transport_policy(), triggered=triggered, component_max_abs_velocity=list(speeds),
                spacing=list(self.grid.spacing), dt=self.dt, rho=self.rho, mu=self.mu,
                viscous_number=self.viscous_number))
        if self._star is None:
            self._star = tuple(torch.empty_like(u) for u in velocity)
        star = self._star
        _abi.call('afsi_mac_predict_f64', *(_abi.ptr(u) for u in velocity), *(_abi.ptr(f) for f in density),
                  *(_abi.ptr(u) for u in star), _abi.i64x3(self.grid.shape), _abi.f64x3(self.grid.spacing),
                  self.dt, self.rho, self.mu, _abi.stream())
        result = self.project(star, pressure_initial)
        return MACFlowResult(result.velocity, result.pressure,
                             dict(result.diagnostics, courant=courant, cell_reynolds=reynolds,
                                  advection_diffusion_number=advection_diffusion,
                                  viscous_number=self.viscous_number))
