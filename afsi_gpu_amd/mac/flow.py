"""MAC projection on the GPU: ``MACFlow.project`` of mac/flow.py:101-120.

rhs = -(rho/dt) D u* (HIP divergence kernel), p = multigrid solve with the
previous pressure as the initial guess, u = u* - (dt/rho) G p (HIP gradient
correction; Neumann boundary-normal faces untouched, mac/grid.py:74-83).
The predictor ``step`` (convection, viscosity) is outside this hot path.
``precision='mixed'`` solves the pressure with FP32 V-cycles inside FP64
iterative refinement (divergence, gradient and velocities stay FP64).
"""
from dataclasses import dataclass
from math import isfinite
import torch

from .. import _abi
from .multigrid import GeometricMultigrid, MixedMultigrid


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


def correct(velocity, pressure, grid, coef):
    """In place u_c -= coef * G_c p (mac/flow.py:66-67)."""
    _abi.call('afsi_grad_correct_3d_f64', *(_abi.ptr(u) for u in velocity), _abi.ptr(pressure),
              _abi.i64x3(grid.shape), _abi.f64x3(grid.spacing), float(coef), int(grid.periodic),
              _abi.stream())
    return velocity


class MACFlow:
    def __init__(self, grid, *, dt, rho=1., mu=1., device='cuda', dtype=torch.float64,
                 options=None, pressure_backend='graph', precision='fp64'):
        if any(not isfinite(v) or v <= 0 for v in (dt, rho, mu)):
            raise ValueError('this explicit-viscous MAC scheme requires positive dt/rho/mu')
        if dtype != torch.float64:
            raise ValueError('the plain projection is FP64')
        self.grid, self.dt, self.rho, self.mu = grid, float(dt), float(rho), float(mu)
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

    @torch.no_grad()
    def project(self, tentative, initial=None):
        """mac/flow.py:283-302 (torch branch): finite and closed-wall checks,
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
