"""Channel projection on the GPU: ``ChannelFlow.project`` of
mac2d/flow.py:92-103.

rhs = -(rho/dt) D u* (HIP divergence kernel), p = channel multigrid with
the previous pressure as the initial guess, u = u* - (dt/rho) G p with the
outlet gradient -2p/hx (HIP correction kernel into fresh tensors). The
explicit predictor ``ChannelFlow.step`` (convection, viscosity, inlet
profile) is outside this hot path.
"""
from dataclasses import dataclass
from math import isfinite
import torch

from .. import _abi
from .multigrid import ChannelMultigrid


@dataclass(frozen=True)
class ChannelResult:
    velocity: tuple
    pressure: torch.Tensor
    diagnostics: dict


def _n(grid):
    return (_abi.c_int64 * 2)(*grid.shape)


def _h(grid):
    return (_abi.c_double * 2)(*grid.spacing)


def divergence(velocity, grid, scale=1., out=None):
    """scale * (diff(u)/hx + diff(v)/hy) (mac2d/grid.py:46-48), FP64."""
    if out is None:
        out = torch.empty(grid.shape, device=velocity[0].device, dtype=torch.float64)
    _abi.call('afsi_divergence_2d_f64', _abi.ptr(velocity[0]), _abi.ptr(velocity[1]),
              _abi.ptr(out), _n(grid), _h(grid), float(scale), int(grid.periodic), _abi.stream())
    return out


def correct(velocity, pressure, grid, coef):
    """In place u -= coef * G p (mac2d/flow.py:88-90)."""
    _abi.call('afsi_grad_correct_2d_f64', _abi.ptr(velocity[0]), _abi.ptr(velocity[1]),
              _abi.ptr(pressure), _n(grid), _h(grid), float(coef), int(grid.periodic),
              _abi.stream())
    return velocity


class ChannelFlow:
    def __init__(self, grid, *, dt=1 / 16000, rho=1., mu=.1, device='cuda', options=None,
                 pressure_backend='graph', dtype=torch.float64):
        if any(not isfinite(v) or v <= 0 for v in (dt, rho, mu)):
            raise ValueError('positive finite dt/rho/mu required')
        self.grid, self.dt, self.rho, self.mu = grid, dt, rho, mu
        self.pressure_solver = ChannelMultigrid(grid, device=device, dtype=dtype, options=options,
                                                backend=pressure_backend)
        self._rhs = torch.empty(grid.shape, device=device, dtype=torch.float64)

    @torch.no_grad()
    def project(self, tentative, initial=None):
        """mac2d/flow.py:92-103 (reference branch checks and messages)."""
        self.grid.check_velocity(tentative)
        if any(not torch.isfinite(v).all() for v in tentative):
            raise ValueError('finite tentative velocity required')
        if tentative[1][:, 0].count_nonzero() or tentative[1][:, -1].count_nonzero():
            raise ValueError('normal wall velocities must be zero')
        tentative = tuple(u.contiguous() for u in tentative)
        rhs = divergence(tentative, self.grid, -self.rho / self.dt, out=self._rhs)
        p, info = self.pressure_solver.solve(rhs, initial)
        velocity = correct(tuple(u.clone() for u in tentative), p, self.grid, self.dt / self.rho)
        return ChannelResult(velocity, p, dict(pressure=info))
