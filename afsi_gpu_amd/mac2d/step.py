"""One IB time step on the channel (the ordering of mac2d/coupling.py:66-109
with the point force given): prepare -> spread -> u* = u + (dt/rho) f with
the wall normal faces kept at zero and the inlet face kept at its value
(the boundary() of mac2d/grid.py:74-79) -> project -> gather -> x += dt U.
"""
import torch

from .flow import ChannelFlow
from . import transfer


class ChannelIBStep:
    def __init__(self, grid, *, dt, rho=1., device='cuda', options=None, pressure_backend='graph'):
        self.grid, self.dt, self.rho = grid, float(dt), float(rho)
        self.flow = ChannelFlow(grid, dt=dt, rho=rho, device=device, options=options,
                                pressure_backend=pressure_backend)

    @torch.no_grad()
    def step(self, x, force, velocity, pressure=None, *, advance=True):
        """Returns (velocity, pressure, point_velocity, info); x advances in place."""
        stencil = transfer.prepare(self.grid, x)
        density = transfer.spread(self.grid, force, stencil)
        coef = self.dt / self.rho
        u = velocity[0] + coef * density[0]
        v = velocity[1] + coef * density[1]
        if not self.grid.periodic:
            u[0] = velocity[0][0]
        v[:, 0] = 0
        v[:, -1] = 0
        result = self.flow.project((u, v), pressure)
        U = transfer.gather(self.grid, result.velocity, stencil)
        if advance:
            x.add_(U, alpha=self.dt)
        return result.velocity, result.pressure, U, result.diagnostics
