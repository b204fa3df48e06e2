"""Oracle: one spread -> project -> interpolate step (FP64).

Test infrastructure only (see ``oracle/__init__.py``).

Follows the ordering of the reference's coupled step (mac/coupling.py:74-85:
prepare -> spread -> flow step -> interpolate -> x + dt*U) with the point
forces given directly (no FE mass solve) and the predictor reduced to the
force term u* = u + (dt/rho) f (convection/viscosity are not on this path):
  f = spread(F)                      mac/compact_transfer.py:172-185
  p = MG(-(rho/dt) D u*, p_prev)     mac/flow.py:101-120
  u = u* - (dt/rho) G p
  U = gather(u)                      mac/compact_transfer.py:187-201
  x_new = x + dt U                   mac/coupling.py:84-85
"""
import numpy as np

from . import ib3d, mac3d


def ib_step(x, F, u, p_prev, grid, mg, *, dt, rho=1.):
    base, phi = ib3d.prepare_compact(x, grid)
    f = ib3d.spread_grid(F, base, phi, grid)
    star = tuple(uc + dt / rho * fc for uc, fc in zip(u, f))
    velocity, p, info = mac3d.project(star, grid.spacing, mg, dt=dt, rho=rho, initial=p_prev)
    U = ib3d.gather_grid(velocity, base, phi, grid)
    return dict(force=f, velocity=velocity, pressure=p, point_velocity=U,
                x=x + dt * U, info=info, tentative=star)


def ib_step_2d(x, F, velocity, p_prev, channel, mg, *, dt, rho=1.):
    """The channel step (mac2d/coupling.py:66-109 ordering, point forces
    given): reflected stencil, spread, u* = u + (dt/rho) f with the
    boundary() of mac2d/grid.py:74-79 (inlet face kept, wall normal faces
    zero), projection, gather, x + dt U."""
    from . import mac2d
    if not mac2d.support_ok(x, channel):
        raise ValueError('valve IB points leave the channel support region')
    ids, w = mac2d.prepare(x, channel)
    f = mac2d.spread(F, ids, w, channel)
    u = velocity[0] + dt / rho * f[0]
    v = velocity[1] + dt / rho * f[1]
    if not channel.periodic:
        u[0] = velocity[0][0]
    v[:, 0] = 0
    v[:, -1] = 0
    (pu, pv), p, info = mac2d.project((u, v), channel.spacing, mg, dt=dt, rho=rho, initial=p_prev)
    U = mac2d.gather((pu, pv), ids, w)
    return dict(force=f, velocity=(pu, pv), pressure=p, point_velocity=U, x=x + dt * U,
                info=info, tentative=(u, v), ids=ids, weights=w)


def fe_ib_step(g, x, F, u, p_prev, grid, mg, *, dt, rho=1., options=None, full=False, mu=1.):
    """The coupled step with the FE quadrature transfer (mac/coupling.py:74-85,
    mac/transfer.py:104-140): M C = F, f = V^-1 K^T W B C at the interaction
    points B x, projection of u + (dt/rho) f, M U = B^T W K u, x + dt U."""
    from . import fe
    options = options or fe.SolverOptions()
    M = fe.mass_matrix(g)
    diag = fe.mass_diagonal(g)
    base, phi = ib3d.prepare_compact(fe.interaction_points(g, x), grid)
    C, spread_info = fe.pcg(lambda v: M @ v, F, diag, None, options)
    f = ib3d.spread_grid(fe.weighted(g, C), base, phi, grid)
    if full:   # MACFlow.step (mac/flow.py:122-157)
        velocity, p, info = mac3d.flow_step(u, f, grid.spacing, mg, dt=dt, rho=rho, mu=mu, initial=p_prev)
    else:
        star = tuple(uc + dt / rho * fc for uc, fc in zip(u, f))
        velocity, p, info = mac3d.project(star, grid.spacing, mg, dt=dt, rho=rho, initial=p_prev)
    rhs = fe.assemble(g, ib3d.gather_grid(velocity, base, phi, grid))
    U, interp_info = fe.pcg(lambda v: M @ v, rhs, diag, None, options)
    return dict(force=f, velocity=velocity, pressure=p, nodal_velocity=U, x=x + dt * U, info=info,
                force_mass=spread_info, velocity_mass=interp_info)
