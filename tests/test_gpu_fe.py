"""GPU parity of the FE interaction-point transfer and the mass PCG against
oracle/fe.py, and the conservation checks of the reference's own test
(tests/test_mac_execution.py:26-66): total force, power, torque, uniform
velocity, support and invalid-field errors, PCG restart/breakdown paths.

Bars: B, W B, B^T W and the mass action within 1e-13 relative; the PCG
stops at an iteration the oracle reaches on the mass matrix moved by a few
units in the last place (the oracle's own count wherever that is unique)
and agrees within 1e-10; transferred
fields within 1e-10 (the PCG tolerance is 1e-12 relative)."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

from oracle import fe as ofe
from oracle import ib3d
from afsi_gpu_amd.mac.fe import FEGeometry, FETransfer, MassSolver, SolverOptions
from afsi_gpu_amd.mac.grid import MACGrid

pytestmark = pytest.mark.gpu


def mesh(which):
    if which == 'sphere':
        nodes, tris = ofe.icosphere(2, .3, (.1, -.2, .05))
        return ofe.p1_surface(nodes, tris), nodes
    if which == 'p2':   # the reference's production element: 10 nodes, 64 points per cell
        nodes, tets = ofe.box_tets(2, (-.4, -.3, -.5), (.5, .4, .3))
        return ofe.p2_solid(nodes, tets)
    nodes, tets = ofe.box_tets(3, (-.4, -.3, -.5), (.5, .4, .3))
    return ofe.p1_solid(nodes, tets), nodes


def on_device(g, gpu):
    t = lambda a, dt=torch.float64: torch.tensor(a, device=gpu, dtype=dt)
    return FEGeometry(t(g.cells, torch.int64), t(g.values), t(g.weights), g.node_count)


def grids():
    return (MACGrid((16, 20, 24), (4., 6., 8.), (-2., -3., -4.)),
            ib3d.Grid3((16, 20, 24), (4., 6., 8.), (-2., -3., -4.)))


def reachable_stops(M, rhs, diagonal, initial, options, count=16, seed=0):
    """Iteration counts at which the oracle PCG stops on M and on ``count``
    symmetric copies of M whose entries are moved by up to two units in the
    last place. Where the recurrence norm crosses the tolerance at a
    rounding-decided iteration (long unrestarted runs lose orthogonality:
    on the P2 box mesh with no restart the stop moves by one iteration under
    such perturbations at every rtol from 1e-8 to 1e-12) the set holds more
    than one count; with restarts every few iterations it holds one, and the
    GPU must take exactly that count."""
    M = sp.csr_matrix(M)
    rng = np.random.default_rng(seed)
    ulps = lambda a: rng.integers(-2, 3, size=a.shape) * np.spacing(np.abs(a))
    stops = set()
    for k in range(count + 1):
        P = M
        if k:
            upper = sp.triu(M, 1).tocsr()
            upper.data = upper.data + ulps(upper.data)
            P = (upper + upper.T + sp.diags(M.diagonal() + ulps(M.diagonal()))).tocsr()
        stops.add(ofe.pcg(lambda v: P @ v, rhs, diagonal, initial, options)[1].iterations)
    return stops


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize('which', ['sphere', 'solid', 'p2'])
def test_fe_operators_match_oracle(gpu, which):
    g, X = mesh(which)
    grid, _ = grids()
    tr = FETransfer(grid, on_device(g, gpu))
    rng = np.random.default_rng(1)
    v = rng.normal(size=(g.node_count, 3))
    U = rng.normal(size=(g.cells.shape[0] * g.values.shape[0], 3))
    vt, Ut = torch.tensor(v, device=gpu), torch.tensor(U, device=gpu)
    assert rel(tr.evaluate(vt).cpu(), ofe.evaluate(g, v)) < 1e-13
    assert rel(tr.weighted_force(vt).cpu(), ofe.weighted(g, v)) < 1e-13
    assert rel(tr.assemble_velocity(Ut).cpu(), ofe.assemble(g, U)) < 1e-13
    assert rel(tr.mass_action(vt).cpu(), ofe.mass_matrix(g) @ v) < 1e-13
    assert rel(tr.diagonal.cpu(), ofe.mass_diagonal(g)) < 1e-14
    assert rel(tr.interaction_points(torch.tensor(X, device=gpu)).cpu(), ofe.interaction_points(g, X)) < 1e-14


def test_mass_pcg_matches_oracle_and_error_paths(gpu):
    g, X = mesh('solid')
    grid, _ = grids()
    tr = FETransfer(grid, on_device(g, gpu))
    opts = SolverOptions(rtol=1e-11, atol=1e-13, recompute_every=12, check_every=4, max_iterations=500)
    oopts = ofe.SolverOptions(rtol=1e-11, atol=1e-13, recompute_every=12, check_every=4, max_iterations=500)
    solver = MassSolver(tr.mass, tr.diagonal, opts)
    rhs, initial = np.sin(5 * X), .1 * np.cos(X)
    M = ofe.mass_matrix(g)
    expected, einfo = ofe.pcg(lambda v: M @ v, rhs, ofe.mass_diagonal(g), initial, oopts)
    rt, it = torch.tensor(rhs, device=gpu), torch.tensor(initial, device=gpu)
    result, info = solver.solve(rt, it)
    print('PCG iterations', info.iterations, 'oracle', einfo.iterations)
    assert info.iterations in reachable_stops(M, rhs, ofe.mass_diagonal(g), initial, oopts)
    assert rel(result.cpu(), expected) < 1e-10
    assert np.linalg.norm(rhs - M @ result.cpu().numpy()) <= info.tolerance * (1 + 1e-9)
    assert abs(info.tolerance / einfo.tolerance - 1) < 1e-12
    saved = result.clone()
    ptrs = [solver.x.data_ptr(), solver.workspace.data_ptr()]
    again, _ = solver.solve(-rt, result)
    assert rel(again.cpu(), -expected) < 1e-10
    assert torch.equal(result, saved)
    assert ptrs == [solver.x.data_ptr(), solver.workspace.data_ptr()]
    zero, zinfo = solver.solve(torch.zeros_like(rt))
    assert zinfo.iterations == 0 and zero.count_nonzero() == 0
    with pytest.raises(ValueError, match='finite'):
        solver.solve(torch.full_like(rt, float('nan')))
    from afsi_gpu_amd.mac.fe import CSRMatrix
    neg = CSRMatrix(tr.mass.row_ptr, tr.mass.cols, (-tr.mass.vals).contiguous(), tr.mass.rows)
    with pytest.raises(RuntimeError, match='breakdown'):
        MassSolver(neg, tr.diagonal, opts).solve(rt)
    with pytest.raises(RuntimeError, match='converge'):
        MassSolver(tr.mass, tr.diagonal, SolverOptions(max_iterations=1, rtol=1e-14)).solve(rt)


def oracle_spread(g, ogrid, x, force):
    M = ofe.mass_matrix(g)
    C, _ = ofe.pcg(lambda v: M @ v, force, ofe.mass_diagonal(g))
    points = ofe.interaction_points(g, x)
    base, phi = ib3d.prepare_compact(points, ogrid)
    return ib3d.spread_grid(ofe.weighted(g, C), base, phi, ogrid)


def oracle_interpolate(g, ogrid, x, velocity):
    M = ofe.mass_matrix(g)
    points = ofe.interaction_points(g, x)
    base, phi = ib3d.prepare_compact(points, ogrid)
    rhs = ofe.assemble(g, ib3d.gather_grid(velocity, base, phi, ogrid))
    U, _ = ofe.pcg(lambda v: M @ v, rhs, ofe.mass_diagonal(g))
    return U


@pytest.mark.parametrize('which', ['sphere', 'solid', 'p2'])
def test_fe_transfer_matches_oracle_and_conserves(gpu, which):
    g, X = mesh(which)
    grid, ogrid = grids()
    tr = FETransfer(grid, on_device(g, gpu), warm_start=True)
    x = X + .013 * np.sin(3 * X)
    xt = torch.tensor(x, device=gpu)
    st = tr.prepare(xt)
    coords = [grid.coordinates(c, device=gpu) for c in range(3)]
    velocity = tuple(torch.sin((c + 1) * coords[c].sum(-1)) for c in range(3))
    for k in range(2):
        force = np.sin(3 * X) + .01 * k
        ft = torch.tensor(force, device=gpu)
        f, fi = tr.spread(ft, st)
        U, ui = tr.interpolate(velocity, st)
        assert fi.residual_norm <= fi.tolerance and ui.residual_norm <= ui.tolerance
        ref_f = oracle_spread(g, ogrid, x, force)
        for c in range(3):
            assert rel(f[c].cpu(), ref_f[c]) < 1e-10, c
        ref_U = oracle_interpolate(g, ogrid, x, [u.cpu().numpy() for u in velocity])
        assert rel(U.cpu(), ref_U) < 1e-10
        total = torch.stack([fc.sum() * grid.volume for fc in f])
        torch.testing.assert_close(total, ft.sum(0), rtol=1e-10, atol=2e-10)
        power = grid.volume * sum((u * fc).sum() for u, fc in zip(velocity, f))
        torch.testing.assert_close((U * ft).sum(), power, rtol=1e-10, atol=2e-10)
        torque = torch.zeros(3, device=gpu, dtype=torch.float64)
        for c, fc in enumerate(f):
            vec = torch.zeros((*fc.shape, 3), device=gpu, dtype=torch.float64)
            vec[..., c] = fc
            torque += torch.linalg.cross(coords[c], vec).sum((0, 1, 2)) * grid.volume
        torch.testing.assert_close(torque, torch.linalg.cross(xt, ft).sum(0), rtol=1e-10, atol=2e-10)
    uniform = tuple(torch.full_like(v, float(c + 1)) for c, v in enumerate(velocity))
    u, _ = tr.interpolate(uniform, st)
    torch.testing.assert_close(u, xt.new_tensor([1., 2., 3.]).expand_as(xt), rtol=1e-10, atol=1e-10)
    with pytest.raises(ValueError, match='support'):
        tr.prepare(xt + 10)
    invalid = list(velocity)
    invalid[0] = torch.full_like(invalid[0], float('nan'))
    with pytest.raises(ValueError, match='invalid'):
        tr.interpolate(tuple(invalid), st)
    with pytest.raises(ValueError, match='invalid FE nodal field'):
        tr.spread(ft[:-1], st)


def test_fe_ib_step_matches_oracle(gpu):
    """One coupled step (mac/coupling.py:74-123 ordering) on a closed box."""
    from afsi_gpu_amd.mac import MGOptions
    from afsi_gpu_amd.step import FEIBStep
    from oracle import mac3d
    from oracle.step import fe_ib_step
    g, X = mesh('sphere')
    grid, ogrid = grids()
    dt = 1e-3
    stepper = FEIBStep(grid, on_device(g, gpu), dt=dt, device=gpu, options=MGOptions(smooth=2, smoother='rbgs'))
    x = X + .013 * np.sin(3 * X)
    F = np.sin(3 * X) * 10
    u = grid.zeros(device=gpu)
    x_new, vel, p, U, info = stepper.step(torch.tensor(x, device=gpu), torch.tensor(F, device=gpu), u,
                                          diagnostics=True)
    mg = mac3d.Multigrid(ogrid.shape, ogrid.spacing, mac3d.MGOptions(smooth=2, smoother='rbgs'), periodic=False)
    ref = fe_ib_step(g, x, F, tuple(np.zeros(ogrid.face_shape(c)) for c in range(3)), None, ogrid, mg, dt=dt)
    err_U = rel(U.cpu(), ref['nodal_velocity'])
    err_u = max(rel(vel[c].cpu(), ref['velocity'][c]) for c in range(3))
    print('FE step: nodal velocity', err_U, 'face velocity', err_u, 'mass PCG iterations',
          info['force_mass'].iterations, info['velocity_mass'].iterations, 'power error', info['power_error'])
    assert err_U < 1e-9 and err_u < 1e-9
    assert np.abs(x_new.cpu().numpy() - ref['x']).max() < 1e-14 + 1e-9 * np.abs(ref['x']).max()
    assert info['power_error'] <= 1e-9 * max(abs(info['solid_power']), 1e-30)
    with pytest.raises(ValueError, match='displacement'):
        FEIBStep(grid, on_device(g, gpu), dt=50., device=gpu).step(
            torch.tensor(x, device=gpu), torch.tensor(F, device=gpu), grid.zeros(device=gpu))


def test_fe_ib_step_with_full_flow_step_matches_oracle(gpu):
    """The coupled step with MACFlow.step (transport screen, explicit
    convection and viscosity, projection) two steps in a row, as
    MACIBStepper.step runs it (mac/coupling.py:74-123)."""
    from afsi_gpu_amd.mac import MGOptions
    from afsi_gpu_amd.step import FEIBStep
    from oracle import mac3d
    from oracle.step import fe_ib_step
    g, X = mesh('p2')
    grid, ogrid = grids()
    dt = 2e-3   # viscous number 2e-3 * 3 * 16 = 0.096
    stepper = FEIBStep(grid, on_device(g, gpu), dt=dt, device=gpu, options=MGOptions(smooth=2, smoother='rbgs'),
                       predictor='full')
    mg = mac3d.Multigrid(ogrid.shape, ogrid.spacing, mac3d.MGOptions(smooth=2, smoother='rbgs'), periodic=False)
    x = X + .013 * np.sin(3 * X)
    F = np.sin(3 * X) * 20
    u = grid.zeros(device=gpu)
    u_ref = tuple(np.zeros(ogrid.face_shape(c)) for c in range(3))
    xt, p, p_ref = torch.tensor(x, device=gpu), None, None
    for k in range(2):
        xt, u, p, U, info = stepper.step(xt, torch.tensor(F, device=gpu), u, p)
        ref = fe_ib_step(g, x, F, u_ref, p_ref, ogrid, mg, dt=dt, full=True)
        x, u_ref, p_ref = ref['x'], ref['velocity'], ref['pressure']
        err_U = rel(U.cpu(), ref['nodal_velocity'])
        err_u = max(rel(u[c].cpu(), u_ref[c]) for c in range(3))
        print('full FE step', k, err_U, err_u, info['pressure']['courant'])
        assert err_U < 1e-9 and err_u < 1e-9


@pytest.mark.parametrize('check,recompute', [(1, 40), (3, 2), (4, 5), (4, 12), (5, 3), (7, 7)])
def test_mass_pcg_schedules_match_oracle(gpu, check, recompute):
    """The PCG reads each check one check period late and restarts and its
    last iteration synchronously; across check and restart schedules
    (restarts before, at and between checks) the iteration count, the status
    and the result follow the oracle's recurrence (fluid/solvers.py:39-122);
    the count is one the oracle reaches under last-place perturbations."""
    g, X = mesh('p2')
    grid, _ = grids()
    tr = FETransfer(grid, on_device(g, gpu))
    M = ofe.mass_matrix(g)
    rhs = np.sin(5 * X) + .3 * np.cos(2 * X[:, ::-1])
    for rtol, limit in ((1e-12, 500), (1e-14, 9)):
        opts = SolverOptions(rtol=rtol, atol=1e-16, recompute_every=recompute, check_every=check,
                             max_iterations=limit)
        oopts = ofe.SolverOptions(rtol=rtol, atol=1e-16, recompute_every=recompute, check_every=check,
                                  max_iterations=limit)
        solver = MassSolver(tr.mass, tr.diagonal, opts)
        try:
            expected, einfo = ofe.pcg(lambda v: M @ v, rhs, ofe.mass_diagonal(g), None, oopts)
        except ofe.PCGError as err:
            with pytest.raises(RuntimeError, match='converge'):
                solver.solve(torch.tensor(rhs, device=gpu))
            print('schedule', check, recompute, limit, 'both fail:', str(err)[:40], flush=True)
            continue
        result, info = solver.solve(torch.tensor(rhs, device=gpu))
        print('schedule', check, recompute, limit, info.iterations, einfo.iterations, flush=True)
        assert info.iterations in reachable_stops(M, rhs, ofe.mass_diagonal(g), None, oopts)
        assert rel(result.cpu(), expected) < 1e-10
