"""The torch-CPU restatement used as the CPU baseline (oracle/torch_cpu.py)
against the NumPy oracle on the same inputs: compact stencils, spread and
gather within 1e-13, one Jacobi V(4,4) cycle within 1e-12, and a solve that
meets the reference's tolerance."""
import numpy as np
import torch

from oracle import ib3d, mac3d, torch_cpu


def setup(n=(16, 12, 20), points=700, seed=3):
    gen = np.random.default_rng(seed)
    lengths = (1., .75, 1.25)
    grid = ib3d.Grid3(n, lengths, periodic=True)
    x = gen.uniform(0, 1, size=(points, 3)) * np.asarray(lengths)
    x[:200] = x[0] + 1e-3 * gen.normal(size=(200, 3))
    x %= np.asarray(lengths)
    F = gen.normal(size=(points, 3))
    return grid, x, F


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


def test_transfers_match_numpy_oracle():
    grid, x, F = setup()
    base, phi = ib3d.prepare_compact(x, grid)
    tb, tphi = torch_cpu.prepare(torch.tensor(x), grid.shape, grid.spacing)
    assert np.array_equal(tb.numpy(), base)
    assert np.abs(tphi.numpy() - phi).max() < 1e-15
    f = torch_cpu.spread(torch.tensor(F), tb, tphi, grid.shape, grid.volume)
    f_ref = ib3d.spread_grid(F, base, phi, grid)
    assert max(rel(f[c], f_ref[c]) for c in range(3)) < 1e-13
    u = tuple(np.random.default_rng(4).normal(size=grid.shape) for _ in range(3))
    U = torch_cpu.gather(tuple(torch.tensor(a) for a in u), tb, tphi, grid.shape)
    assert rel(U, ib3d.gather_grid(u, base, phi, grid)) < 1e-13


def test_jacobi_cycle_and_solve_match_numpy_oracle():
    shape, h = (16, 16, 32), (1 / 16, 1 / 16, 2 / 32)
    gen = np.random.default_rng(5)
    rhs = gen.normal(size=shape)
    rhs -= rhs.mean()
    start = gen.normal(size=shape)
    ref = mac3d.Multigrid(shape, h, mac3d.MGOptions(smooth=4, smoother='jacobi'), periodic=True)
    mg = torch_cpu.JacobiMultigrid(shape, h, smooth=4)
    assert mg.shapes == ref.shapes
    got = mg.cycle(0, torch.tensor(start - start.mean()), torch.tensor(rhs)).numpy()
    want = ref.cycle(0, start - start.mean(), rhs)
    assert rel(got, want) < 1e-12
    p, cycles = mg.solve(torch.tensor(rhs))
    r = rhs - mac3d.negative_laplacian(p.numpy(), h, True)
    assert np.linalg.norm(r) <= 1e-10 * np.linalg.norm(rhs)
    _, info = ref.solve(rhs)
    assert cycles == info['cycles']


def test_step_matches_numpy_oracle():
    from oracle.step import ib_step
    grid, x, F = setup((16, 16, 16), 600, 7)
    grid = ib3d.Grid3((16, 16, 16), (1., 1., 1.), periodic=True)
    x %= 1.
    u = tuple(np.zeros(grid.shape) for _ in range(3))
    ref = ib_step(x, F, u, None, grid,
                  mac3d.Multigrid(grid.shape, grid.spacing, mac3d.MGOptions(smooth=4, smoother='jacobi'), True),
                  dt=1e-3)
    mg = torch_cpu.JacobiMultigrid(grid.shape, grid.spacing, smooth=4)
    x1, vel, p, U, cycles = torch_cpu.ib_step(torch.tensor(x), torch.tensor(F),
                                              tuple(torch.tensor(a) for a in u), None, grid.shape,
                                              grid.spacing, mg, dt=1e-3)
    assert cycles == ref['info']['cycles']
    assert rel(U, ref['point_velocity']) < 1e-11
    assert max(rel(vel[c], ref['velocity'][c]) for c in range(3)) < 1e-11
