"""IBStep with the pressure kept in the solver's own buffers
(resident_pressure=True: the right-hand side formed in the solver's buffer,
the pressure returned as a view and passed back as the next warm start)
against the default path that copies both ways: three steps, FP64 and
mixed, the same cycle counts and velocities, pressures and point
velocities within 1e-12 (the spread's global atomics make two runs differ
in summation order only)."""
import numpy as np
import pytest
import torch

from afsi_gpu_amd import workloads
from afsi_gpu_amd.mac import MACGrid, MGOptions
from afsi_gpu_amd.step import IBStep

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_resident_pressure_is_bit_identical(gpu, precision):
    gen = np.random.default_rng(17)
    x0 = workloads.shell(gen, (.3, .2, .25), 30000)
    F = gen.normal(size=x0.shape)
    grid = MACGrid((32, 32, 32), (1., 1., 1.), periodic=True)
    opts = MGOptions(smooth=2, smoother='rbgs')
    runs = []
    for resident in (False, True):
        stepper = IBStep(grid, len(x0), dt=1e-3, device=gpu, options=opts, precision=precision,
                         resident_pressure=resident)
        x = torch.tensor(x0, device=gpu)
        f = torch.tensor(F, device=gpu)
        u, p = grid.zeros(device=gpu), None
        out = []
        for _ in range(3):
            u, p, U, info = stepper.step(x, f, u, p)
            out.append((U.clone(), p.clone(), tuple(c.clone() for c in u), info['pressure']['cycles']))
        runs.append(out)
    rel = lambda a, b: float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
    for (U0, p0, u0, c0), (U1, p1, u1, c1) in zip(*runs):
        assert c0 == c1
        assert rel(U1, U0) < 1e-12 and rel(p1, p0) < 1e-12
        assert all(rel(b, a) < 1e-12 for a, b in zip(u0, u1))


@pytest.mark.parametrize('periodic', [True, False])
@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_graph_steps_match_plain_launches(gpu, precision, periodic):
    """graphs=True (binning, spread, right-hand side and solver start in one
    captured graph; solver finish, correction, interpolation and advance in
    another) against the same steps as plain launches, the two steppers
    alternating so that eager work runs between the replays: five steps
    (the first from a zero pressure, the mixed solver's first outer
    iteration launched ahead from the second on), equal cycle counts,
    velocities, pressures, point velocities and positions within 1e-12."""
    gen = np.random.default_rng(23)
    x0 = workloads.shell(gen, (.3, .2, .25), 30000)
    F = 1e-2 * gen.normal(size=x0.shape)  # points move ~1/100 cell per step (none reaches a wall)
    grid = MACGrid((32, 32, 32), (1., 1., 1.), periodic=periodic)
    steppers = [IBStep(grid, len(x0), dt=1e-3, device=gpu, precision=precision, resident_pressure=True,
                       graphs=graphs) for graphs in (False, True)]
    state = [[torch.tensor(x0, device=gpu), torch.tensor(F, device=gpu), grid.zeros(device=gpu), None]
             for _ in steppers]
    rel = lambda a, b: float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
    for k in range(5):
        out = []
        for stepper, st in zip(steppers, state):
            x, f, u, p = st
            u, p, U, info = stepper.step(x, f, u, p)
            st[2], st[3] = u, p
            out.append((U.clone(), p.clone(), tuple(c.clone() for c in u), x.clone(), info['pressure']['cycles']))
        (U0, p0, u0, x0_, c0), (U1, p1, u1, x1, c1) = out
        assert c0 == c1, k
        assert rel(U1, U0) < 1e-12 and rel(p1, p0) < 1e-12, k
        assert all(rel(b, a) < 1e-12 for a, b in zip(u0, u1)), k
        assert rel(x1, x0_) < 1e-14, k
    assert len(steppers[1].graphs) >= 2
