"""GPU explicit MAC predictor and MACFlow.step against the oracle
(mac/flow.py:59-64,122-157; mac/grid.py:96-134): the fused predictor kernel
bit-identical to the FP64 restatement, the flow step within 1e-10, the
reference's gradient-body-force test (tests/test_mac.py:103-113), and the
transport and viscous guards."""
from math import pi

import numpy as np
import pytest
import torch

from oracle import mac3d
from afsi_gpu_amd import _abi
from afsi_gpu_amd.mac import MACGrid, MGOptions
from afsi_gpu_amd.mac.flow import MACFlow, MACTransportGuardError

pytestmark = pytest.mark.gpu


def random_fields(grid, seed, scale=1.):
    gen = np.random.default_rng(seed)
    return tuple(scale * gen.normal(size=grid.face_shape(c)) for c in range(3))


@pytest.mark.parametrize('shape,lengths', [((12, 10, 14), (1.2, 1., 1.4)), ((8, 8, 8), (1., 1., 1.)),
                                           ((5, 7, 4), (.5, .7, .4))])
def test_predictor_bit_exact(gpu, shape, lengths):
    grid = MACGrid(shape, lengths)
    h = grid.spacing
    u = random_fields(grid, 1)
    f = random_fields(grid, 2, 3.)
    dt, rho, mu = 1e-3, 1.3, .7
    ref = mac3d.predict(u, f, h, dt=dt, rho=rho, mu=mu)
    ut = [torch.tensor(a, device=gpu) for a in u]
    ft = [torch.tensor(a, device=gpu) for a in f]
    st = [torch.empty_like(a) for a in ut]
    _abi.call('afsi_mac_predict_f64', *(_abi.ptr(a) for a in ut), *(_abi.ptr(a) for a in ft),
              *(_abi.ptr(a) for a in st), _abi.i64x3(shape), _abi.f64x3(h), dt, rho, mu, _abi.stream())
    for c in range(3):
        got = st[c].cpu().numpy()
        assert np.array_equal(got, ref[c]), (c, np.abs(got - ref[c]).max())


def test_flow_step_matches_oracle_and_guards(gpu):
    shape, lengths = (16, 12, 20), (1.6, 1.2, 2.)
    grid = MACGrid(shape, lengths)
    h = grid.spacing
    dt = 5e-4   # viscous number dt * sum 1/h^2 = 0.15 (limit 0.25)
    u = mac3d.zero_normal(random_fields(grid, 3, .5))
    f = random_fields(grid, 4)
    opts = MGOptions(smooth=2, smoother='rbgs')
    mg = mac3d.Multigrid(shape, h, mac3d.MGOptions(smooth=2, smoother='rbgs'))
    v_ref, p_ref, info_ref = mac3d.flow_step(u, f, h, mg, dt=dt)
    flow = MACFlow(grid, dt=dt, device=gpu, options=opts)
    res = flow.step(tuple(torch.tensor(a, device=gpu) for a in u), tuple(torch.tensor(a, device=gpu) for a in f))
    for a, b in zip(res.velocity, v_ref):
        assert np.linalg.norm(a.cpu().numpy() - b) / np.linalg.norm(b) < 1e-10
    for key in ('courant', 'cell_reynolds', 'advection_diffusion_number'):
        assert abs(res.diagnostics[key] / info_ref[key] - 1) < 1e-14
    # transport guard: too fast a flow stops the step before its predictor
    fast = tuple(torch.tensor(a * 1e4, device=gpu) for a in u)
    with pytest.raises(MACTransportGuardError, match='courant') as err:
        flow.step(fast, tuple(torch.tensor(a, device=gpu) for a in f))
    assert 'courant' in err.value.diagnostics['triggered']
    bad = [torch.tensor(a, device=gpu) for a in f]
    bad[1][0, 0, 0] = float('nan')
    with pytest.raises(ValueError, match='finite'):
        flow.step(tuple(torch.tensor(a, device=gpu) for a in u), tuple(bad))
    with pytest.raises(ValueError, match='viscous'):
        MACFlow(grid, dt=1., device=gpu)


def test_gradient_body_force_has_zero_projected_velocity(gpu):
    """tests/test_mac.py:103-113."""
    grid = MACGrid((16, 16, 16), (2., 2., 2.))
    flow = MACFlow(grid, dt=1e-4, device=gpu)
    xyz = grid.coordinates(device=gpu)
    p = torch.cos(pi * xyz[..., 0] / 2) * torch.cos(pi * xyz[..., 2] / 2)
    grad = mac3d.gradient(p.cpu().numpy(), grid.spacing)
    result = flow.step(grid.zeros(device=gpu), tuple(torch.tensor(g, device=gpu) for g in grad))
    torch.testing.assert_close(result.pressure - result.pressure.mean(), p - p.mean(), rtol=1e-8, atol=1e-9)
    assert max(u.abs().max().item() for u in result.velocity) < 1e-11
