"""P2 fused cell-block transfer on the GPU (csrc/p2.hip) against the oracle:
shared tables, fused spread and gather-assemble (single and multi-group
rules, chunked rules of more than 64 points, elements beyond the LDS patch),
the full FETransfer(fused=True) spread/interpolate with mass solves,
adjointness, conservation, affine reproduction and the reference's error
paths (tests/test_p2_cuda_execution.py:28-84)."""
import numpy as np
import pytest
import torch

from oracle import fe as ofe
from oracle import ib3d
from afsi_gpu_amd.mac import p2_transfer as p2
from afsi_gpu_amd.mac.fe import FEGeometry, FETransfer
from afsi_gpu_amd.mac.grid import MACGrid

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def t(a, gpu, dtype=torch.float64):
    return torch.tensor(np.ascontiguousarray(a), device=gpu, dtype=dtype)


def mesh(n=2, rule=4, lo=(-.4, -.3, -.5), hi=(.5, .4, .3)):
    nodes, tets = ofe.box_tets(n, lo, hi)
    return ofe.p2_solid(nodes, tets, rule)


def grids(shape=(16, 20, 24), lengths=(4., 6., 8.), origin=(-2., -3., -4.)):
    return MACGrid(shape, lengths, origin), ib3d.Grid3(shape, lengths, origin)


def on_device(g, gpu):
    return FEGeometry(t(g.cells, gpu, torch.int64), t(g.values, gpu), t(g.weights, gpu), g.node_count)


def split_rule(g, gpu, rules=(3, 5)):
    """Two groups: the first half of the elements with tet_rule(rules[0]), the
    rest with tet_rule(rules[1]) (125 points: two table chunks). Physical
    weights 6 |T| w, with |T| = the row sum of the 64-point weights."""
    half = len(g.cells) // 2
    volume = g.weights.sum(1)
    groups, oracle_groups = [], []
    for part, n in ((slice(0, half), rules[0]), (slice(half, None), rules[1])):
        pts, w = ofe.tet_rule(n)
        N = ofe.p2_basis(pts)
        W = 6 * volume[part, None] * w[None]
        oracle_groups.append(ofe.Geometry(g.cells[part], N, W, g.node_count))
        groups.append(p2.RuleGroup(t(g.cells[part], gpu, torch.int64), t(N, gpu), t(W, gpu), n))
    return p2.Rule(tuple(groups)), oracle_groups


def oracle_tables(groups, x, ogrid):
    pts = np.concatenate([ofe.p2_points(og, x) for og in groups])
    base, phi, valid = ofe.p2_shared_tables(pts, ogrid)
    return base, phi, valid, pts


def oracle_spread(groups, C, ogrid, x):
    base, phi, _, _ = oracle_tables(groups, x, ogrid)
    cb, cp = ofe.shared_expanded(base, phi)
    force = np.concatenate([ofe.weighted(og, C) for og in groups])
    return ib3d.spread_grid(force, cb, cp, ogrid)


def oracle_assemble(groups, velocity, ogrid, x):
    base, phi, _, _ = oracle_tables(groups, x, ogrid)
    cb, cp = ofe.shared_expanded(base, phi)
    gathered = ib3d.gather_grid(velocity, cb, cp, ogrid)
    out, offset = np.zeros((groups[0].node_count, 3)), 0
    for og in groups:
        count = og.weights.size
        out += ofe.assemble(og, gathered[offset:offset + count])
        offset += count
    return out


class Holder:
    """The attributes the p2 functions read from a transfer."""

    def __init__(self, grid, node_count, gpu):
        self.grid = grid
        self.diagonal = torch.zeros((node_count, 3), device=gpu, dtype=torch.float64)


def velocity_fields(grid, gpu):
    coords = [grid.coordinates(c, device=gpu) for c in range(3)]
    return tuple(torch.sin((c + 1) * coords[c].sum(-1) + .3 * c) for c in range(3))


@pytest.mark.parametrize('split', [False, True])
def test_p2_tables_spread_and_gather_match_oracle(gpu, split):
    g, X = mesh()
    grid, ogrid = grids()
    x = X + .013 * np.sin(3 * X)
    if split:
        rule, groups = split_rule(g, gpu)
    else:
        rule, groups = p2.rule_from_geometry(on_device(g, gpu)), [g]
    holder = Holder(grid, g.node_count, gpu)
    xt = t(x, gpu)
    base, phi = p2.prepare(holder, xt, rule)
    rb, rp, valid, pts = oracle_tables(groups, x, ogrid)
    assert valid.all()
    print('p2 tables', rule.point_count, 'points; max phi diff', float(np.abs(phi.cpu().numpy() - rp).max()),
          flush=True)
    np.testing.assert_array_equal(base.cpu().numpy(), rb)
    np.testing.assert_allclose(phi.cpu().numpy(), rp, rtol=0, atol=1e-15)
    stencil = p2.SharedStencil(base, phi, rule)
    C = np.sin(1.3 * X) + .1
    fields = p2.spread(holder, t(C, gpu), stencil)
    ref = oracle_spread(groups, C, ogrid, x)
    for c in range(3):
        print('p2 spread', c, rel(fields[c].cpu(), ref[c]), flush=True)
        assert rel(fields[c].cpu(), ref[c]) < 1e-12, c
    velocity = velocity_fields(grid, gpu)
    rhs = p2.assemble_velocity(holder, velocity, stencil)
    ref_rhs = oracle_assemble(groups, [u.cpu().numpy() for u in velocity], ogrid, x)
    print('p2 assemble', rel(rhs.cpu(), ref_rhs), flush=True)
    assert rel(rhs.cpu(), ref_rhs) < 1e-12
    # adjointness: sum_n C . rhs = V sum_faces u . f
    solid = float((t(C, gpu) * rhs).sum())
    fluid = float(grid.volume * sum((u * f).sum() for u, f in zip(velocity, fields)))
    assert abs(solid - fluid) <= 1e-12 * max(abs(solid), 1.)


def test_p2_fused_matches_unfused_and_handles_large_elements(gpu):
    """One element larger than the LDS patch (direct global accumulation)
    next to small ones; the fused operators equal the GPU's unfused compact
    path on the expanded tables and the oracle."""
    grid, ogrid = grids((32, 32, 32), (4., 4., 4.), (-2., -2., -2.))
    big_nodes = np.array([[-1.2, -1.1, -1.0], [1.3, -1.0, -.9], [-1.0, 1.2, -.8], [-.9, -.8, 1.3]])
    small_nodes, small_tets = ofe.box_tets(1, (.2, .2, .2), (.6, .5, .7))
    nodes = np.concatenate([big_nodes, small_nodes])
    tets = np.concatenate([np.array([[0, 1, 2, 3]]), small_tets + 4])
    g, X = ofe.p2_solid(nodes, tets)
    x = X + .02 * np.cos(2 * X)
    tr = FETransfer(grid, on_device(g, gpu), fused=True)
    st = tr.prepare(t(x, gpu))
    C = np.cos(X) - .2
    Ct = t(C, gpu)
    fields = p2.spread(tr, Ct, st)
    ref = oracle_spread([g], C, ogrid, x)
    unfused = tr.spread_grid(tr.weighted_force(Ct), st)
    for c in range(3):
        assert rel(fields[c].cpu(), ref[c]) < 1e-12, c
        assert rel(fields[c].cpu(), unfused[c].cpu()) < 1e-12, c
    velocity = velocity_fields(grid, gpu)
    rhs = p2.assemble_velocity(tr, velocity, st)
    assert rel(rhs.cpu(), oracle_assemble([g], [u.cpu().numpy() for u in velocity], ogrid, x)) < 1e-12
    assert rel(rhs.cpu(), tr.assemble_velocity(tr.gather_grid(velocity, st)).cpu()) < 1e-12


def test_fe_transfer_fused_end_to_end(gpu):
    """FETransfer(fused=True): spread/interpolate with the mass solves match
    the unfused FETransfer and the oracle; total force and power conserved;
    an affine velocity field is reproduced at the nodes
    (tests/test_p2_cuda_execution.py:46-54)."""
    g, X = mesh()
    grid, ogrid = grids()
    geometry = on_device(g, gpu)
    fused = FETransfer(grid, geometry, fused=True, warm_start=True)
    plain = FETransfer(grid, geometry)
    x = X + .013 * np.sin(3 * X)
    xt = t(x, gpu)
    sf, sp = fused.prepare(xt), plain.prepare(xt)
    assert isinstance(sf, p2.SharedStencil)
    force = np.sin(2 * X)
    ft = t(force, gpu)
    velocity = velocity_fields(grid, gpu)
    for k in range(2):
        f1, i1 = fused.spread(ft + .01 * k, sf)
        f2, _ = plain.spread(ft + .01 * k, sp)
        for c in range(3):
            assert rel(f1[c].cpu(), f2[c].cpu()) < 1e-11, c
        total = torch.stack([grid.volume * v.sum() for v in f1])
        torch.testing.assert_close(total, (ft + .01 * k).sum(0), atol=5e-10, rtol=3e-10)
        U1, _ = fused.interpolate(velocity, sf)
        U2, _ = plain.interpolate(velocity, sp)
        assert rel(U1.cpu(), U2.cpu()) < 1e-11
        power = grid.volume * sum((u * f).sum() for u, f in zip(velocity, f1))
        torch.testing.assert_close((U1 * (ft + .01 * k)).sum(), power, rtol=1e-10, atol=2e-10)
    A = xt.new_tensor([[.02, .01, 0.], [0., -.03, .01], [.01, 0., .04]])
    constant = xt.new_tensor([.3, -.2, .5])
    affine = tuple((grid.coordinates(c, device=gpu) @ A.T + constant)[..., c] for c in range(3))
    projected, _ = fused.interpolate(affine, sf)
    torch.testing.assert_close(projected, xt @ A.T + constant, atol=3e-11, rtol=3e-10)


def test_p2_error_paths_and_buffer_ownership(gpu):
    g, X = mesh()
    grid, _ = grids()
    geometry = on_device(g, gpu)
    tr = FETransfer(grid, geometry, fused=True)
    xt = t(X, gpu)
    old = tr.prepare(xt)
    saved = old.phi.clone()
    new = tr.prepare(xt + .05)
    assert old.base.data_ptr() != new.base.data_ptr()
    torch.testing.assert_close(old.phi, saved, atol=0, rtol=0)
    ones = tuple(torch.ones_like(v) for v in grid.zeros(device=gpu))
    before = p2.assemble_velocity(tr, ones, old)
    with pytest.raises(ValueError, match='support'):
        tr.prepare(xt - 2.)
    with pytest.raises(ValueError, match='invalid FE nodal field'):
        tr.prepare(xt * float('nan'))
    with pytest.raises(ValueError, match='support'):
        p2.prepare(tr, xt * float('nan'), tr.rule)
    after = p2.assemble_velocity(tr, ones, old)
    # element order of the node accumulation varies between launches
    torch.testing.assert_close(after, before, atol=2e-14, rtol=3e-13)
    # caller buffers: reused in place, validated
    base, phi = torch.empty_like(old.base), torch.empty_like(old.phi)
    b2, p2_ = p2.prepare(tr, xt, tr.rule, buffers=(base, phi))
    assert b2.data_ptr() == base.data_ptr()
    torch.testing.assert_close(p2_, saved, atol=0, rtol=0)
    with pytest.raises(ValueError, match='invalid P2 stencil buffers'):
        p2.prepare(tr, xt, tr.rule, buffers=(base[:, :-1], phi))
    with pytest.raises(ValueError, match='invalid P2 stencil buffers'):
        p2.prepare(tr, xt, tr.rule, buffers=(base, phi.float()))
    # construction errors
    p1_nodes, p1_tets = ofe.box_tets(1, (-.4, -.3, -.5), (.5, .4, .3))
    p1 = ofe.p1_solid(p1_nodes, p1_tets)
    with pytest.raises(ValueError, match='10-node'):
        FETransfer(grid, on_device(p1, gpu), fused=True)
    with pytest.raises(ValueError, match='10-node'):
        FETransfer(MACGrid((16, 20, 24), (4., 6., 8.), (-2., -3., -4.), periodic=True), geometry, fused=True)
    bad = p2.Rule((p2.RuleGroup(geometry.cells, geometry.values, geometry.weights[:, :-1]),))
    with pytest.raises(ValueError, match='weights mismatch'):
        p2.check_rule(bad, g.node_count, geometry.weights)


def test_p2_fp32_operators_match_fp64_oracle(gpu):
    """FP32 entry points (positions, basis, weights, tables and fields in
    FP32; p2_transfer.cu dispatches both types) against the FP64 oracle at
    the reference's FP32 tolerances (tests/test_p2_cuda_execution.py:38-45):
    tables, spread and gather-assemble, and the adjointness of the two."""
    g, X = mesh()
    grid, ogrid = grids()
    x = (X + .013 * np.sin(3 * X)).astype(np.float32).astype(np.float64)
    geometry = FEGeometry(t(g.cells, gpu, torch.int64), t(g.values, gpu, torch.float32),
                          t(g.weights, gpu, torch.float32), g.node_count)
    rule = p2.rule_from_geometry(geometry)
    holder = Holder(grid, g.node_count, gpu)
    holder.diagonal = holder.diagonal.float()
    base, phi = p2.prepare(holder, t(x, gpu, torch.float32), rule)
    assert phi.dtype == torch.float32
    og = ofe.Geometry(g.cells, g.values.astype(np.float32).astype(np.float64),
                      g.weights.astype(np.float32).astype(np.float64), g.node_count)
    rb, rp, valid, _ = oracle_tables([og], x, ogrid)
    np.testing.assert_array_equal(base.cpu().numpy(), rb)
    np.testing.assert_allclose(phi.cpu().numpy(), rp, rtol=8e-5, atol=3e-5)
    stencil = p2.SharedStencil(base, phi, rule)
    C = (np.sin(1.3 * X) + .1).astype(np.float32)
    fields = p2.spread(holder, t(C, gpu, torch.float32), stencil)
    ref = oracle_spread([og], C.astype(np.float64), ogrid, x)
    for c in range(3):
        assert fields[c].dtype == torch.float32
        assert rel(fields[c].double().cpu(), ref[c]) < 1e-5, c
    velocity = tuple(v.float() for v in velocity_fields(grid, gpu))
    rhs = p2.assemble_velocity(holder, velocity, stencil)
    ref_rhs = oracle_assemble([og], [u.double().cpu().numpy() for u in velocity], ogrid, x)
    assert rhs.dtype == torch.float32 and rel(rhs.double().cpu(), ref_rhs) < 1e-5
    solid = float((t(C, gpu) * rhs.double()).sum())
    fluid = float(grid.volume * sum((u.double() * f.double()).sum() for u, f in zip(velocity, fields)))
    assert abs(solid - fluid) <= 1e-5 * max(abs(solid), 1.)


def test_fe_ib_step_fused_matches_unfused(gpu):
    """FEIBStep(fused=True) runs the coupled step (mac/coupling.py:74-123)
    through the P2 fused kernels; two steps in a row agree with the unfused
    transfer's step, and the solid and fluid powers balance."""
    from afsi_gpu_amd.mac import MGOptions
    from afsi_gpu_amd.step import FEIBStep
    g, X = mesh()
    grid, _ = grids()
    opts = MGOptions(smooth=2, smoother='rbgs')
    runs = {}
    for fused in (False, True):
        stepper = FEIBStep(grid, on_device(g, gpu), dt=1e-3, device=gpu, options=opts, fused=fused, warm_start=True)
        x = t(X + .013 * np.sin(3 * X), gpu)
        F = t(np.sin(3 * X) * 10, gpu)
        u, p = grid.zeros(device=gpu), None
        for _ in range(2):
            x, u, p, U, info = stepper.step(x, F, u, p, diagnostics=True)
            assert info['power_error'] <= 1e-9 * max(abs(info['solid_power']), 1e-30)
        runs[fused] = (x, u, U)
    (xa, ua, Ua), (xb, ub, Ub) = runs[False], runs[True]
    assert rel(Ub.cpu(), Ua.cpu()) < 1e-10
    assert max(rel(b.cpu(), a.cpu()) for a, b in zip(ua, ub)) < 1e-10
    assert float((xb - xa).abs().max()) < 1e-13
