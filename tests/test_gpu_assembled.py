"""CSR-assembled P1 operator on the GPU (csrc/assembled.hip) against the
oracle: the device hash table (collisions, duplicates, zero skipping,
overflow, large keys with cancelling repeats; tests/test_csr_hash_assembly.py:
22-60), the contracted operator B_c against the oracle's dense B_c, and the
AssembledTransfer spread/interpolate against FETransfer and the oracle."""
import numpy as np
import pytest
import torch

from oracle import assembled as oas
from oracle import fe as ofe
from oracle import ib3d
from afsi_gpu_amd.mac.assembled import AssembledTransfer, HashWorkspace, accumulate, sorted_entries
from afsi_gpu_amd.mac.fe import FEGeometry, FETransfer
from afsi_gpu_amd.mac.grid import MACGrid

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def t(a, gpu, dtype=torch.float64):
    return torch.tensor(np.ascontiguousarray(a), device=gpu, dtype=dtype)


def test_hash_table_collisions_overflow_and_large_keys(gpu):
    colliders = [k for k in range(10000) if oas.bucket(k, 128) == 3][:40]
    keys = t(colliders * 7 + [987654], gpu, torch.int64)
    values = torch.full(keys.shape, .125, device=gpu, dtype=torch.float64)
    values[-1] = 0
    work = HashWorkspace()
    work.reset(128, values)
    accumulate(keys, values, work)
    assert work.flag.item() == 0
    k, v = sorted_entries(work)
    assert k.tolist() == sorted(colliders)
    assert torch.equal(v, torch.full((40,), .875, device=gpu, dtype=torch.float64))
    # every key landed in its SplitMix64 bucket or within the probe run after it
    tk = work.keys.cpu().numpy()
    for key in colliders:
        slot = int(np.nonzero(tk == key)[0][0])
        assert (slot - oas.bucket(key, 128)) % 128 < 40
    work.reset(8, values)
    accumulate(torch.arange(20, device=gpu), values[:20].clone().fill_(1.), work)
    assert work.flag.item() == 1
    unique = torch.arange(1024, device=gpu, dtype=torch.int64) * 2147483649 + 4294967296
    repeats = 128
    keys = unique.repeat(repeats)
    signs = torch.where(torch.arange(repeats, device=gpu) % 2 == 0, 1., -.5).to(torch.float64)
    values = signs.repeat_interleave(len(unique)) * .125
    for _ in range(3):
        work.reset(4096, values)
        accumulate(keys, values, work)
        assert work.flag.item() == 0
        k, v = sorted_entries(work)
        assert torch.equal(k, unique)
        assert torch.equal(v, torch.full_like(v, repeats / 32))


def setup(gpu, which='solid'):
    if which == 'solid':
        nodes, tets = ofe.box_tets(3, (-.4, -.3, -.5), (.5, .4, .3))
        g = ofe.p1_solid(nodes, tets)
    else:
        nodes, tris = ofe.icosphere(2, .6, (.1, -.2, .05))
        g = ofe.p1_surface(nodes, tris)
    geometry = FEGeometry(t(g.cells, gpu, torch.int64), t(g.values, gpu), t(g.weights, gpu), g.node_count)
    grid = MACGrid((16, 20, 24), (4., 6., 8.), (-2., -3., -4.))
    ogrid = ib3d.Grid3((16, 20, 24), (4., 6., 8.), (-2., -3., -4.))
    return g, nodes, geometry, grid, ogrid


@pytest.mark.parametrize('which', ['solid', 'sphere'])
def test_assembled_operator_matches_dense_oracle(gpu, which):
    g, X, geometry, grid, ogrid = setup(gpu, which)
    x = X + .013 * np.sin(3 * X)
    tr = AssembledTransfer(grid, geometry)
    st = tr.prepare(t(x, gpu))
    pts = ofe.evaluate(g, x)
    for c in range(3):
        B = oas.dense_operator(g, pts, ogrid, c)
        m = st.gather[c]
        dense = torch.sparse_csr_tensor(m.row_ptr.long(), m.cols.long(), m.vals, (g.node_count, B.shape[1]))
        got = dense.to_dense().cpu().numpy()
        print('assembled B', which, c, int(m.vals.numel()), rel(got, B), flush=True)
        assert rel(got, B) < 1e-13
        assert int(m.vals.numel()) == int((B != 0).sum())
        mt = st.spread[c]
        dt = torch.sparse_csr_tensor(mt.row_ptr.long(), mt.cols.long(), mt.vals, (B.shape[1], g.node_count))
        assert torch.equal(dt.to_dense(), dense.to_dense().T)


def test_assembled_transfer_matches_fe_transfer_and_oracle(gpu):
    g, X, geometry, grid, ogrid = setup(gpu)
    x = X + .013 * np.sin(3 * X)
    xt = t(x, gpu)
    asm = AssembledTransfer(grid, geometry, warm_start=True)
    plain = FETransfer(grid, geometry, warm_start=True)
    sa, sp = asm.prepare(xt), plain.prepare(xt)
    coords = [grid.coordinates(c, device=gpu) for c in range(3)]
    velocity = tuple(torch.sin((c + 1) * coords[c].sum(-1)) for c in range(3))
    F = t(np.sin(3 * X), gpu)
    fa, ia = asm.spread(F, sa)
    fp, _ = plain.spread(F, sp)
    M = ofe.mass_matrix(g)
    C, _ = ofe.pcg(lambda v: M @ v, np.sin(3 * X), ofe.mass_diagonal(g))
    base, phi = ib3d.prepare_compact(ofe.evaluate(g, x), ogrid)
    ref = ib3d.spread_grid(ofe.weighted(g, C), base, phi, ogrid)
    for c in range(3):
        assert rel(fa[c].cpu(), fp[c].cpu()) < 1e-11
        assert rel(fa[c].cpu(), ref[c]) < 1e-10
    Ua, _ = asm.interpolate(velocity, sa)
    Up, _ = plain.interpolate(velocity, sp)
    assert rel(Ua.cpu(), Up.cpu()) < 1e-11
    total = torch.stack([grid.volume * v.sum() for v in fa])
    torch.testing.assert_close(total, F.sum(0), rtol=1e-10, atol=2e-10)
    assert asm.csr_builds == 1
    with pytest.raises(RuntimeError, match='budget exceeded'):
        AssembledTransfer(grid, geometry, max_entries=100).prepare(xt)
    with pytest.raises(ValueError, match='support'):
        asm.prepare(xt + 10)


def test_dropin_contract_and_hash_accumulate(gpu):
    """The reference-API ops (mac/cuda_ib.py:93-112) called the way
    mac/hash_transfer_assembly.py:80-130 calls them: per cell group with its
    point offset and stencil boxes, into one hash table per component; the
    table holds the oracle's dense B_c. The cells are split into two groups
    so the second runs at a nonzero offset."""
    from types import SimpleNamespace
    from afsi_gpu_amd.mac import cuda_ib
    from afsi_gpu_amd.mac.transfer import from_points
    g, X, geometry, grid, ogrid = setup(gpu)
    x = X + .013 * np.sin(3 * X)
    pts = ofe.evaluate(g, x)
    stencil = from_points(t(pts, gpu), grid)
    cells, weights = t(g.cells, gpu, torch.int64), t(g.weights, gpu)
    E, Q = weights.shape
    half = E // 2
    groups = [SimpleNamespace(values=t(g.values, gpu), weights=weights[:half].contiguous(), cells=cells[:half]),
              SimpleNamespace(values=t(g.values, gpu), weights=weights[half:].contiguous(), cells=cells[half:])]
    for c in range(3):
        shape = grid.face_shape(c)
        work = HashWorkspace()
        work.reset(1 << 16, weights)
        offset = 0
        for group in groups:
            e = group.weights.shape[0]
            base = stencil.base[c, offset:offset + e * Q].view(e, Q, 3)
            low = base.amin(1)
            width = base.amax(1) - low + 4
            sizes = width.prod(-1)
            prefix = torch.cat((sizes.new_zeros(1), sizes.cumsum(0)))
            cuda_ib.contract(stencil, group, c, offset, low, width, prefix, int(prefix[-1]), shape, work)
            offset += group.weights.numel()
        assert work.flag.item() == 0
        keys, vals = sorted_entries(work)
        nf = int(np.prod(shape))
        B = oas.dense_operator(g, pts, ogrid, c)
        got = np.zeros_like(B)
        k = keys.cpu().numpy()
        got[k // nf, k % nf] = vals.cpu().numpy()
        assert rel(got, B) < 1e-13
        # hash_accumulate of the same entries doubles every value
        cuda_ib.hash_accumulate(keys, vals, work)
        k2, v2 = sorted_entries(work)
        assert torch.equal(k2, keys) and torch.equal(v2, 2 * vals)
    with pytest.raises(NotImplementedError):
        cuda_ib.contract(stencil, groups[0], 0, 0, low, width, prefix, 0, grid.face_shape(0), work, cache=object())
