"""The spread entry point that allocates nothing
(afsi_spread_tiled_noalloc_f64) against afsi_spread_tiled_f64 through the C
ABI: two calls in a row on one GPU, the second with fewer points on a
smaller closed grid, each within 1e-12 relative of the allocating entry
point and of the oracle; the 128^3 clustered-shells workload within 1e-12
of the allocating spread; zero points; and a grid of too many tiles
rejected with an error before any launch."""
import numpy as np
import pytest
import torch

from oracle import ib3d
from afsi_gpu_amd import _abi, workloads
from afsi_gpu_amd.mac import MACGrid
from afsi_gpu_amd.mac.transfer import Binner, TiledTransfer

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


def noalloc(x, F, grid, tile, bins, out, flags):
    _abi.call('afsi_spread_tiled_noalloc_f64', _abi.ptr(x), _abi.ptr(F), x.shape[0], grid.abi(),
              _abi.i32x3(tile), _abi.ptr(bins.perm), _abi.ptr(bins.cell_ptr), _abi.ptr(out[0]),
              _abi.ptr(out[1]), _abi.ptr(out[2]), _abi.ptr(flags), _abi.stream())


def case(gpu, shape, periodic, count, seed, tile=(8, 8, 8)):
    gen = np.random.default_rng(seed)
    lengths = (1., 1.25, .75)
    grid = MACGrid(shape, lengths, periodic=periodic)
    og = ib3d.Grid3(shape, lengths, periodic=periodic)
    h = np.asarray(og.spacing)
    x = 2.5 * h + gen.uniform(0, 1, size=(count, 3)) * (np.asarray(lengths) - 5 * h)
    x[: count // 4] = x[0] + 1e-3 * gen.normal(size=(count // 4, 3))  # a dense tile
    F = gen.normal(size=(count, 3))
    xt, Ft = torch.tensor(x, device=gpu), torch.tensor(F, device=gpu)
    binner = Binner(grid, count, gpu, tile)
    bins = binner(xt)
    ref = TiledTransfer(grid, gpu, tile, flags=binner.flags).spread(xt, Ft, bins)
    out = tuple(torch.full(grid.face_shape(c), float('nan'), device=gpu, dtype=torch.float64) for c in range(3))
    noalloc(xt, Ft, grid, tile, bins, out, binner.flags)
    torch.cuda.synchronize()
    assert int(binner.flags.item()) == 0
    base, phi = ib3d.prepare_compact(x, og)
    oracle = ib3d.spread_grid(F, base, phi, og)
    for c in range(3):
        assert bool(torch.isfinite(out[c]).all())
        assert rel(out[c].cpu(), ref[c].cpu()) < 1e-12
        assert rel(out[c].cpu(), oracle[c]) < 1e-12


def test_noalloc_spread_twice_then_smaller(gpu):
    case(gpu, (24, 20, 28), True, 30000, 1)
    case(gpu, (16, 12, 12), False, 4000, 2, tile=(5, 3, 7))


def test_noalloc_spread_clustered_shells_128_and_empty(gpu):
    x_np, F_np = workloads.clustered_shells_128()
    grid = MACGrid((128,) * 3, (1., 1., 1.), periodic=True)
    x, F = torch.tensor(x_np, device=gpu), torch.tensor(F_np, device=gpu)
    binner = Binner(grid, len(x_np), gpu)
    bins = binner(x)
    ref = TiledTransfer(grid, gpu, flags=binner.flags).spread(x, F, bins)
    out = grid.zeros(device=gpu)
    noalloc(x, F, grid, (8, 8, 8), bins, out, binner.flags)
    assert int(binner.flags.item()) == 0
    assert max(rel(a.cpu(), b.cpu()) for a, b in zip(out, ref)) < 1e-12
    none = torch.zeros((0, 3), device=gpu, dtype=torch.float64)
    empty = Binner(grid, 0, gpu)(none)
    noalloc(none, none, grid, (8, 8, 8), empty, out, binner.flags)
    assert all(int(c.count_nonzero()) == 0 for c in out)


def test_noalloc_spread_rejects_too_many_tiles(gpu):
    grid = MACGrid((256, 256, 256), (1., 1., 1.), periodic=True)
    x = torch.full((1, 3), .5, device=gpu, dtype=torch.float64)
    binner = Binner(grid, 1, gpu, (2, 2, 2))
    bins = binner(x)
    out = grid.zeros(device=gpu)  # full size: no launch could leave the arrays
    with pytest.raises(RuntimeError, match='exceeds the no-allocation'):
        noalloc(x, x.clone(), grid, (2, 2, 2), bins, out, binner.flags)
