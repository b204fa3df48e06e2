"""Pin the binning oracle (oracle/binning.py) without a GPU.

The key is an integer function of the reference's stencil base; the checks
are independent: the cell brackets the point on the physical lattice, the
tile-major key decodes back to the cell, perm sorts the keys and keeps
equal keys in input order (stability, checked by brute force), and the CSR
offsets count points per key (brute-force loop).
"""
import numpy as np
import pytest

from oracle import binning
from oracle.ib3d import Grid3, prepare_compact


@pytest.mark.parametrize('periodic', [False, True])
@pytest.mark.parametrize('tile', [(8, 8, 8), (4, 8, 16), (5, 3, 7)])
def test_cells_keys_and_csr(periodic, tile):
    grid = Grid3((20, 24, 32), (2., 3., 4.), (-1., .5, 2.), periodic)
    gen = np.random.default_rng(17)
    lo = np.asarray(grid.origin) + 2.2 * np.asarray(grid.spacing)
    hi = np.asarray(grid.origin) + np.asarray(grid.lengths) - 2.2 * np.asarray(grid.spacing)
    x = gen.uniform(lo, hi, size=(3000, 3))
    x[::7] = x[3]  # many exact duplicates -> long runs of equal keys
    if periodic:
        x[1] = np.asarray(grid.origin) - .3 * np.asarray(grid.spacing)
    cells = binning.cell_coords(x, grid)
    h, o, n = np.asarray(grid.spacing), np.asarray(grid.origin), np.asarray(grid.shape)
    unwrapped = np.floor((x - o) / h)
    assert (np.mod(unwrapped, n) == cells).all()
    base, _ = prepare_compact(x, grid)
    for a in range(3):
        expected = base[a][:, a] + 1
        if periodic:
            expected %= grid.shape[a]
        assert (cells[:, a] == expected).all()
    keys, perm, ptr = binning.bin_points(x, grid, tile)
    # decode tile-major keys
    tvol = int(np.prod(tile))
    nt = binning.tile_counts(grid.shape, tile)
    tid, local = keys // tvol, keys % tvol
    t = np.stack([tid // (nt[1] * nt[2]), (tid // nt[2]) % nt[1], tid % nt[2]], 1)
    l = np.stack([local // (tile[1] * tile[2]), (local // tile[2]) % tile[1], local % tile[2]], 1)
    assert (t * np.asarray(tile) + l == cells).all()
    # perm: sorted and stable
    assert sorted(perm.tolist()) == list(range(len(x)))
    sk = keys[perm]
    assert (np.diff(sk) >= 0).all()
    for i in range(1, len(perm)):
        if sk[i] == sk[i - 1]:
            assert perm[i] > perm[i - 1]
    # CSR by brute force
    counts = {}
    for k in keys.tolist():
        counts[k] = counts.get(k, 0) + 1
    assert ptr[0] == 0 and ptr[-1] == len(x)
    assert len(ptr) == binning.key_space(grid.shape, tile) + 1
    for k in list(counts)[:200]:
        assert ptr[k + 1] - ptr[k] == counts[k]
        assert (sk[ptr[k]:ptr[k + 1]] == k).all()
