"""Oracle: integer cell keys and the stable CSR of Lagrangian points.

Test infrastructure only (see ``oracle/__init__.py``).

The reference has no point binning. Its integer work on this path is the
stencil base ``floor((x-o_c)/h-1)`` (mac/compact_transfer.py:21-23;
mac/transfer.py:86-87) and the Cartesian face ids (mac/transfer.py:96-99).
The cell of a point is the face-lattice base plus one along each axis, i.e.
``base[a][:, a] + 1`` of the compact stencil, so every key is a pure integer
function of the reference's own base. The CSR is the build's own: keys in a
tile-major cell numbering, ``perm = argsort(key, stable)`` and
``cell_ptr = [0, cumsum(bincount(key))]`` (the bincount->cumsum idiom of
mac/cached_transfer_assembly.py:67-69).
"""
import numpy as np
from .ib3d import prepare_compact


def tile_counts(shape, tile):
    return tuple(-(-n // t) for n, t in zip(shape, tile))


def cell_coords(points, grid):
    """Integer cell (P,3) = face-lattice base + 1, wrapped on a periodic grid."""
    base, _ = prepare_compact(points, grid)
    cells = np.stack([base[a][:, a] + 1 for a in range(3)], -1)
    if grid.periodic:
        cells = cells % np.asarray(grid.shape)
    return cells


def tile_major_keys(cells, shape, tile):
    """key = tile_id * (Tx*Ty*Tz) + local cell id; tiles and cells in each
    tile both x-major (z fastest)."""
    tx, ty, tz = tile
    ntx, nty, ntz = tile_counts(shape, tile)
    ci, cj, ck = cells[:, 0], cells[:, 1], cells[:, 2]
    tid = ((ci // tx) * nty + cj // ty) * ntz + ck // tz
    local = ((ci % tx) * ty + cj % ty) * tz + ck % tz
    return tid * (tx * ty * tz) + local


def key_space(shape, tile):
    ntx, nty, ntz = tile_counts(shape, tile)
    return ntx * nty * ntz * tile[0] * tile[1] * tile[2]


def bin_points(points, grid, tile):
    """Returns (keys int64, perm int64, cell_ptr int64[key_space+1])."""
    cells = cell_coords(points, grid)
    keys = tile_major_keys(cells, grid.shape, tile)
    perm = np.argsort(keys, kind='stable')
    counts = np.bincount(keys, minlength=key_space(grid.shape, tile))
    cell_ptr = np.concatenate(([0], np.cumsum(counts)))
    return keys, perm, cell_ptr
