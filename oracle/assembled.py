"""CPU restatement of the CSR-assembled quadrature operator for P1 meshes
(test infrastructure; used only by tests).

mac/assembled_transfer.py:1-8: for component c,
    B_c[a, f] = sum_e sum_q W_eq N_qa phi_c(f - x_eq),  a = cells[e, local],
spread = B_c^T (M^-1 b) / V, gather = B_c u. The reduction over cells goes
through the bounded table of mac/hash_transfer_assembly.py:13-70: SplitMix64
finaliser bucket, linear probing of at most 128 slots, -1 = empty.
"""
import numpy as np

from . import ib3d

MASK = (1 << 64) - 1


def splitmix_finalise(key):
    """The SplitMix64 output mix of a 64-bit value (hash_transfer_assembly.py:17-23)."""
    h = int(key) & MASK
    h = ((h ^ (h >> 30)) * 0xbf58476d1ce4e5b9) & MASK
    h = ((h ^ (h >> 27)) * 0x94d049bb133111eb) & MASK
    return h ^ (h >> 31)


def bucket(key, capacity):
    return splitmix_finalise(key) & (capacity - 1)


def table_accumulate(keys, values, capacity, probes=128):
    """Sequential insertion (hash_transfer_assembly.py:55-70): returns the
    key and value arrays of the table and the overflow flag."""
    tk = np.full(capacity, -1, dtype=np.int64)
    tv = np.zeros(capacity)
    flag = 0
    for key, value in zip(np.asarray(keys).tolist(), np.asarray(values).tolist()):
        if value == 0:
            continue
        if key < 0:
            flag = 1
            continue
        slot = bucket(key, capacity)
        for _ in range(probes):
            if tk[slot] in (-1, key):
                tk[slot] = key
                tv[slot] += value
                break
            slot = (slot + 1) & (capacity - 1)
        else:
            flag = 1
    return tk, tv, flag


def sorted_entries(tk, tv):
    """Occupied slots in ascending key order (hash_transfer_assembly.py:73-78)."""
    keep = tk >= 0
    order = np.argsort(tk[keep], kind='stable')
    return tk[keep][order], tv[keep][order]


def dense_operator(g, points, grid, c):
    """Dense B_c (node_count, faces_c) from the compact stencil of the
    cell-major interaction points."""
    base, phi = ib3d.prepare_compact(points, grid)
    ids, w = ib3d.expanded_component(base, phi, grid, c)
    nf = int(np.prod(grid.face_shape(c)))
    B = np.zeros((g.node_count, nf))
    E, Q = g.weights.shape
    for a in range(g.cells.shape[1]):
        rows = np.repeat(g.cells[:, a], Q * 64)
        vals = (g.weights.reshape(-1)[:, None] * np.tile(g.values[:, a], E)[:, None] * w).reshape(-1)
        np.add.at(B, (rows, ids.reshape(-1)), vals)
    return B
