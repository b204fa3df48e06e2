"""CSR-assembled quadrature IB transfer for P1 meshes on the GPU: the frozen
operator of mac/assembled_transfer.py:1-8 and :138-211
(``AssembledP1Transfer`` with hash assembly and device contraction),

    B_c[a, f] = sum_e sum_q W_eq N_qa phi_c(f - x_eq),

built once per ``prepare`` and then applied by CSR products in both
directions: spread = B_c^T (M^-1 b) / V, interpolation solves M U = B u.

Construction per component (``csrc/assembled.hip``): the cells are
contracted in parallel, one wave per cell accumulating its stencil box in
LDS, into a bounded device hash table (``HashWorkspace``; a probe overflow
discards the whole table and retries at twice the capacity, up to the
entry budget, mac/hash_transfer_assembly.py:132-155); the occupied keys are
then sorted once and split into the node-major CSR B and its face-major
transpose. Only the table's flag and the entry count reach the host.
"""
from dataclasses import dataclass
from math import prod

import torch

from .. import _abi
from .fe import CSRMatrix, FETransfer
from .transfer import from_points


def power_of_two(n):
    return 1 << max(0, (int(n) - 1).bit_length())


@dataclass
class HashWorkspace:
    """Reusable device table (mac/hash_transfer_assembly.py:26-50)."""
    capacity: int = 0
    keys: torch.Tensor = None
    values: torch.Tensor = None
    flag: torch.Tensor = None
    allocations: int = 0
    reuses: int = 0

    def reset(self, capacity, like):
        if capacity < 1 or capacity & (capacity - 1):
            raise ValueError('power-of-two hash capacity required')
        if (self.capacity != capacity or self.values is None or self.values.device != like.device or
                self.values.dtype != like.dtype):
            self.keys = torch.empty(capacity, device=like.device, dtype=torch.int64)
            self.values = torch.empty(capacity, device=like.device, dtype=like.dtype)
            self.flag = torch.empty((), device=like.device, dtype=torch.int32)
            self.capacity = capacity
            self.allocations += 1
        else:
            self.reuses += 1
        self.keys.fill_(-1)
        self.values.zero_()
        self.flag.zero_()

    def abi(self):
        return _abi.HashTable(_abi.ptr(self.keys), _abi.ptr(self.values), self.capacity, _abi.ptr(self.flag))

    @property
    def storage_bytes(self):
        return self.capacity * (8 + self.values.element_size()) + 4 if self.values is not None else 0


def accumulate(keys, values, workspace):
    """values[slot(key)] += value for raw (key, value) pairs (ib_accumulate)."""
    if keys.dtype != torch.int64 or values.dtype != torch.float64 or keys.shape != values.shape:
        raise ValueError('hash input must be int64 keys and float64 values of one shape')
    keys, values = keys.contiguous(), values.contiguous()
    _abi.call('afsi_hash_accumulate_f64', _abi.ptr(keys), _abi.ptr(values), keys.numel(), workspace.abi(),
              _abi.stream())


def sorted_entries(workspace):
    """Occupied slots in ascending key order (hash_transfer_assembly.py:73-78)."""
    occupied = workspace.keys >= 0
    keys, values = workspace.keys[occupied], workspace.values[occupied]
    keys, permutation = torch.sort(keys)
    return keys, values[permutation]


def csr_apply(matrix, x, x_stride, y, y_stride, scale):
    _abi.call('afsi_csr_spmv_f64', matrix.abi, _abi.ptr(x), x_stride, _abi.ptr(y), y_stride, float(scale),
              _abi.stream())


def _csr(rows, cols, vals, nrows):
    if vals.numel() >= 2 ** 31:
        raise RuntimeError('assembled IB entry budget exceeded')
    row_ptr = torch.zeros(nrows + 1, dtype=torch.int64, device=vals.device)
    row_ptr[1:] = torch.cumsum(torch.bincount(rows, minlength=nrows), 0)
    return CSRMatrix(row_ptr.to(torch.int32), cols.to(torch.int32).contiguous(), vals.contiguous(), nrows)


@dataclass(frozen=True)
class AssembledStencil:
    """B_c (node-major) and B_c^T (face-major) per component, with the
    assembly statistics (mac/assembled_transfer.py:15-27)."""
    gather: tuple
    spread: tuple
    assembly: dict

    @property
    def storage_bytes(self):
        return sum(m.row_ptr.numel() * 4 + m.cols.numel() * 4 + m.vals.numel() * 8
                   for m in self.gather + self.spread)


class AssembledTransfer(FETransfer):
    """FETransfer whose prepare assembles the P1 quadrature operator into CSR
    (mac/assembled_transfer.py:122-211; assembly 'hash', contraction on the
    device). Same spread/interpolate results as FETransfer up to roundoff."""

    def __init__(self, grid, geometry, *, max_entries=32_000_000, **kwargs):
        if type(max_entries) is not int or max_entries < 1:
            raise ValueError('positive CSR entry budget required')
        if geometry.cells.shape[-1] > 4 or grid.periodic:
            raise ValueError('the assembled transfer takes P1 cells (at most 4 nodes) in a closed MAC box')
        super().__init__(grid, geometry, **kwargs)
        self.max_entries = max_entries
        self.hash_workspaces = tuple(HashWorkspace() for _ in range(3))
        self.csr_builds = 0

    def _component(self, c, base, phi, remaining):
        g = self.geometry
        shape = self.grid.face_shape(c)
        nf = prod(shape)
        E, Q = g.weights.shape
        cb = base[c].view(E, Q, 3)
        sizes = (cb.amax(1) - cb.amin(1) + 4).prod(-1)
        raw = int(sizes.sum()) * g.cells.shape[1]
        limit = power_of_two(2 * remaining)
        workspace = self.hash_workspaces[c]
        capacity = min(limit, max(workspace.capacity, power_of_two(max(128, raw // 8))))
        attempts = 0
        grid = self.grid.abi()
        while True:
            workspace.reset(capacity, g.weights)
            attempts += 1
            _abi.call('afsi_p1_contract_f64', self._g, _abi.ptr(base), _abi.ptr(phi), base.shape[1], 0, c, grid,
                      workspace.abi(), _abi.stream())
            flag = int(workspace.flag.item())
            if flag & 2:
                raise ValueError('invalid P1 cell table or stencil box for the assembled operator')
            if not flag:
                break
            if capacity >= limit:
                raise RuntimeError('assembled IB hash probe/budget exceeded; raise max_entries')
            capacity = min(limit, 2 * capacity)
        keys, values = sorted_entries(workspace)
        if len(keys) > remaining:
            raise RuntimeError('assembled IB entry budget exceeded; no hash entries were truncated')
        if not bool(torch.isfinite(values).all()):
            raise FloatingPointError('invalid hash IB entries: nonfinite values')
        rows, cols = keys // nf, keys % nf
        B = _csr(rows, cols, values, g.node_count)
        order = torch.sort(cols * g.node_count + rows).indices
        BT = _csr(cols[order], rows[order], values[order], nf)
        return B, BT, dict(raw_cell_entries=raw, nnz=len(keys), hash_capacity=capacity, hash_attempts=attempts,
                           hash_workspace_bytes=workspace.storage_bytes)

    @torch.no_grad()
    def assemble_stencil(self, points):
        stencil = from_points(points.contiguous(), self.grid, check=False)
        gather, spread, stats = [], [], []
        remaining = self.max_entries
        for c in range(3):
            B, BT, info = self._component(c, stencil.base, stencil.phi, remaining)
            remaining -= info['nnz']
            gather.append(B)
            spread.append(BT)
            stats.append(info)
        self.csr_builds += 1
        return AssembledStencil(tuple(gather), tuple(spread),
                                dict(components=stats, total_nnz=sum(s['nnz'] for s in stats),
                                     max_entries=self.max_entries))

    @torch.no_grad()
    def prepare(self, x):
        points = self.interaction_points(x)
        self.check_support(points)
        return self.assemble_stencil(points)

    @torch.no_grad()
    def spread(self, nodal_force, stencil):
        self._nodal(nodal_force)
        if not isinstance(stencil, AssembledStencil):
            return super().spread(nodal_force, stencil)
        coefficient, info = self.solve_mass(nodal_force, self._force_coefficient if self.warm_start else None)
        if self.warm_start:
            self._force_coefficient = coefficient
        coefficient = coefficient.contiguous()
        density = []
        for c, BT in enumerate(stencil.spread):
            out = torch.empty(self.grid.face_shape(c), device=coefficient.device, dtype=coefficient.dtype)
            csr_apply(BT, coefficient[:, c], 3, out, 1, 1 / self.grid.volume)
            density.append(out)
        return tuple(density), info

    @torch.no_grad()
    def interpolate(self, velocity, stencil):
        if not isinstance(stencil, AssembledStencil):
            return super().interpolate(velocity, stencil)
        self.grid.check_velocity(velocity)
        if any(u.device != self.geometry.weights.device or u.dtype != self.geometry.weights.dtype
               or not torch.isfinite(u).all() for u in velocity):
            raise ValueError('invalid MAC interpolation field')
        rhs = torch.empty((self.geometry.node_count, 3), device=velocity[0].device, dtype=torch.float64)
        for c, (B, u) in enumerate(zip(stencil.gather, velocity)):
            csr_apply(B, u.contiguous(), 1, rhs[:, c], 3, 1.)
        result, info = self.solve_mass(rhs, self._velocity_coefficient if self.warm_start else None)
        if self.warm_start:
            self._velocity_coefficient = result
        return result, info
