"""P2 fused cell-block transfer on the GPU: ``prepare``, ``spread`` and
``assemble_velocity`` of mac/p2_cuda_transfer.py:27-67, with the shared
face/centre stencil tables of mac/shared_stencil.py:7-19.

The interaction points of 10-node tetrahedra are never materialised:
``prepare`` evaluates them in registers and writes the shared tables (base
(2,P,3) int64, phi (2,P,3,4); lattice 0 faces, lattice 1 centres), and the
spread and gather kernels (``csrc/p2.hip``) work one element at a time
through an LDS patch of the element's stencil box. A rule is a tuple of
quadrature groups; a group's rows occupy ``offset .. offset + E*Q`` of the
tables, in group order, as in the reference.

Kernel-side validation (a stencil box leaving the face arrays, a node id
out of range) skips the element and sets a device flag; ``prepare`` reads
that flag together with its own support flag in its one host
synchronisation and raises there. There is no CPU fallback.
"""
from ctypes import c_void_p
from dataclasses import dataclass

import torch

from .. import _abi
from .cuda_ib import require_cuda
from .transfer import CompactStencil

SUPPORT_MESSAGE = 'MAC IB support reaches a wall or is nonfinite; enlarge/refine the fluid box'


@dataclass(frozen=True)
class RuleGroup:
    """One interaction quadrature group: ``cells`` (E,10) int64,
    ``values`` N (Q,10) at the reference points, ``weights`` W (E,Q)."""
    cells: torch.Tensor
    values: torch.Tensor
    weights: torch.Tensor
    order: int = 0


@dataclass(frozen=True)
class Rule:
    groups: tuple

    @property
    def point_count(self):
        return sum(g.weights.numel() for g in self.groups)


@dataclass(frozen=True)
class SharedStencil:
    """Shared tables of one prepare and the rule that produced them."""
    base: torch.Tensor
    phi: torch.Tensor
    rule: Rule
    layout = 'shared'

    @property
    def storage_bytes(self):
        return self.base.numel() * self.base.element_size() + self.phi.numel() * self.phi.element_size()

    def expanded(self):
        """Per-component compact tables (mac/shared_stencil.py:13-19):
        component c takes lattice 0 along axis c and lattice 1 elsewhere."""
        base = torch.stack([torch.stack([self.base[int(a != c), :, a] for a in range(3)], -1) for c in range(3)])
        phi = torch.stack([torch.stack([self.phi[int(a != c), :, a, :] for a in range(3)], 1) for c in range(3)])
        return CompactStencil(base.contiguous(), phi.contiguous())


def rule_from_geometry(geometry):
    """The single-group rule of an FE geometry whose cells have ten nodes."""
    return Rule((RuleGroup(geometry.cells.contiguous(), geometry.values.contiguous(),
                           geometry.weights.contiguous()),))


def check_rule(rule, node_count, like, *, ids=True):
    """Host validation of a rule's tables: shapes, dtypes and devices, and
    with ``ids`` the node-id range (two reductions; once per rule)."""
    for g in rule.groups:
        if (g.cells.dtype != torch.int64 or g.cells.ndim != 2 or g.cells.shape[1] != 10 or
                g.values.ndim != 2 or g.values.shape[1] != 10 or g.values.shape[0] == 0):
            raise ValueError('P2 transfer requires ten basis functions')
        if g.weights.shape != (g.cells.shape[0], g.values.shape[0]):
            raise ValueError('P2 integration weights mismatch')
        for t in (g.cells, g.values, g.weights):
            if t.device != like.device or not t.is_contiguous():
                raise ValueError('P2 IB requires contiguous tensors on one device')
        if g.values.dtype != like.dtype or g.weights.dtype != like.dtype:
            raise ValueError('P2 IB dtype mismatch')
        if ids and g.cells.numel() and (int(g.cells.min()) < 0 or int(g.cells.max()) >= node_count):
            raise ValueError('P2 cell table refers to a missing node')


def _status(transfer, device):
    """Device words: [0] prepare support flags, [1] sticky kernel flags."""
    status = getattr(transfer, '_p2_status', None)
    if status is None or status.device != device:
        status = torch.zeros(2, dtype=torch.int32, device=device)
        transfer._p2_status = status
    return status


def _group(g, node_count):
    return _abi.P2Group(_abi.ptr(g.cells), _abi.ptr(g.values), _abi.ptr(g.weights), g.cells.shape[0],
                        g.values.shape[0], node_count)


def _tables(base, phi, offset):
    return _abi.P2Tables(_abi.ptr(base), _abi.ptr(phi), base.shape[1], offset)


def _suffix(dtype):
    if dtype == torch.float64:
        return 'f64'
    if dtype == torch.float32:
        return 'f32'
    raise RuntimeError('P2 IB requires FP32/FP64')


def _numel(shape):
    n = 1
    for s in shape:
        n *= s
    return n


def _word(status, i):
    return c_void_p(status.data_ptr() + 4 * i)


def _check_stencil(stencil, like):
    total = stencil.rule.point_count
    if (stencil.base.shape != (2, total, 3) or stencil.phi.shape != (2, total, 3, 4) or
            stencil.base.dtype != torch.int64 or stencil.phi.dtype != like.dtype or
            stencil.base.device != like.device or stencil.phi.device != like.device or
            not stencil.base.is_contiguous() or not stencil.phi.is_contiguous()):
        raise RuntimeError('P2 IB requires shared face/center tables')
    check_rule(stencil.rule, 0, like, ids=False)


def prepare(transfer, x, rule, *, buffers=None):
    """Shared tables of the interaction points of ``x`` (mac/p2_cuda_transfer.py:27-46).

    ``buffers`` = (base, phi) to reuse; raises the reference's 'invalid P2
    stencil buffers' and 'support' ValueErrors."""
    require_cuda(x)
    suffix = _suffix(x.dtype)
    if x.ndim != 2 or x.shape[1] != 3:
        raise ValueError('P2 positions must be (N,3)')
    total = rule.point_count
    if buffers is None:
        base = torch.empty((2, total, 3), device=x.device, dtype=torch.int64)
        phi = x.new_empty((2, total, 3, 4))
    else:
        base, phi = buffers
        if (base.shape != (2, total, 3) or phi.shape != (2, total, 3, 4) or
                base.dtype != torch.int64 or phi.dtype != x.dtype or
                base.device != x.device or phi.device != x.device or
                not base.is_contiguous() or not phi.is_contiguous()):
            raise ValueError('invalid P2 stencil buffers')
    check_rule(rule, x.shape[0], x, ids=False)
    status = _status(transfer, x.device)
    status[0].zero_()
    x = x.contiguous()
    grid = transfer.grid.abi()
    offset = 0
    for g in rule.groups:
        _abi.call(f'afsi_p2_prepare_{suffix}', _abi.ptr(x), _group(g, x.shape[0]), grid, _tables(base, phi, offset),
                  _word(status, 0), _abi.stream())
        offset += g.weights.numel()
    support, skipped = status.tolist()
    if skipped:
        status[1].zero_()
        raise RuntimeError(f'P2 transfer skipped elements with invalid tables (flags {skipped})')
    if support & 2:
        raise ValueError('P2 cell table refers to a missing node')
    if support:
        raise ValueError(SUPPORT_MESSAGE)
    return base, phi


def spread(transfer, coefficient, stencil):
    """Face force densities of the nodal force coefficient (mac/p2_cuda_transfer.py:49-56):
    fresh zeroed face arrays accumulated group by group."""
    require_cuda(coefficient)
    suffix = _suffix(coefficient.dtype)
    if coefficient.ndim != 2 or coefficient.shape[1] != 3:
        raise RuntimeError('invalid P2 nodal field')
    _check_stencil(stencil, coefficient)
    fields = transfer.grid.zeros(device=coefficient.device, dtype=coefficient.dtype)
    coefficient = coefficient.contiguous()
    status = _status(transfer, coefficient.device)
    grid = transfer.grid.abi()
    offset = 0
    for g in stencil.rule.groups:
        _abi.call(f'afsi_p2_spread_{suffix}', _abi.ptr(coefficient), _group(g, coefficient.shape[0]),
                  _tables(stencil.base, stencil.phi, offset), grid, float(transfer.grid.volume),
                  *(_abi.ptr(f) for f in fields), _word(status, 1), _abi.stream())
        offset += g.weights.numel()
    return fields


def assemble_velocity(transfer, velocity, stencil):
    """B^T W K u accumulated per element (mac/p2_cuda_transfer.py:59-67)."""
    rhs = torch.zeros_like(transfer.diagonal)
    suffix = _suffix(rhs.dtype)
    velocity = tuple(v.contiguous() for v in velocity)
    if len(velocity) != 3 or any(v.dtype != rhs.dtype or v.device != rhs.device or
                                 v.numel() != _numel(transfer.grid.face_shape(c)) for c, v in enumerate(velocity)):
        raise RuntimeError('P2 MAC field length mismatch')
    _check_stencil(stencil, rhs)
    status = _status(transfer, rhs.device)
    grid = transfer.grid.abi()
    offset = 0
    for g in stencil.rule.groups:
        _abi.call(f'afsi_p2_gather_assemble_{suffix}', *(_abi.ptr(v) for v in velocity), _group(g, rhs.shape[0]),
                  _tables(stencil.base, stencil.phi, offset), grid, _abi.ptr(rhs), _word(status, 1),
                  _abi.stream())
        offset += g.weights.numel()
    return rhs
