"""Drop-in IB backend: the call signatures of afsi_torch ``mac/cuda_ib.py``.

``build``, ``require_cuda``, ``indexed_gather``, ``indexed_spread``,
``compact_gather``, ``compact_spread``, ``contract`` and
``hash_accumulate`` keep the reference's names,
argument meaning, ownership (fresh outputs; inputs made contiguous) and
error behaviour (mac/cuda_ib.py:38-117; mac/csrc/transfer.cpp:10-57;
mac/csrc/ib.cpp:14-54), and
run hand-written gfx950 kernels from ``libafsi_hip.so`` on the current
stream. Out-of-range indices do not trap the device as the reference's
CUDA_KERNEL_ASSERT does (mac/csrc/transfer.cu:19,41); they set a device
flag, and every call reads it back and raises ``RuntimeError('IB index out
of range')`` as the failed assert surfaces in the reference. ``check=False``
skips that one host synchronisation for callers that validate stencils
themselves. The spreads are deterministic and use no atomics: points or
links are sorted by target and every output entry sums its contributors in
a fixed order.
"""
import threading
import torch

from .. import _abi

_lock = threading.Lock()
_loaded = False


def build(*, verbose=False):
    """mac/cuda_ib.py:38-60: load the prebuilt library once; refuse to run
    without a GPU or during graph capture."""
    global _loaded
    if _loaded:
        return
    if not torch.cuda.is_available():
        raise RuntimeError('IB HIP backend requires a ROCm-enabled PyTorch and an available GPU')
    if torch.cuda.is_current_stream_capturing():
        raise RuntimeError('load the IB HIP library before graph capture')
    with _lock:
        if _loaded:
            return
        _abi.load()
        if verbose:
            print(f'afsi_hip ABI v{_abi.load().afsi_abi_version()} loaded from {_abi.LIBRARY}')
        _loaded = True


def require_cuda(tensor):
    """mac/cuda_ib.py:63-66."""
    if not tensor.is_cuda:
        raise ValueError('ib_backend=cuda requires CUDA tensors; select reference explicitly on CPU')
    build()


def _same(t, f, dtype):
    """mac/csrc/transfer.cpp:10-14 checks, as RuntimeError."""
    if not t.is_cuda or t.device != f.device:
        raise RuntimeError('IB transfer requires one CUDA device')
    if t.dtype != dtype:
        raise RuntimeError('IB transfer requires contiguous matching tensors')
    if t.requires_grad:
        raise RuntimeError('IB forward transfer does not register autograd')


def _field(f):
    if f.dtype not in (torch.float32, torch.float64):
        raise RuntimeError('IB supports float32/float64')
    _same(f, f, f.dtype)


def _suffix(dtype):
    return 'f64' if dtype == torch.float64 else 'f32'


def _flags(device):
    return torch.zeros((), dtype=torch.int32, device=device)


def _raise_flags(flags, check):
    if check and int(flags.item()):
        raise RuntimeError('IB index out of range')


def indexed_gather(field, indices, weights, *, check=True):
    """U[p,c] = sum_k field[indices[p,k],c]*weights[p,k] (mac/cuda_ib.py:69-71)."""
    require_cuda(field)
    field, indices, weights = field.contiguous(), indices.contiguous(), weights.contiguous()
    _field(field)
    _same(indices, field, torch.int64)
    _same(weights, field, field.dtype)
    if field.dim() != 2 or field.size(1) <= 0:
        raise RuntimeError('IB indexed field must be N x C')
    if indices.dim() != 2 or indices.size(1) <= 0 or indices.shape != weights.shape:
        raise RuntimeError('IB indexed stencil shape mismatch')
    if field.size(0) <= 0:
        raise RuntimeError('empty IB grid')
    out = torch.empty((indices.size(0), field.size(1)), device=field.device, dtype=field.dtype)
    flags = _flags(field.device)
    _abi.call(f'afsi_indexed_gather_{_suffix(field.dtype)}', _abi.ptr(field), field.size(0),
              field.size(1), _abi.ptr(indices), _abi.ptr(weights), indices.size(0),
              indices.size(1), _abi.ptr(out), _abi.ptr(flags), _abi.stream())
    _raise_flags(flags, check)
    return out


def indexed_spread(field, indices, weights, size, volume=1., *, check=True):
    """Fresh zeroed (size,C) with out[ids[p,k],c] += (f[p,c]*w[p,k])/volume
    (mac/cuda_ib.py:74-76)."""
    require_cuda(field)
    field, indices, weights = field.contiguous(), indices.contiguous(), weights.contiguous()
    _field(field)
    _same(indices, field, torch.int64)
    _same(weights, field, field.dtype)
    if field.dim() != 2 or field.size(1) <= 0:
        raise RuntimeError('IB indexed field must be N x C')
    if indices.dim() != 2 or indices.size(1) <= 0 or indices.shape != weights.shape:
        raise RuntimeError('IB indexed stencil shape mismatch')
    size, volume = int(size), float(volume)
    if size <= 0 or not volume > 0 or volume == float('inf'):
        raise RuntimeError('invalid IB output size/volume')
    if field.size(0) != indices.size(0):
        raise RuntimeError('IB spread point count mismatch')
    out = torch.empty((size, field.size(1)), device=field.device, dtype=field.dtype)
    flags = _flags(field.device)
    nbytes = int(_abi.load().afsi_indexed_spread_workspace_bytes(field.size(0), indices.size(1), size))
    workspace = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=field.device)
    _abi.call(f'afsi_indexed_spread_{_suffix(field.dtype)}', _abi.ptr(field), field.size(0),
              field.size(1), _abi.ptr(indices), _abi.ptr(weights), indices.size(1), size, volume,
              _abi.ptr(out), _abi.ptr(workspace), workspace.numel(), _abi.ptr(flags), _abi.stream())
    _raise_flags(flags, check)
    return out


def _compact_args(stencil, c, dtype, device):
    base, phi = stencil.base[c].contiguous(), stencil.phi[c].contiguous()
    if base.dim() != 2 or base.size(1) != 3:
        raise RuntimeError('invalid compact IB field/base')
    if phi.dim() != 3 or phi.size(0) != base.size(0) or phi.size(1) != 3 or phi.size(2) != 4:
        raise RuntimeError('invalid compact IB weights')
    if base.dtype != torch.int64 or phi.dtype != dtype or base.device != device or phi.device != device:
        raise RuntimeError('IB transfer requires contiguous matching tensors')
    return base, phi


def compact_gather(grid, velocity, stencil, *, check=True):
    """(P,3) point velocities from three face arrays (mac/cuda_ib.py:79-83)."""
    require_cuda(velocity[0])
    dtype, device = velocity[0].dtype, velocity[0].device
    count = stencil.base.shape[1]
    out = torch.empty((count, 3), device=device, dtype=dtype)
    flags = _flags(device)
    for c, u in enumerate(velocity):
        u = u.contiguous()
        _field(u)
        nx, ny, nz = grid.face_shape(c)
        if u.numel() != nx * ny * nz:
            raise RuntimeError('compact IB field length mismatch')
        base, phi = _compact_args(stencil, c, dtype, device)
        _abi.call(f'afsi_compact_gather_{_suffix(dtype)}', _abi.ptr(u), _abi.ptr(base),
                  _abi.ptr(phi), count, nx, ny, nz, _abi.c_void_p(out.data_ptr() + c * out.element_size()),
                  3, _abi.ptr(flags), _abi.stream())
    _raise_flags(flags, check)
    return out


def compact_spread(grid, force, stencil, *, check=True):
    """Three fresh face arrays f_c += (w*F_c)/V (mac/cuda_ib.py:86-90)."""
    require_cuda(force)
    force = force.contiguous()
    _field(force)
    dtype, device = force.dtype, force.device
    count = stencil.base.shape[1]
    if force.shape != (count, 3):
        raise RuntimeError('compact IB field length mismatch')
    flags = _flags(device)
    result = []
    nbytes = max(int(_abi.load().afsi_compact_spread_workspace_bytes(count, *grid.face_shape(c)))
                 for c in range(3))
    workspace = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=device)
    for c in range(3):
        shape = grid.face_shape(c)
        out = torch.empty(shape, device=device, dtype=dtype)
        base, phi = _compact_args(stencil, c, dtype, device)
        _abi.call(f'afsi_compact_spread_{_suffix(dtype)}',
                  _abi.c_void_p(force.data_ptr() + c * force.element_size()), 3, _abi.ptr(base),
                  _abi.ptr(phi), count, *shape, float(grid.volume), _abi.ptr(out), _abi.ptr(workspace),
                  workspace.numel(), _abi.ptr(flags), _abi.stream())
        result.append(out)
    _raise_flags(flags, check)
    return tuple(result)


def _hash_table(workspace):
    """The reference's HashWorkspace (keys int64, values, flag int32;
    mac/hash_transfer_assembly.py:26-50) as the C ABI's table."""
    keys, values, flag = workspace.keys, workspace.values, workspace.flag
    if keys is None or values is None or flag is None:
        raise RuntimeError('IB hash workspace is not allocated')
    n = keys.numel()
    if keys.dtype != torch.int64 or values.dtype != torch.float64 or values.numel() != n:
        raise RuntimeError('IB hash shape mismatch')
    if n <= 0 or n & (n - 1):
        raise RuntimeError('IB requires power-of-two hash capacity')
    if flag.numel() != 1 or flag.dtype != torch.int32:
        raise RuntimeError('IB flag must contain one int32')
    for tensor in (keys, values, flag):
        if not tensor.is_cuda or not tensor.is_contiguous():
            raise RuntimeError('IB tensors must be contiguous and share a CUDA device')
    return _abi.HashTable(_abi.ptr(keys), _abi.ptr(values), n, _abi.ptr(flag))


def contract(stencil, group, component, offset, low, width, prefix, count, face_shape, workspace, cache=None):
    """mac/cuda_ib.py:93-106: integrate one cell group's P1 contributions
    B_c[node, face] += W_eq N_qa phi(face - x_eq) into the device hash table
    under the key node * faces_c + face (mac/csrc/ib.cpp:58-65), without
    materialising raw entries. ``low``/``width``/``prefix``/``count`` are
    the reference's per-cell stencil boxes; they are checked for shape, and
    each cell's wave derives its own box from the stencil. A probe overflow
    sets bit 1 of ``workspace.flag`` (the reference's retry signal); an
    invalid cell table or a box outside the face arrays sets bit 2."""
    if cache is not None:
        raise NotImplementedError('the cached contraction (cached_contract_) is not provided; pass cache=None')
    if getattr(stencil, 'layout', None) == 'shared':
        raise RuntimeError('IB contraction takes the compact (3,P,3) stencil layout')
    values, weights, cells = group.values, group.weights, group.cells
    require_cuda(weights)
    base, phi = stencil.base, stencil.phi
    if base.dim() != 3 or base.size(0) != 3 or base.size(2) != 3 or base.dtype != torch.int64:
        raise RuntimeError('IB base shape mismatch')
    if phi.dim() != 4 or phi.shape[:3] != base.shape or phi.size(3) != 4 or phi.dtype != torch.float64:
        raise RuntimeError('IB phi shape mismatch')
    if values.dim() != 2 or values.size(1) != cells.size(1) or values.size(1) > 4:
        raise RuntimeError('P1 shape values must be Q x 4')
    if cells.dim() != 2 or cells.size(1) != values.size(1):
        raise RuntimeError('P1 cells must be E x 4')
    E, Q = weights.shape
    if cells.size(0) != E or values.size(0) != Q:
        raise RuntimeError('IB quadrature shape mismatch')
    if low.shape != (E, 3) or width.shape != low.shape or prefix.numel() != E + 1:
        raise RuntimeError('IB bounds shape mismatch')
    offset, count, component = int(offset), int(count), int(component)
    if not (0 <= offset <= base.size(1)) or weights.numel() > base.size(1) - offset:
        raise RuntimeError('IB site offset out of range')
    if count < 0 or (count and not E) or component not in (0, 1, 2):
        raise RuntimeError('invalid IB site count or component')
    nf = tuple(int(n) for n in face_shape)
    if len(nf) != 3 or min(nf) <= 0:
        raise RuntimeError('invalid IB face shape')
    cells64 = cells.to(torch.int64).contiguous()
    vals = values.to(torch.float64).contiguous()
    wts = weights.to(torch.float64).contiguous()
    # node ids are checked against the largest count whose keys fit int64
    # (no host read of the cell table)
    node_bound = ((1 << 62) // (nf[0] * nf[1] * nf[2])) - 1
    geometry = _abi.FEGeometry(_abi.ptr(cells64), _abi.ptr(vals), _abi.ptr(wts), E, cells.size(1), Q,
                               node_bound, None, None)
    grid = _abi.Grid3()
    for a in range(3):
        grid.n[a] = nf[a] - (1 if a == component else 0)
        grid.h[a] = 1.
    grid.periodic = 0
    if E:
        _abi.call('afsi_p1_contract_f64', geometry, _abi.ptr(base.contiguous()), _abi.ptr(phi.contiguous()),
                  base.size(1), offset, component, grid, _hash_table(workspace), _abi.stream())


def hash_accumulate(keys, values, workspace):
    """mac/cuda_ib.py:109-112: values[slot(key)] += value for raw (key,
    value) pairs, zeros skipped, SplitMix64 slots with linear probing
    (mac/csrc/ib.cu:16-54,154-166); an overflow sets workspace.flag."""
    require_cuda(values)
    if keys.dim() != 1 or values.dim() != 1 or keys.numel() != values.numel():
        raise RuntimeError('IB hash shape mismatch')
    if keys.dtype != torch.int64 or values.dtype != torch.float64:
        raise RuntimeError('IB hash keys must be int64 and values float64')
    keys, values = keys.contiguous(), values.contiguous()
    _abi.call('afsi_hash_accumulate_f64', _abi.ptr(keys), _abi.ptr(values), keys.numel(), _hash_table(workspace),
              _abi.stream())
