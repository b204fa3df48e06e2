"""Wall-reflected 16-link IB transfer of the channel (mac2d/transfer.py:61-127).

``prepare`` builds the reference's per-component (P,16) int64 ids and FP64
signed weights on the GPU (``afsi_channel_prepare_f64``); ``spread`` and
``gather`` are the reference's ``_spread``/``_gather`` with point forces
given directly, through the generic indexed kernels (the reference's CUDA
backend path, mac2d/transfer.py:108-127). The FE quadrature and consistent
mass solve around them are outside this hot path.
"""
from dataclasses import dataclass
import torch

from .. import _abi
from ..mac import cuda_ib


@dataclass(frozen=True)
class Stencil:
    indices: tuple
    weights: tuple


def prepare(grid, points):
    """mac2d/transfer.py:83-101: support check, then the reflected stencil."""
    cuda_ib.require_cuda(points)
    if points.dim() != 2 or points.shape[1] != 2 or points.dtype != torch.float64:
        raise ValueError('channel IB points must be (P,2) float64')
    points = points.contiguous()
    P = points.shape[0]
    kw = dict(device=points.device)
    ids = [torch.empty((P, 16), dtype=torch.int64, **kw) for _ in range(2)]
    w = [torch.empty((P, 16), dtype=torch.float64, **kw) for _ in range(2)]
    flags = torch.zeros((), dtype=torch.int32, **kw)
    g = grid.abi()
    _abi.call('afsi_channel_prepare_f64', _abi.ptr(points), P, g, _abi.ptr(ids[0]), _abi.ptr(w[0]),
              _abi.ptr(ids[1]), _abi.ptr(w[1]), _abi.ptr(flags), _abi.stream())
    if int(flags.item()) & 1:
        raise ValueError('valve IB points leave the channel support region')
    return Stencil(tuple(ids), tuple(w))


def spread(grid, force, stencil):
    """Face densities f_c = sum (F_c w)/V per component (mac2d/transfer.py:108-115)."""
    cuda_ib.require_cuda(force)
    force = force.contiguous()
    return tuple(cuda_ib.indexed_spread(force[:, c:c + 1].contiguous(), stencil.indices[c],
                                        stencil.weights[c], grid.face_shape(c)[0] *
                                        grid.face_shape(c)[1], grid.volume).reshape(grid.face_shape(c))
                 for c in range(2))


def gather(grid, velocity, stencil):
    """U[p, c] = sum_k u_c[ids] w (mac2d/transfer.py:123-127)."""
    return torch.cat([cuda_ib.indexed_gather(u.reshape(-1, 1), ids, w)
                      for u, ids, w in zip(velocity, stencil.indices, stencil.weights)], -1)
