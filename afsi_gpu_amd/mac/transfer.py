"""Point-level 3D IB transfer on MAC faces: stencils, binning, tiled kernels.

* ``CompactStencil`` / ``from_points``: the reference's compact separable
  stencil (mac/compact_transfer.py:11-26,144-150), built on the GPU with
  the reference's two-branch Peskin formula; consumed by
  ``cuda_ib.compact_gather`` / ``compact_spread``.
* ``bin_points``: build-own stable CSR of points by tile-major cell key.
* ``TiledTransfer``: spread / interpolate all three components straight
  from positions over the CSR (weights in registers, per-tile LDS
  accumulation), the fast path the IB step uses. ``spread='gather'``
  selects the per-face-tile gather spread instead (every face produced by
  the one workgroup that owns its tile and written once, no atomics in HBM;
  slower on clustered workloads).
"""
from dataclasses import dataclass
import torch

from .. import _abi
from .cuda_ib import require_cuda

DEFAULT_TILE = (8, 8, 8)


@dataclass(frozen=True)
class CompactStencil:
    base: torch.Tensor  # (3,P,3) int64, lower neighbour per staggered lattice
    phi: torch.Tensor   # (3,P,3,4), four 1D weights per axis

    @property
    def storage_bytes(self):
        return self.base.numel() * self.base.element_size() + self.phi.numel() * self.phi.element_size()


def _check_points(points):
    if points.ndim != 2 or points.shape[1] != 3 or points.dtype != torch.float64:
        raise ValueError('positions must be (N,3) float64')


def from_points(points, grid, *, check=True):
    """Compact stencil of FP64 points (mac/compact_transfer.py:144-150).

    With ``check`` the support predicate of mac/compact_transfer.py:94-96,115-117
    is enforced (one host synchronisation), raising the reference's
    'support' ValueError."""
    require_cuda(points)
    _check_points(points)
    points = points.contiguous()
    count = points.shape[0]
    base = torch.empty((3, count, 3), dtype=torch.int64, device=points.device)
    phi = torch.empty((3, count, 3, 4), dtype=torch.float64, device=points.device)
    flags = torch.zeros((), dtype=torch.int32, device=points.device)
    g = grid.abi()
    _abi.call('afsi_prepare_compact_f64', _abi.ptr(points), count, g, _abi.ptr(base), _abi.ptr(phi),
              _abi.ptr(flags), _abi.stream())
    if check and int(flags.item()):
        raise ValueError('MAC IB support reaches a wall or is nonfinite; enlarge/refine the fluid box')
    return CompactStencil(base, phi)


@dataclass
class PointBins:
    """Stable CSR of points: keys_sorted (P) uint32 as int32 storage,
    perm (P) int32 = stable argsort(keys), cell_ptr (key_space+1) int32."""
    tile: tuple
    keys: torch.Tensor
    perm: torch.Tensor
    cell_ptr: torch.Tensor
    flags: torch.Tensor

    def check(self):
        return raise_flags(int(self.flags.item()))


F_SUPPORT, F_INDEX, F_NONFINITE = 1, 2, 4


def raise_flags(bits):
    """The reference's errors for the kernels' flag bits: a point whose
    support reaches a closed wall or is nonfinite raises its ValueError
    (mac/compact_transfer.py:94-96,115-117); a stencil index outside the
    face arrays raises RuntimeError (the device assert of
    mac/csrc/transfer.cu:19,41)."""
    if bits & (F_SUPPORT | F_NONFINITE):
        raise ValueError('MAC IB support reaches a wall or is nonfinite; enlarge/refine the fluid box')
    if bits & F_INDEX:
        raise RuntimeError('IB index out of range')
    return bits


class Binner:
    """Reusable buffers for ``afsi_bin_points`` at a fixed point count."""

    def __init__(self, grid, count, device, tile=DEFAULT_TILE):
        self.grid, self.count, self.tile = grid, int(count), tuple(int(t) for t in tile)
        self._g = grid.abi()
        self._tile = _abi.i32x3(self.tile)
        _abi.load()
        space = _abi.load().afsi_bin_key_space(self._g, self._tile)
        if space <= 0:
            raise ValueError('invalid binning tile')
        self.key_space = int(space)
        self.keys = torch.empty(self.count, dtype=torch.int32, device=device)
        self.perm = torch.empty(self.count, dtype=torch.int32, device=device)
        self.cell_ptr = torch.empty(self.key_space + 1, dtype=torch.int32, device=device)
        self.flags = torch.zeros((), dtype=torch.int32, device=device)
        nbytes = int(_abi.load().afsi_bin_workspace_bytes(self.count))
        self.workspace = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=device)

    def __call__(self, points):
        if points.shape != (self.count, 3) or points.dtype != torch.float64:
            raise ValueError('positions must match the binner (N,3) float64')
        points = points.contiguous()
        self.flags.zero_()
        _abi.call('afsi_bin_points', _abi.ptr(points), self.count, self._g, self._tile,
                  _abi.ptr(self.keys), _abi.ptr(self.perm), _abi.ptr(self.cell_ptr),
                  _abi.ptr(self.workspace), self.workspace.numel(), _abi.ptr(self.flags),
                  _abi.stream())
        return PointBins(self.tile, self.keys, self.perm, self.cell_ptr, self.flags)


def bin_points(points, grid, tile=DEFAULT_TILE):
    """One-shot binning (fresh buffers)."""
    require_cuda(points)
    _check_points(points)
    return Binner(grid, points.shape[0], points.device, tile)(points)


class TiledTransfer:
    """Spread / interpolate over binned points for a fixed grid and tile."""

    def __init__(self, grid, device, tile=DEFAULT_TILE, flags=None, spread='atomic'):
        """``flags``: an int32 device word to OR the kernels' error bits
        into (e.g. the binner's, which it clears at every binning); by
        default a word of its own. ``spread``: 'atomic' (point-binned LDS
        accumulation, the fast path) or 'gather' (per-face-tile gather)."""
        if spread not in ('atomic', 'gather'):
            raise ValueError("spread must be 'atomic' or 'gather'")
        self.method = spread
        self.grid, self.tile = grid, tuple(int(t) for t in tile)
        self._g = grid.abi()
        self._tile = _abi.i32x3(self.tile)
        nbytes = int(_abi.load().afsi_spread_workspace_bytes(self._g, self._tile))
        if nbytes == 0:
            raise ValueError('invalid grid or tile for the tiled transfer')
        self.device = torch.device(device)
        self.workspace = torch.empty(nbytes, dtype=torch.uint8, device=device)
        self.spread_workspace = None
        self.flags = torch.zeros((), dtype=torch.int32, device=device) if flags is None else flags

    def _spread_workspace(self, count):
        """Workspace of the gather spread for ``count`` points (halo counts,
        work list, scratch slots of split tiles); grown on demand."""
        nbytes = int(_abi.load().afsi_spread_gather_workspace_bytes(self._g, self._tile, count))
        if nbytes == 0:
            raise ValueError('invalid grid or tile for the tiled spread')
        if self.spread_workspace is None or self.spread_workspace.numel() < nbytes:
            self.spread_workspace = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        return self.spread_workspace

    def spread(self, points, force, bins, out=None):
        """Three face arrays f_c = sum_p w_c(p) F_c(p) / V (every face
        overwritten)."""
        if out is None:
            out = self.grid.zeros(device=points.device, dtype=torch.float64)
        if self.method == 'atomic':
            name, ws = 'afsi_spread_tiled_f64', self.workspace
        else:
            name, ws = 'afsi_spread_gather_f64', self._spread_workspace(points.shape[0])
        _abi.call(name, _abi.ptr(points), _abi.ptr(force), points.shape[0],
                  self._g, self._tile, _abi.ptr(bins.perm), _abi.ptr(bins.cell_ptr),
                  _abi.ptr(out[0]), _abi.ptr(out[1]), _abi.ptr(out[2]), _abi.ptr(ws),
                  ws.numel(), _abi.ptr(self.flags), _abi.stream())
        return out

    def interpolate(self, points, velocity, bins=None, out=None):
        """(P,3) velocities U_c(p) = sum w_c(p) u_c in the caller's point order."""
        if out is None:
            out = torch.empty_like(points)
        if bins is not None and tuple(bins.tile) != self.tile:
            raise ValueError('bins were built with another tile')
        perm = bins.perm if bins is not None else None
        cell_ptr = bins.cell_ptr if bins is not None else None
        _abi.call('afsi_interp_tiled_f64', _abi.ptr(points), points.shape[0], self._g, self._tile,
                  _abi.ptr(perm), _abi.ptr(cell_ptr), _abi.ptr(velocity[0]), _abi.ptr(velocity[1]),
                  _abi.ptr(velocity[2]), _abi.ptr(out), _abi.ptr(self.workspace),
                  self.workspace.numel(), _abi.ptr(self.flags), _abi.stream())
        return out
