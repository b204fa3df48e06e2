"""FE interaction-point transfer on the GPU: the quadrature IB equations

    M F = b,   f = V^-1 K^T W B F,   M U = B^T W K u

of mac/transfer.py:1-5 (``FETransfer``) with the consistent FE mass solved by
the Jacobi PCG of fluid/solvers.py:39-122 / mac/mass_solver.py:39-103.

B (FE evaluation at the interaction points), W B, B^T W, the CSR mass action
and the PCG loop are HIP kernels (``csrc/fe.hip``); K is the tiled
Peskin-4 spread and interpolation of ``mac.transfer`` at the interaction
points. With ``fused=True`` (10-node tetrahedra in a closed box) the
transfer instead runs the P2 fused cell-block kernels of ``mac.p2_transfer``
(the reference's production P2 path, mac/adaptive_transfer.py:334-405 with
a CUDA-prepared shared stencil): no interaction points, point forces or bins
are materialised. The mass matrix and the node-incidence lists are
assembled once at construction (static geometry). There is no CPU fallback:
without the HIP library every call raises.
"""
from dataclasses import dataclass
from math import isfinite

import torch

from .. import _abi
from . import cuda_ib, p2_transfer
from .transfer import DEFAULT_TILE, Binner, TiledTransfer


@dataclass(frozen=True)
class FEGeometry:
    """Interaction quadrature of an FE mesh: ``cells`` (E, A) int64 node ids,
    ``values`` N (Q, A) basis values at the reference quadrature points,
    ``weights`` W (E, Q) physical quadrature weights, ``node_count``. The
    fields ``FETransfer`` reads from the reference's geometry objects
    (mac/transfer.py:31-41)."""
    cells: torch.Tensor
    values: torch.Tensor
    weights: torch.Tensor
    node_count: int


@dataclass(frozen=True)
class SolverOptions:
    """fluid/solvers.py:13-27."""
    rtol: float = 1e-10
    atol: float = 1e-12
    max_iterations: int = 1000
    recompute_every: int = 40
    check_every: int = 1

    def __post_init__(self):
        if (not isfinite(self.rtol) or not isfinite(self.atol) or self.rtol < 0 or
                self.atol < 0 or self.rtol + self.atol == 0):
            raise ValueError('finite nonnegative tolerances, at least one positive, required')
        for name in ('max_iterations', 'recompute_every', 'check_every'):
            n = getattr(self, name)
            if not isinstance(n, int) or isinstance(n, bool) or n < 1:
                raise ValueError(f'{name} must be a positive integer')


@dataclass(frozen=True)
class SolveInfo:
    """fluid/solvers.py:30-36."""
    iterations: int
    residual_norm: float
    rhs_norm: float
    tolerance: float


class CSRMatrix:
    """Device CSR (int32 row_ptr/cols, FP64 values) and its ABI view."""

    def __init__(self, row_ptr, cols, vals, rows):
        self.row_ptr, self.cols, self.vals, self.rows = row_ptr, cols, vals, int(rows)
        self.abi = _abi.CSR(_abi.ptr(row_ptr), _abi.ptr(cols), _abi.ptr(vals), self.rows, int(vals.numel()))

    @classmethod
    def from_coo(cls, rows, cols, vals, size):
        """Sum duplicates (coalesce), then compress rows."""
        a = torch.sparse_coo_tensor(torch.stack((rows, cols)), vals, (size, size)).coalesce()
        r, c = a.indices()
        counts = torch.bincount(r, minlength=size)
        row_ptr = torch.zeros(size + 1, dtype=torch.int64, device=vals.device)
        row_ptr[1:] = torch.cumsum(counts, 0)
        if row_ptr[-1].item() >= 2 ** 31:
            raise ValueError('mass matrix exceeds 2^31 nonzeros')
        return cls(row_ptr.to(torch.int32).contiguous(), c.to(torch.int32).contiguous(),
                   a.values().contiguous(), size)

    def mm(self, x, out=None):
        out = torch.empty_like(x) if out is None else out
        _abi.call('afsi_csr_spmv3_f64', self.abi, _abi.ptr(x), _abi.ptr(out), _abi.stream())
        return out


_PCG_MESSAGES = {
    1: 'PCG breakdown: operator/preconditioner must be positive definite and finite',
    3: 'nonfinite true PCG residual',
}


class MassSolver:
    """Buffered Jacobi PCG for a constant SPD matrix (mac/mass_solver.py:39-103):
    same recurrence, restart/check schedule and true-residual acceptance as
    pcg(); the workspace is allocated once."""

    def __init__(self, mass, diagonal, options=None):
        self.mass, self.diagonal = mass, diagonal.contiguous()
        self.options = options or SolverOptions()
        if not torch.isfinite(self.diagonal).all() or (self.diagonal <= 0).any():
            raise ValueError('Jacobi diagonal must be finite and positive')
        self.x = torch.empty_like(self.diagonal)
        nbytes = int(_abi.load().afsi_mass_pcg_workspace_bytes(mass.rows))
        self.workspace = torch.empty(nbytes, dtype=torch.uint8, device=self.diagonal.device)
        o = self.options
        self._opt = _abi.PCGOptions(o.rtol, o.atol, o.max_iterations, o.recompute_every, o.check_every, 0)

    @torch.no_grad()
    def solve(self, rhs, initial=None):
        for v in (rhs,) if initial is None else (rhs, initial):
            if v.shape != self.x.shape or v.dtype != self.x.dtype or v.device != self.x.device:
                raise ValueError('mass solve shape, dtype or device mismatch')
        finite = torch.isfinite(rhs).all()
        if initial is not None:
            finite = finite & torch.isfinite(initial).all()
        if not finite:
            raise ValueError('finite mass RHS and initial guess required')
        rhs = rhs.contiguous()
        initial = None if initial is None else initial.contiguous()
        info = _abi.PCGInfo()
        _abi.call('afsi_mass_pcg_f64', self.mass.abi, _abi.ptr(self.diagonal), _abi.ptr(rhs),
                  None if initial is None else _abi.ptr(initial), _abi.ptr(self.x), self._opt,
                  _abi.ptr(self.workspace), self.workspace.numel(), info, _abi.stream())
        if info.status in _PCG_MESSAGES:
            raise RuntimeError(_PCG_MESSAGES[info.status])
        if info.status == 2:
            raise RuntimeError(f'PCG did not converge after {self.options.max_iterations} iterations: '
                               f'true residual {info.residual_norm:.6e} > tolerance {info.tolerance:.6e}')
        return self.x.clone(), SolveInfo(info.iterations, info.residual_norm, info.rhs_norm, info.tolerance)


@dataclass
class FEStencil:
    """Interaction points and their binned CSR (the tiled form of the
    reference's MACStencil, mac/transfer.py:13-16)."""
    points: torch.Tensor
    bins: object


class FETransfer:
    """GPU mirror of ``FETransfer`` (mac/transfer.py:19-140): same
    constructor arguments, methods, warm start and error messages."""

    def __init__(self, grid, geometry, *, options=None, warm_start=False, tile=DEFAULT_TILE, fused=False):
        self.grid, self.geometry = grid, geometry
        if type(fused) is not bool:
            raise ValueError('fused must be a bool')
        if fused and (geometry.cells.shape[-1] != 10 or grid.periodic):
            raise ValueError('the fused P2 transfer requires 10-node cells in a closed MAC box')
        self.fused = fused
        if type(warm_start) is not bool:
            raise ValueError('warm_start must be a bool')
        self.warm_start = warm_start
        self._force_coefficient = None
        self._velocity_coefficient = None
        cells, N, W = geometry.cells, geometry.values, geometry.weights
        device = W.device
        if device.type != 'cuda':
            raise RuntimeError('FETransfer runs on the HIP device (no CPU fallback)')
        if (W <= 0).any():
            raise ValueError('positive interaction quadrature weights required')
        self.options = (SolverOptions(rtol=1e-12, atol=1e-13, max_iterations=500, check_every=4)
                        if options is None else options)
        if torch.linalg.matrix_rank(N.cpu()) < N.shape[1]:
            raise ValueError('interaction quadrature underintegrates the consistent FE mass')
        E, A = cells.shape
        Q = N.shape[0]
        local = torch.einsum('eq,qa,qb->eab', W, N, N)
        rows = cells[:, :, None].expand_as(local).reshape(-1)
        cols = cells[:, None, :].expand_as(local).reshape(-1)
        self.mass = CSRMatrix.from_coo(rows, cols, local.reshape(-1), geometry.node_count)
        self.diagonal = W.new_zeros(geometry.node_count).index_add(
            0, cells.reshape(-1), local.diagonal(dim1=-2, dim2=-1).reshape(-1))[:, None].expand(-1, 3).contiguous()
        flat = cells.reshape(-1)
        order = torch.sort(flat, stable=True).indices
        node_ptr = torch.zeros(geometry.node_count + 1, dtype=torch.int64, device=device)
        node_ptr[1:] = torch.cumsum(torch.bincount(flat, minlength=geometry.node_count), 0)
        self._node_ptr = node_ptr.to(torch.int32)
        self._node_items = order.to(torch.int32).contiguous()
        self._cells = cells.contiguous()
        self._values = N.contiguous()
        self._weights = W.contiguous()
        self._g = _abi.FEGeometry(_abi.ptr(self._cells), _abi.ptr(self._values), _abi.ptr(self._weights), E, A, Q,
                                  geometry.node_count, _abi.ptr(self._node_ptr), _abi.ptr(self._node_items))
        self.points_count = E * Q
        self.solver = MassSolver(self.mass, self.diagonal, self.options)
        self.tile = tile
        if fused:
            self.rule = p2_transfer.rule_from_geometry(geometry)
            p2_transfer.check_rule(self.rule, geometry.node_count, W)
        else:
            self.binner = Binner(grid, self.points_count, device, tile)
            self.transfer = TiledTransfer(grid, device, tile)

    def reset_warm_start(self):
        """Discard cached PCG initial guesses when starting a new trajectory."""
        self._force_coefficient = None
        self._velocity_coefficient = None

    def _nodal(self, value):
        g = self.geometry
        if (value.shape != (g.node_count, 3) or value.device != g.weights.device or
                value.dtype != g.weights.dtype or not torch.isfinite(value).all()):
            raise ValueError('invalid FE nodal field')

    def evaluate(self, value):
        """(E*Q, 3) values sum_a N_qa value[cells[e,a]]."""
        out = value.new_empty((self.points_count, 3))
        _abi.call('afsi_fe_evaluate_f64', self._g, _abi.ptr(value.contiguous()), _abi.ptr(out), _abi.stream())
        return out

    def mass_action(self, value):
        return self.mass.mm(value.contiguous())

    def solve_mass(self, rhs, initial):
        return self.solver.solve(rhs, initial)

    def weighted_force(self, value):
        out = value.new_empty((self.points_count, 3))
        _abi.call('afsi_fe_weighted_f64', self._g, _abi.ptr(value.contiguous()), _abi.ptr(out), _abi.stream())
        return out

    def interaction_points(self, x):
        self._nodal(x)
        return self.evaluate(x)

    def check_support(self, points):
        if self.grid.periodic:
            if not torch.isfinite(points).all():
                raise ValueError('MAC IB support reaches a wall; enlarge/refine the fluid box')
            return
        h = points.new_tensor(self.grid.spacing)
        scaled = (points - points.new_tensor(self.grid.origin)) / h
        if (not torch.isfinite(points).all() or (scaled < 2).any() or
                (scaled >= points.new_tensor(self.grid.shape) - 2).any()):
            raise ValueError('MAC IB support reaches a wall; enlarge/refine the fluid box')

    @torch.no_grad()
    def prepare(self, x):
        if self.fused:
            self._nodal(x)
            base, phi = p2_transfer.prepare(self, x, self.rule)
            return p2_transfer.SharedStencil(base, phi, self.rule)
        points = self.interaction_points(x)
        self.check_support(points)
        bins = self.binner(points)
        bins.check()
        return FEStencil(points, bins)

    @torch.no_grad()
    def spread(self, nodal_force, stencil):
        self._nodal(nodal_force)
        coefficient, info = self.solve_mass(nodal_force, self._force_coefficient if self.warm_start else None)
        if self.warm_start:
            self._force_coefficient = coefficient
        if isinstance(stencil, p2_transfer.SharedStencil):
            return p2_transfer.spread(self, coefficient, stencil), info
        return self.spread_grid(self.weighted_force(coefficient), stencil), info

    def spread_grid(self, force_q, stencil):
        if isinstance(stencil, p2_transfer.SharedStencil):
            return cuda_ib.compact_spread(self.grid, force_q.contiguous(), stencil.expanded())
        return self.transfer.spread(stencil.points, force_q.contiguous(), stencil.bins)

    def gather_grid(self, velocity, stencil):
        if isinstance(stencil, p2_transfer.SharedStencil):
            return cuda_ib.compact_gather(self.grid, tuple(u.contiguous() for u in velocity), stencil.expanded())
        return self.transfer.interpolate(stencil.points, [u.contiguous() for u in velocity], stencil.bins)

    def assemble_velocity(self, point_velocity):
        out = point_velocity.new_empty((self.geometry.node_count, 3))
        _abi.call('afsi_fe_assemble_f64', self._g, _abi.ptr(point_velocity.contiguous()), _abi.ptr(out),
                  _abi.stream())
        return out

    @torch.no_grad()
    def interpolate(self, velocity, stencil):
        self.grid.check_velocity(velocity)
        if any(u.device != self.geometry.weights.device or u.dtype != self.geometry.weights.dtype
               or not torch.isfinite(u).all() for u in velocity):
            raise ValueError('invalid MAC interpolation field')
        if isinstance(stencil, p2_transfer.SharedStencil):
            rhs = p2_transfer.assemble_velocity(self, velocity, stencil)
        else:
            rhs = self.assemble_velocity(self.gather_grid(velocity, stencil))
        result, info = self.solve_mass(rhs, self._velocity_coefficient if self.warm_start else None)
        if self.warm_start:
            self._velocity_coefficient = result
        return result, info
