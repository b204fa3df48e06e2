"""Many channel instances stepped together (build-own).

``BatchedChannelStep`` advances a list of 2D channels of uneven size by one
IB step each (the ordering of ``ChannelIBStep.step``, mac2d/coupling.py:
66-109 with point forces given) with one launch per phase over all
instances:

* one stencil launch for every point of every instance (global face ids
  into the concatenated u and v face arrays);
* one spread for both components of every point (``afsi_chb_spread_f64``:
  points binned by stencil-base tile, FP64 LDS accumulation, one atomic per
  face and tile; ``spread='indexed'`` selects the deterministic
  ``cuda_ib.indexed_spread`` per component instead) and one indexed gather
  per component (``cuda_ib.indexed_gather``);
* the predictor and the boundary() of mac2d/grid.py:74-79 as masked
  elementwise updates of the concatenated faces;
* one divergence and one gradient-correction launch;
* the pressure solves (each instance's own FP64 hierarchy): blocks of
  ``check_every`` V-cycles run one workgroup per unconverged instance, and
  per-instance residual statistics decide convergence after each block with
  the contract of ``ChannelMultigrid.solve`` (same tolerance, checks and
  messages, cycle count per instance).

Positions and forces live in batch-owned buffers (``x[i]`` / ``force[i]``
views); face velocities likewise (``velocity(i)``).
"""
from ctypes import byref, c_int64, c_size_t
from math import isfinite
import torch

from .. import _abi
from ..mac import cuda_ib
from ..mac.multigrid import MGOptions
from .multigrid import ChannelMultigrid


class BatchedChannelStep:
    def __init__(self, grids, counts, *, dt, rho=1., device='cuda', options=None, spread='tiled'):
        if not grids or len(grids) != len(counts):
            raise ValueError('one point count per channel grid required')
        if spread not in ('tiled', 'indexed'):
            raise ValueError("spread must be 'tiled' or 'indexed'")
        self.spread = spread
        self.device = torch.device(device)
        self.grids = list(grids)
        self.counts = [int(c) for c in counts]
        self.dt, self.rho = float(dt), float(rho)
        self.options = options or MGOptions(smooth=2, smoother='rbgs')
        n = len(self.grids)
        self.solvers = [ChannelMultigrid(g, device=self.device, dtype=torch.float64, options=self.options,
                                         backend='workspace') for g in self.grids]
        total = sum(self.counts)
        kw = dict(device=self.device, dtype=torch.float64)
        self.x_all = torch.zeros((total, 2), **kw)
        self.force_all = torch.zeros((total, 2), **kw)
        begins = [0]
        for c in self.counts:
            begins.append(begins[-1] + c)
        self.begins = begins
        self.x = [self.x_all[a:b] for a, b in zip(begins[:-1], begins[1:])]
        self.force = [self.force_all[a:b] for a, b in zip(begins[:-1], begins[1:])]
        self.items = (_abi.CHBItem * n)()
        for i, (g, sol) in enumerate(zip(self.grids, self.solvers)):
            self.items[i].x = self.x[i].data_ptr() if self.counts[i] else None
            self.items[i].points = self.counts[i]
            self.items[i].grid = g.abi()
            self.items[i].plan = sol.plan
        self.batch = _abi.CHBBatch()
        self.batch.count = n
        self.batch.items = self.items
        pts, nu, nv, cells, nbytes = c_int64(), c_int64(), c_int64(), c_int64(), c_size_t()
        _abi.call('afsi_chb_sizes', byref(self.batch), byref(pts), byref(nu), byref(nv), byref(cells),
                  byref(nbytes))
        self.nu, self.nv, self.cells = nu.value, nv.value, cells.value
        self.workspace = torch.empty(max(nbytes.value, 1), dtype=torch.uint8, device=self.device)
        self.batch.workspace = self.workspace.data_ptr()
        self.batch.workspace_bytes = nbytes.value
        _abi.call('afsi_chb_prepare', byref(self.batch), _abi.stream())
        nbytes = int(_abi.load().afsi_chb_spread_workspace_bytes(byref(self.batch)))
        self.spread_workspace = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=self.device)
        self.u = torch.zeros(self.nu, **kw)
        self.v = torch.zeros(self.nv, **kw)
        self.fu = torch.empty(self.nu, **kw)
        self.fv = torch.empty(self.nv, **kw)
        self.ids = [torch.empty((total, 16), dtype=torch.int64, device=self.device) for _ in range(2)]
        self.w = [torch.empty((total, 16), **kw) for _ in range(2)]
        self.flags = torch.zeros((), dtype=torch.int32, device=self.device)
        self.stats = torch.zeros((n, 4), **kw)
        self.active = torch.zeros(n, dtype=torch.int32, device=self.device)
        # per point 1/V of its grid; face masks of the predictor's boundary()
        inv_volume = torch.cat([torch.full((c, 1), 1. / g.volume, **kw) for g, c in zip(self.grids, self.counts)])
        self.inv_volume = inv_volume
        u_mask, v_mask = [], []
        for g in self.grids:
            mu = torch.ones(g.face_shape(0), **kw)
            if not g.periodic:
                mu[0] = 0      # the inlet face keeps its velocity
            mv = torch.ones(g.face_shape(1), **kw)
            mv[:, 0] = 0       # wall normal faces are zero
            mv[:, -1] = 0
            u_mask.append(mu.reshape(-1))
            v_mask.append(mv.reshape(-1))
        self.u_mask, self.v_mask = torch.cat(u_mask), torch.cat(v_mask)
        self.u_off, self.v_off = [0], [0]
        for g in self.grids:
            self.u_off.append(self.u_off[-1] + g.face_shape(0)[0] * g.face_shape(0)[1])
            self.v_off.append(self.v_off[-1] + g.face_shape(1)[0] * g.face_shape(1)[1])
        self.cycles = [0] * n

    def velocity(self, i):
        """Views of instance i's (u, v) faces in the batch buffers."""
        g = self.grids[i]
        return (self.u[self.u_off[i]:self.u_off[i + 1]].view(g.face_shape(0)),
                self.v[self.v_off[i]:self.v_off[i + 1]].view(g.face_shape(1)))

    def pressure(self, i):
        return self.solvers[i].p[0]

    def reset(self):
        """Fluid at rest: zero face velocities and pressures (the next solve
        starts from p = 0)."""
        self.u.zero_()
        self.v.zero_()
        for sol in self.solvers:
            sol.p[0].zero_()
        self.cycles = [0] * len(self.grids)

    def _stats(self, which):
        _abi.call('afsi_chb_stats_f64', byref(self.batch), which, _abi.ptr(self.stats), _abi.stream())

    def _solve(self):
        """ChannelMultigrid.solve (mac2d/multigrid.py:83-134) for every
        instance at once, from each instance's previous pressure."""
        opts = self.options
        self._stats(0)
        rhs_stats = self.stats.tolist()
        self._stats(1)
        p_stats = self.stats.tolist()
        if any(r[3] or p[3] for r, p in zip(rhs_stats, p_stats)):
            raise ValueError('finite pressure RHS/initial guess required')
        for g, sol, (s1, _, a1, _) in zip(self.grids, self.solvers, rhs_stats):
            count = g.shape[0] * g.shape[1]
            if g.periodic and abs(s1 / count) > 1e-12 + 1e-10 * a1 / count:
                raise ValueError('periodic channel pressure RHS is incompatible: net flux must vanish')
        self._stats(0)
        _abi.call('afsi_chb_stats_f64', byref(self.batch), 3, _abi.ptr(self.stats), _abi.stream())
        self._stats(1)
        _abi.call('afsi_chb_stats_f64', byref(self.batch), 4, _abi.ptr(self.stats), _abi.stream())
        self._stats(0)
        tol = [max(opts.atol, opts.rtol * q ** .5) for _, q, _, _ in self.stats.tolist()]
        self._stats(2)
        res = [q ** .5 for _, q, _, _ in self.stats.tolist()]
        if any(not isfinite(r) for r in res):
            raise RuntimeError('nonfinite channel pressure residual or RHS norm')
        cycles = [0] * len(self.grids)
        active = [r > t for r, t in zip(res, tol)]
        while any(active):
            if max(c for c, a in zip(cycles, active) if a) >= opts.max_cycles:
                raise RuntimeError('channel MG did not converge')
            count = min(opts.check_every, opts.max_cycles - max(cycles))
            self.active.copy_(torch.tensor(active, dtype=torch.int32))
            _abi.call('afsi_chb_cycles_f64', byref(self.batch), _abi.ptr(self.active), count, _abi.stream())
            self._stats(2)
            st = self.stats.tolist()
            for i, a in enumerate(active):
                if not a:
                    continue
                cycles[i] += count
                res[i] = st[i][1] ** .5
                if st[i][3] or not isfinite(res[i]):
                    raise RuntimeError('nonfinite channel pressure residual')
                active[i] = res[i] > tol[i]
        self._stats(1)
        _abi.call('afsi_chb_stats_f64', byref(self.batch), 4, _abi.ptr(self.stats), _abi.stream())
        self.cycles = cycles
        return cycles, res, tol

    @torch.no_grad()
    def step(self, *, advance=True):
        """One step of every instance; returns the (P_total, 2) point
        velocities (rows in instance order) and the solve's cycle counts."""
        self.flags.zero_()
        _abi.call('afsi_chb_stencil_f64', byref(self.batch), _abi.ptr(self.ids[0]), _abi.ptr(self.w[0]),
                  _abi.ptr(self.ids[1]), _abi.ptr(self.w[1]), _abi.ptr(self.flags), _abi.stream())
        if int(self.flags.item()) & 1:
            raise ValueError('valve IB points leave the channel support region')
        scaled = self.force_all * self.inv_volume
        if self.spread == 'tiled':
            _abi.call('afsi_chb_spread_f64', byref(self.batch), _abi.ptr(scaled), _abi.ptr(self.fu),
                      _abi.ptr(self.fv), _abi.ptr(self.spread_workspace), self.spread_workspace.numel(),
                      _abi.ptr(self.flags), _abi.stream())
            fu, fv = self.fu, self.fv
        else:
            fu = cuda_ib.indexed_spread(scaled[:, :1], self.ids[0], self.w[0], self.nu, check=False)
            fv = cuda_ib.indexed_spread(scaled[:, 1:], self.ids[1], self.w[1], self.nv, check=False)
        coef = self.dt / self.rho
        self.u.add_(coef * fu.view(-1) * self.u_mask)
        self.v.add_(coef * fv.view(-1)).mul_(self.v_mask)
        _abi.call('afsi_chb_divergence_f64', byref(self.batch), _abi.ptr(self.u), _abi.ptr(self.v),
                  -self.rho / self.dt, _abi.stream())
        cycles, res, tol = self._solve()
        if int(self.flags.item()) & 2:
            raise RuntimeError('channel spread: a stencil left its tile box')
        _abi.call('afsi_chb_correct_f64', byref(self.batch), _abi.ptr(self.u), _abi.ptr(self.v), coef,
                  _abi.stream())
        U = torch.cat((cuda_ib.indexed_gather(self.u.view(-1, 1), self.ids[0], self.w[0], check=False),
                       cuda_ib.indexed_gather(self.v.view(-1, 1), self.ids[1], self.w[1], check=False)), 1)
        if advance:
            self.x_all.add_(U, alpha=self.dt)
        return U, dict(cycles=cycles, residual_norm=res, tolerance=tol)
