"""Autotuning of launch shapes (build-own).

* Red-black sweeps: per multigrid level, the tile (2x2, 4x4, 8x4 or 8x8
  rows) and the
  persistent grid size (one tile per workgroup, or fewer workgroups that
  pipeline several tiles each). Every variant computes the same values
  (the per-cell arithmetic does not depend on the tiling), so tuning only
  changes speed. Cached by (level shape, dtype, boundary type).
* Transfer tile: the binning/spread/interpolation tile edge (4 or 8 cells)
  by point density, timed on the caller's points. Cached by (grid shape,
  boundary type, log2 of the points per cell).

The caches live in this process; ``export_cache``/``import_cache`` move
them between runs as plain JSON-able dicts.
"""
import math
import torch

from . import _abi

_SWEEPS = {}
_TILES = {}
# sweep_tile code -> (x, y) rows of a tile (afsi_mg3d_plan.sweep_tile)
_SWEEP_TILES = {8: (8, 8), 84: (8, 4), 4: (4, 4), 2: (2, 2)}


def _time(fn, repeats):
    fn()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    for _ in range(repeats):
        fn()
    end.record()
    torch.cuda.synchronize()
    return start.elapsed_time(end) / repeats


def tune_sweeps(solver, repeats=20):
    """Choose sweep_tile / sweep_blocks for every level of a fused
    ``GeometricMultigrid`` (red-black) and write them into its plan.
    Returns {level: (tile, blocks, microseconds)}."""
    if not solver.fused or solver.options.smoother != 'rbgs':
        return {}
    plan = solver.plan
    name = f'afsi_mg3d_sweep_{solver.suffix}'
    chosen = {}
    for level, shape in enumerate(solver.shapes):
        n0, n1, n2 = shape
        if n2 > 128 or n2 % 2 or n0 * n1 * n2 <= 4096:
            continue   # fallback kernel, or solved inside the LDS coarse cycle
        key = (tuple(shape), solver.suffix, solver.periodic)
        if key in _SWEEPS:
            tile, blocks, us = _SWEEPS[key]
        else:
            best = None
            for tile, (tx, ty) in _SWEEP_TILES.items():
                if n0 % tx or n1 % ty:
                    continue
                ntiles = (n0 // tx) * (n1 // ty)
                for blocks in sorted({ntiles, max(8, ntiles // 2 // 8 * 8), max(8, ntiles // 4 // 8 * 8)},
                                     reverse=True):
                    plan.sweep_tile[level], plan.sweep_blocks[level] = tile, blocks
                    us = 1e3 * _time(lambda: _abi.call(name, plan, level, _abi.stream()), repeats)
                    if best is None or us < best[2]:
                        best = (tile, blocks, us)
            tile, blocks, us = best
            _SWEEPS[key] = best
        plan.sweep_tile[level], plan.sweep_blocks[level] = tile, blocks
        chosen[level] = (tile, blocks, us)
    return chosen


def tune_transfer_tile(grid, x, force, *, repeats=5, candidates=(8, 4)):
    """Tile edge for binning + spread + interpolation of these points on
    this grid; returns (edge, {edge: milliseconds})."""
    from .mac.transfer import TiledTransfer, bin_points
    density = x.shape[0] / math.prod(grid.shape)
    key = (tuple(grid.shape), bool(grid.periodic), round(math.log2(max(density, 1e-6))))
    if key in _TILES:
        return _TILES[key]
    times = {}
    vel = grid.zeros(device=x.device)
    for edge in candidates:
        if any(n % edge for n in grid.shape):
            continue
        tile = (edge,) * 3
        tr = TiledTransfer(grid, x.device, tile)

        def run():
            bins = bin_points(x, grid, tile)
            tr.spread(x, force, bins)
            tr.interpolate(x, vel, bins)
        times[edge] = _time(run, repeats)
    best = min(times, key=times.get)
    _TILES[key] = (best, times)
    return best, times


def export_cache():
    return dict(sweeps={repr(k): v for k, v in _SWEEPS.items()},
                tiles={repr(k): v for k, v in _TILES.items()})


def import_cache(data):
    import ast
    for k, v in data.get('sweeps', {}).items():
        _SWEEPS[ast.literal_eval(k)] = tuple(v)
    for k, v in data.get('tiles', {}).items():
        _TILES[ast.literal_eval(k)] = (v[0], {int(a): b for a, b in v[1].items()})
