"""GPU parity of stencil preparation, binning and IB transfers vs the oracle.

Bars: integer work (bases, keys, perm, CSR) bit-exact; FP64 transfers
within 1e-12 relative (l2 over the output, and max-abs relative to the
output's max) of oracle/ib3d.py on the same seeded inputs.
"""
import numpy as np
import pytest
import torch

from oracle import binning as obin
from oracle import ib3d
from afsi_gpu_amd import _abi, workloads
from afsi_gpu_amd.mac import cuda_ib
from afsi_gpu_amd.mac.grid import MACGrid
from afsi_gpu_amd.mac.transfer import TiledTransfer, bin_points, from_points

pytestmark = pytest.mark.gpu
TOL = 1e-12


def rel_err(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    scale = max(np.linalg.norm(b), 1e-300)
    return np.linalg.norm(a - b) / scale


def max_err(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def grids(periodic):
    return (MACGrid((16, 20, 24), (4., 6., 8.), (-2., -3., -4.), periodic),
            ib3d.Grid3((16, 20, 24), (4., 6., 8.), (-2., -3., -4.), periodic))


def points_in(grid, n, seed, margin=2.5):
    gen = np.random.default_rng(seed)
    lo = np.asarray(grid.origin) + margin * np.asarray(grid.spacing)
    hi = np.asarray(grid.origin) + np.asarray(grid.lengths) - margin * np.asarray(grid.spacing)
    return gen.uniform(lo, hi, size=(n, 3)), gen.normal(size=(n, 3))


@pytest.mark.parametrize('periodic', [False, True])
def test_prepare_compact_bit_exact(gpu, periodic):
    g, og = grids(periodic)
    x, _ = points_in(og, 5000, 1)
    x[:50] = np.round(x[:50] / np.asarray(og.spacing)) * np.asarray(og.spacing)  # exact nodes
    base, phi = ib3d.prepare_compact(x, og)
    st = from_points(torch.tensor(x, device=gpu), g, check=False)
    assert np.array_equal(st.base.cpu().numpy(), base)
    gphi = st.phi.cpu().numpy()
    diff = np.abs(gphi - phi)
    print('phi entries differing from the oracle:', int((diff > 0).sum()), 'max', diff.max())
    assert diff.max() <= 1e-15


def test_prepare_rejects_wall_support(gpu):
    g, og = grids(False)
    x = torch.tensor([[og.origin[0] + .1 * og.spacing[0], 0., 0.]], device=gpu, dtype=torch.float64)
    with pytest.raises(ValueError, match='support'):
        from_points(x, g)


@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
def test_compact_gather_spread(gpu, dtype):
    # The compact ops keep the reference's closed-box indexing (no wrap).
    g, og = grids(False)
    x, F = points_in(og, 3000, 2)
    base, phi = ib3d.prepare_compact(x, og)
    gen = np.random.default_rng(3)
    u = tuple(gen.normal(size=og.face_shape(c)) for c in range(3))
    U_ref = ib3d.gather_grid(u, base, phi, og)
    f_ref = ib3d.spread_grid(F, base, phi, og)
    st = from_points(torch.tensor(x, device=gpu), g)
    st = type(st)(st.base, st.phi.to(dtype))
    U = cuda_ib.compact_gather(g, tuple(torch.tensor(a, device=gpu, dtype=dtype) for a in u), st,
                               check=True)
    f = cuda_ib.compact_spread(g, torch.tensor(F, device=gpu, dtype=dtype), st, check=True)
    tol = TOL if dtype == torch.float64 else 2e-6
    assert rel_err(U.cpu(), U_ref) < tol
    for c in range(3):
        assert rel_err(f[c].cpu(), f_ref[c]) < tol
        assert tuple(f[c].shape) == og.face_shape(c)


@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
@pytest.mark.parametrize('neighbors,components', [(16, 1), (64, 3), (37, 2)])
def test_indexed_signed_duplicate_weights(gpu, dtype, neighbors, components):
    gen = np.random.default_rng(31)
    ids = gen.integers(0, 53, size=(19, neighbors))
    w = gen.normal(size=ids.shape)
    u = gen.normal(size=(53, components))
    f = gen.normal(size=(19, components))
    v = .017
    t = lambda a: torch.tensor(a, device=gpu, dtype=dtype)
    tid = torch.tensor(ids, device=gpu)
    got = cuda_ib.indexed_gather(t(u), tid, t(w), check=True)
    den = cuda_ib.indexed_spread(t(f), tid, t(w), 53, v, check=True)
    tol = TOL if dtype == torch.float64 else 3e-6
    assert rel_err(got.cpu(), ib3d.indexed_gather(u, ids, w)) < tol
    assert rel_err(den.cpu(), ib3d.indexed_spread(f, ids, w, 53, v)) < tol


@pytest.mark.parametrize('components', [1, 3])
def test_indexed_spread_skewed_nodes(gpu, components):
    """Nodes that receive thousands of links (spanning several of the
    segmented sum's 512-link spans), nodes with one link and nodes with
    none, plus out-of-range ids (skipped, flagged): against a direct sum."""
    gen = np.random.default_rng(41)
    size, points, neighbors = 300, 9000, 16
    hot = gen.integers(0, size, size=3)
    ids = np.where(gen.random((points, neighbors)) < .4, hot[gen.integers(0, 3, size=(points, neighbors))],
                   gen.integers(0, size, size=(points, neighbors)))
    ids[gen.random(ids.shape) < .01] = size + 5        # out of range
    ids[:, :][ids == 7] = 8                            # node 7 gets no links
    w = gen.normal(size=ids.shape)
    f = gen.normal(size=(points, components))
    v = .031
    valid = ids < size
    want = np.zeros((size, components))
    for c in range(components):
        contrib = (f[:, c:c + 1] * w) / v
        want[:, c] = np.bincount(ids[valid], weights=contrib[valid], minlength=size)
    assert np.bincount(ids[valid], minlength=size).max() > 2 * 512
    out = cuda_ib.indexed_spread(torch.tensor(f, device=gpu), torch.tensor(ids, device=gpu),
                                 torch.tensor(w, device=gpu), size, v, check=False)
    got = out.cpu().numpy()
    assert got[7].tolist() == [0.0] * components
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    with pytest.raises(RuntimeError, match='IB index out of range'):
        cuda_ib.indexed_spread(torch.tensor(f, device=gpu), torch.tensor(ids, device=gpu),
                               torch.tensor(w, device=gpu), size, v, check=True)


def test_indexed_empty_and_rejects_cpu(gpu):
    f = torch.ones((0, 1), device=gpu, dtype=torch.float64)
    ids = torch.zeros((0, 2), dtype=torch.int64, device=gpu)
    w = torch.ones((0, 2), device=gpu, dtype=torch.float64)
    out = cuda_ib.indexed_spread(f, ids, w, 3)
    assert out.shape == (3, 1) and out.count_nonzero().item() == 0
    with pytest.raises(ValueError, match='requires CUDA'):
        cuda_ib.indexed_gather(torch.zeros(4, 1), torch.zeros(1, 2, dtype=torch.int64), torch.ones(1, 2))
    bad = torch.tensor([[0, 99]], device=gpu)
    with pytest.raises(RuntimeError, match='out of range'):
        cuda_ib.indexed_gather(torch.ones((4, 1), device=gpu, dtype=torch.float64), bad,
                               torch.ones((1, 2), device=gpu, dtype=torch.float64), check=True)


@pytest.mark.parametrize('periodic', [False, True])
@pytest.mark.parametrize('tile', [(8, 8, 8), (4, 8, 16), (5, 3, 7)])
def test_binning_bit_exact(gpu, periodic, tile):
    g, og = grids(periodic)
    x, _ = points_in(og, 20000, 4)
    x[::5] = x[7]  # long runs of equal keys exercise stability
    if periodic:
        x[1] = np.asarray(og.origin) - .3 * np.asarray(og.spacing)
        x[2] = np.asarray(og.origin) + np.asarray(og.lengths) + .1
    keys, perm, ptr = obin.bin_points(x, og, tile)
    b = bin_points(torch.tensor(x, device=gpu), g, tile)
    b.check()
    assert np.array_equal(b.perm.cpu().numpy().astype(np.int64), perm)
    assert np.array_equal(b.keys.cpu().numpy().astype(np.uint32).astype(np.int64), keys[perm])
    assert np.array_equal(b.cell_ptr.cpu().numpy().astype(np.int64), ptr)


def test_binning_truncated_shell_bit_exact(gpu):
    x, _ = workloads.truncated_shell_64()
    og = ib3d.Grid3((64, 64, 64), (1., 1., 1.), periodic=True)
    g = MACGrid((64, 64, 64), (1., 1., 1.), periodic=True)
    keys, perm, ptr = obin.bin_points(x, og, (8, 8, 8))
    b = bin_points(torch.tensor(x, device=gpu), g, (8, 8, 8))
    assert np.array_equal(b.perm.cpu().numpy().astype(np.int64), perm)
    assert np.array_equal(b.cell_ptr.cpu().numpy().astype(np.int64), ptr)


def test_binning_empty(gpu):
    g, _ = grids(True)
    b = bin_points(torch.zeros((0, 3), device=gpu, dtype=torch.float64), g)
    assert b.cell_ptr.count_nonzero().item() == 0


@pytest.mark.parametrize('method', ['atomic', 'gather'])
@pytest.mark.parametrize('periodic', [False, True])
@pytest.mark.parametrize('tile', [(8, 8, 8), (4, 4, 4), (5, 3, 7)])
def test_tiled_spread_and_interp(gpu, periodic, tile, method):
    g, og = grids(periodic)
    x, F = points_in(og, 20000, 5)
    x[:3000] = x[0] + 1e-3 * np.random.default_rng(6).normal(size=(3000, 3))  # dense cluster
    if periodic:
        x[1] = np.asarray(og.origin) - .3 * np.asarray(og.spacing)
    base, phi = ib3d.prepare_compact(x, og)
    f_ref = ib3d.spread_grid(F, base, phi, og)
    gen = np.random.default_rng(7)
    u = tuple(gen.normal(size=og.face_shape(c)) for c in range(3))
    U_ref = ib3d.gather_grid(u, base, phi, og)
    xt, Ft = torch.tensor(x, device=gpu), torch.tensor(F, device=gpu)
    bins = bin_points(xt, g, tile)
    tr = TiledTransfer(g, gpu, tile, spread=method)
    # every face is overwritten: stale values go
    stale = tuple(torch.full(g.face_shape(c), float('nan'), device=gpu, dtype=torch.float64) for c in range(3))
    f = tr.spread(xt, Ft, bins, out=stale)
    ut = tuple(torch.tensor(a, device=gpu) for a in u)
    U = tr.interpolate(xt, ut, bins)
    assert int(tr.flags.item()) == 0
    assert all(bool(torch.isfinite(fc).all()) for fc in f)
    # unbinned path (one thread per point, HBM gathers): same weights and
    # summation order as the LDS-staged binned path, so identical bits
    assert torch.equal(tr.interpolate(xt, ut, None), U)
    for c in range(3):
        assert rel_err(f[c].cpu(), f_ref[c]) < TOL
        assert max_err(f[c].cpu(), f_ref[c]) < 10 * TOL
    assert rel_err(U.cpu(), U_ref) < TOL
    assert max_err(U.cpu(), U_ref) < 10 * TOL



@pytest.mark.parametrize('periodic', [False, True])
def test_tiled_transfer_empty_and_single_point(gpu, periodic):
    """No points: the spread fields are zero and the interpolation is (0, 3).
    One point: spread and interpolation equal the oracle's."""
    g, og = grids(periodic)
    tr = TiledTransfer(g, gpu, (8, 8, 8))
    none = torch.zeros((0, 3), device=gpu, dtype=torch.float64)
    bins = bin_points(none, g, (8, 8, 8))
    f = tr.spread(none, none.clone(), bins)
    assert all(fc.shape == tuple(g.face_shape(c)) and int(fc.count_nonzero()) == 0 for c, fc in enumerate(f))
    gen = np.random.default_rng(8)
    u = tuple(gen.normal(size=og.face_shape(c)) for c in range(3))
    ut = tuple(torch.tensor(a, device=gpu) for a in u)
    assert tr.interpolate(none, ut, bins).shape == (0, 3)
    x, F = points_in(og, 1, 9)
    base, phi = ib3d.prepare_compact(x, og)
    xt, Ft = torch.tensor(x, device=gpu), torch.tensor(F, device=gpu)
    bins = bin_points(xt, g, (8, 8, 8))
    f = tr.spread(xt, Ft, bins)
    U = tr.interpolate(xt, ut, bins)
    assert int(tr.flags.item()) == 0
    f_ref = ib3d.spread_grid(F, base, phi, og)
    for c in range(3):
        assert rel_err(f[c].cpu(), f_ref[c]) < TOL
    assert rel_err(U.cpu(), ib3d.gather_grid(u, base, phi, og)) < TOL

def test_tiled_transfer_truncated_shell(gpu):
    x, F = workloads.truncated_shell_64()
    og = ib3d.Grid3((64, 64, 64), (1., 1., 1.), periodic=True)
    g = MACGrid((64, 64, 64), (1., 1., 1.), periodic=True)
    base, phi = ib3d.prepare_compact(x, og)
    f_ref = ib3d.spread_grid(F, base, phi, og)
    U_ref = ib3d.gather_grid(f_ref, base, phi, og)
    xt, Ft = torch.tensor(x, device=gpu), torch.tensor(F, device=gpu)
    bins = bin_points(xt, g)
    tr = TiledTransfer(g, gpu)
    f = tr.spread(xt, Ft, bins)
    U = tr.interpolate(xt, f, bins)
    errs = [rel_err(f[c].cpu(), f_ref[c]) for c in range(3)] + [rel_err(U.cpu(), U_ref)]
    print('truncated-shell spread/interp l2-relative errors:', errs)
    assert max(errs) < TOL


def test_sqrt_and_division_round_like_numpy(gpu):
    # In-bounds check of the compiler's FP64 sqrt/div expansion: the
    # two-branch Peskin phi exercises both on 200k random distances.
    gen = np.random.default_rng(12)
    g, og = grids(True)
    x = gen.uniform(0, 1, size=(200000, 3)) * np.asarray(og.lengths) + np.asarray(og.origin)
    _, phi = ib3d.prepare_compact(x, og)
    st = from_points(torch.tensor(x, device=gpu), g, check=False)
    gphi = st.phi.cpu().numpy()
    assert int((gphi != phi).sum()) == 0


def test_transfer_tile_autotuner(gpu):
    from afsi_gpu_amd import autotune
    x, F = workloads.truncated_shell_64(50000)
    g = MACGrid((64, 64, 64), (1., 1., 1.), periodic=True)
    edge, times = autotune.tune_transfer_tile(g, torch.tensor(x, device=gpu), torch.tensor(F, device=gpu))
    assert edge in (4, 8) and set(times) == {4, 8}
    print('transfer tile times (ms):', times)



def test_compact_spread_deterministic_and_dense(gpu):
    """The base-tiled compact spread with a dense cluster of points sharing
    one stencil base: within 1e-12 of the oracle, and calls agree to
    rounding (the order of the LDS and HBM additions is not fixed, as in the
    reference's atomic CUDA op)."""
    g, og = grids(False)
    x, F = points_in(og, 6000, 21)
    x[:2000] = x[0] + 1e-4 * np.random.default_rng(22).normal(size=(2000, 3))
    base, phi = ib3d.prepare_compact(x, og)
    f_ref = ib3d.spread_grid(F, base, phi, og)
    st = from_points(torch.tensor(x, device=gpu), g)
    Ft = torch.tensor(F, device=gpu)
    first = cuda_ib.compact_spread(g, Ft, st)
    second = cuda_ib.compact_spread(g, Ft, st)
    for c in range(3):
        assert rel_err(first[c].cpu(), second[c].cpu().numpy()) < 1e-14
        assert rel_err(first[c].cpu(), f_ref[c]) < TOL


def test_compact_out_of_range_raises(gpu):
    """A stencil reaching outside the face array is the reference's device
    assert: the drop-in ops raise RuntimeError by default."""
    g, og = grids(False)
    x, F = points_in(og, 10, 23)
    st = from_points(torch.tensor(x, device=gpu), g)
    st.base[0, 3, 0] = -2
    with pytest.raises(RuntimeError, match='out of range'):
        cuda_ib.compact_spread(g, torch.tensor(F, device=gpu), st)
    u = tuple(torch.zeros(og.face_shape(c), device=gpu, dtype=torch.float64) for c in range(3))
    with pytest.raises(RuntimeError, match='out of range'):
        cuda_ib.compact_gather(g, u, st)


@pytest.mark.parametrize('shape,tile', [((4, 16, 16), (8, 8, 8)), ((16, 6, 4), (8, 8, 8))])
def test_tiled_transfer_periodic_axis_shorter_than_tile(gpu, shape, tile):
    """Periodic axes shorter than the tile: every staged box entry wraps into
    the face arrays (the host-side walk in test_abi checks the offsets);
    spread and interpolation equal the oracle's."""
    g = MACGrid(shape, (1., 1., 1.), periodic=True)
    og = ib3d.Grid3(shape, (1., 1., 1.), periodic=True)
    gen = np.random.default_rng(24)
    x = gen.uniform(0, 1, size=(3000, 3))
    F = gen.normal(size=(3000, 3))
    base, phi = ib3d.prepare_compact(x, og)
    f_ref = ib3d.spread_grid(F, base, phi, og)
    u = tuple(gen.normal(size=og.face_shape(c)) for c in range(3))
    U_ref = ib3d.gather_grid(u, base, phi, og)
    xt, Ft = torch.tensor(x, device=gpu), torch.tensor(F, device=gpu)
    bins = bin_points(xt, g, tile)
    tr = TiledTransfer(g, gpu, tile)
    f = tr.spread(xt, Ft, bins)
    U = tr.interpolate(xt, tuple(torch.tensor(a, device=gpu) for a in u), bins)
    assert int(tr.flags.item()) == 0
    for c in range(3):
        assert rel_err(f[c].cpu(), f_ref[c]) < TOL
    assert rel_err(U.cpu(), U_ref) < TOL


def test_ib_step_raises_on_wall_support_before_changing_state(gpu):
    """Closed box: a point whose stencil reaches a wall raises the
    reference's ValueError (mac/compact_transfer.py:94-96,115-117); the
    velocity and the points are left as they were."""
    from afsi_gpu_amd.mac import MGOptions
    from afsi_gpu_amd.step import IBStep
    g = MACGrid((16, 16, 16), (1., 1., 1.))
    gen = np.random.default_rng(25)
    x = gen.uniform(.3, .7, size=(500, 3))
    x[7] = (.02, .5, .5)
    xt = torch.tensor(x, device=gpu)
    F = torch.tensor(gen.normal(size=(500, 3)), device=gpu)
    stepper = IBStep(g, 500, dt=1e-3, device=gpu, options=MGOptions(smooth=2, smoother='rbgs'))
    u = g.zeros(device=gpu)
    for c in range(3):
        u[c].copy_(torch.randn(u[c].shape, device=gpu, dtype=torch.float64))
        u[c][tuple(slice(0, 1) if a == c else slice(None) for a in range(3))] = 0
        u[c][tuple(slice(-1, None) if a == c else slice(None) for a in range(3))] = 0
    before = [c.clone() for c in u]
    with pytest.raises(ValueError, match='support reaches a wall'):
        stepper.step(xt, F, u)
    assert all(torch.equal(a, b) for a, b in zip(u, before))
    assert torch.equal(xt.cpu(), torch.tensor(x))
    # the same stepper runs once the point moves inside the support band
    xt[7] = torch.tensor((.5, .5, .5), device=gpu, dtype=torch.float64)
    stepper.step(xt, F, u)


@pytest.mark.parametrize('periodic', [True, False])
@pytest.mark.parametrize('scale', [1., 1e-265, 1e-280, 1e265, 1e275, 'wide'])
def test_tiled_spread_force_magnitudes(gpu, periodic, scale):
    """The spread's LDS box is 64-bit fixed point with a power-of-two scale
    per work item from its largest force; items whose bound max|F|/V/8 lies
    outside [2^-900, 2^900] use FP64 LDS atomics. Forces near both ends of
    the fixed range (1e-265, 1e265), beyond them (1e-280, 1e275) and spread
    over 16 decades within one item ('wide') all meet the 1e-12 bar against
    the oracle (the fixed-point error is relative to the item's largest
    contribution, so the bar holds in the l2 and max norms of the output)."""
    g, og = grids(periodic)
    x, F = points_in(og, 20000, 31)
    x[:3000] = x[0] + 1e-3 * np.random.default_rng(32).normal(size=(3000, 3))  # dense cluster
    if scale == 'wide':
        F = F * 10. ** np.random.default_rng(33).uniform(-8, 8, size=(len(F), 1))
    else:
        F = F * scale
    base, phi = ib3d.prepare_compact(x, og)
    f_ref = ib3d.spread_grid(F, base, phi, og)
    xt, Ft = torch.tensor(x, device=gpu), torch.tensor(F, device=gpu)
    bins = bin_points(xt, g)
    tr = TiledTransfer(g, gpu)
    f = tr.spread(xt, Ft, bins)
    assert int(tr.flags.item()) == 0
    for c in range(3):
        assert bool(torch.isfinite(f[c]).all())
        # normalised first: squared norms of 1e266 overflow, of 1e-278 underflow
        top = np.abs(f_ref[c]).max()
        assert 0 < top < np.inf
        got, want = f[c].cpu().numpy() / top, f_ref[c] / top
        assert rel_err(got, want) < TOL
        assert max_err(got, want) < 10 * TOL


@pytest.mark.parametrize('periodic', [True, False])
def test_tiled_spread_nonfinite_and_zero_forces(gpu, periodic):
    """A NaN force component reaches exactly the faces the oracle's FP64 sum
    makes NaN (its work item falls back to FP64 LDS atomics), the other
    faces stay within the bar; all-zero forces give all-zero face arrays."""
    g, og = grids(periodic)
    x, F = points_in(og, 20000, 41)
    x[:3000] = x[0] + 1e-3 * np.random.default_rng(42).normal(size=(3000, 3))
    F[1500, 1] = np.nan
    base, phi = ib3d.prepare_compact(x, og)
    f_ref = ib3d.spread_grid(F, base, phi, og)
    xt, Ft = torch.tensor(x, device=gpu), torch.tensor(F, device=gpu)
    bins = bin_points(xt, g)
    tr = TiledTransfer(g, gpu)
    f = tr.spread(xt, Ft, bins)
    for c in range(3):
        got = f[c].cpu().numpy()
        nan_ref = np.isnan(f_ref[c])
        assert np.array_equal(np.isnan(got), nan_ref), c
        assert nan_ref.any() == (c == 1)
        assert rel_err(got[~nan_ref], f_ref[c][~nan_ref]) < TOL
    zero = tr.spread(xt, torch.zeros_like(Ft), bins)
    assert all(bool((fc == 0).all()) for fc in zero)
    assert int(tr.flags.item()) == 0
