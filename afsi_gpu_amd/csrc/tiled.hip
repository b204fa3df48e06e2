// Tiled Peskin-4 spread and interpolation over the binned point CSR (the
// per-face-tile gather spread, an alternative without atomics in HBM, is in
// gather_spread.hip).
//
// Spread: the work list is (tile, chunk of <= SPREAD_CHUNK points); a
// persistent grid walks it. A work item accumulates the contributions of its
// points for all three components into an LDS box covering every face
// the tile's points can touch (tile + 2 cells on each side; one less on the
// component's own axis), then adds each nonzero LDS entry to HBM with one
// FP64 global atomic. Dense tiles therefore collapse many point
// contributions into one global update per face, and a point is read and
// its weights evaluated exactly once. Time is set by the LDS atomics
// (192 per point); the box is accumulated in 64-bit fixed point, whose
// integer LDS atomics run about 1.8x the lane rate of FP64 ones for
// scattered addresses (4.4 against 2.4 lanes/clock/CU on gfx950).
//
// Interp: work items of <= INTERP_CHUNK points; the tile's face box is
// staged in LDS and each binned point gathers its 64 links per component
// from there (link order of the reference's compact kernel); results go to
// the caller's point order.
//
// Weights: per axis f = s - base - 1 and the four Peskin weights from the
// shared radical sqrt(1+4f-4f^2); base = floor(s-1) exactly as the
// reference (mac/compact_transfer.py:21-26) with s=(x-o)/h.
#include "binshare.hpp"
#include <mutex>
#include <vector>

namespace afsi {

template <class T>
int launch_fill(T* x, int64_t count, T value, hipStream_t s);

constexpr int SPREAD_THREADS = 256;
// Interpolation workgroups: the staged face box is the same for any thread
// count, so 512 threads double the wavefronts per CU that the LDS allows
// (VGPRs then set the occupancy).
constexpr int INTERP_THREADS = 512;

// Points per work item: the spread's items are short so that items of
// dense tiles spread over more workgroups (its time is set by LDS atomic
// throughput, and small items balance it); interpolation stages the whole
// face box per item, so its items are longer.
constexpr int SPREAD_CHUNK = 512;
constexpr int INTERP_CHUNK = 2048;

struct TileSpec {
    int32_t t[3];
    int64_t nt[3];
    int64_t tvol;
    int64_t ntiles;
    int32_t dim[3][3];   // LDS box extent of component c along axis a
    int32_t off[3];      // LDS offset (in doubles) of component c
    int32_t lds;         // total LDS doubles
};

static TileSpec make_spec(const afsi_grid3& g, const int32_t tile[3]) {
    TileSpec s{};
    s.tvol = 1;
    s.ntiles = 1;
    for (int a = 0; a < 3; ++a) {
        s.t[a] = tile[a];
        s.nt[a] = cdiv(g.n[a], tile[a]);
        s.tvol *= tile[a];
        s.ntiles *= s.nt[a];
    }
    int32_t total = 0;
    for (int c = 0; c < 3; ++c) {
        int32_t vol = 1;
        for (int a = 0; a < 3; ++a) {
            s.dim[c][a] = tile[a] + 4 - (a == c ? 1 : 0);
            vol *= s.dim[c][a];
        }
        s.off[c] = total;
        total += vol;
    }
    s.lds = total;
    return s;
}

// Box extents as compile-time constants for cubic tiles of edge T (T = 0:
// read them from the TileSpec at run time).
template <int T>
__host__ __device__ inline int box_dim(const TileSpec& ts, int c, int a) {
    return T > 0 ? T + 4 - (a == c ? 1 : 0) : ts.dim[c][a];
}
template <int T>
__host__ __device__ inline int box_off(const TileSpec& ts, int c) {
    return T > 0 ? c * (T + 3) * (T + 4) * (T + 4) : ts.off[c];
}
template <int T>
__device__ inline int tile_edge(const TileSpec& ts, int a) {
    return T > 0 ? T : ts.t[a];
}

__host__ __device__ inline int wrap32(int i, int n) {
    const int r = i % n;
    return r < 0 ? r + n : r;
}

struct PointWeights {
    int bf[3], bc[3];  // face-lattice and centre-lattice bases
    double wf[3][4], wc[3][4];
};

// sum_ijk wx[i] wy[j] wz[k] u[i*sx + j*sy + k], nested separably with fused
// multiply-adds: 84 FP64 operations for the 64 taps.
__device__ __forceinline__ double separable_sum(const double* __restrict__ u, int sx, int sy, const double* wx,
                                                const double* wy, const double* wz) {
    double sum = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const double* row = u + i * sx + j * sy;
            double t = row[0] * wz[0];
#pragma unroll
            for (int k = 1; k < 4; ++k) t = __builtin_fma(row[k], wz[k], t);
            s = __builtin_fma(wy[j], t, s);
        }
        sum = __builtin_fma(wx[i], s, sum);
    }
    return sum;
}

// Bases and weights of one point (mac/compact_transfer.py:21-26 bases).
// The face lattice gives s = (x - o)/h, base = floor(s - 1) and f = s - base
// - 1 in [0, 1). The centre lattice sits half a cell higher, s - 1/2, so its
// base is the face base when f >= 1/2 (fraction f - 1/2, exact) and one
// below otherwise (fraction f + 1/2): one division and one floor per axis.
__device__ inline void point_weights(const double xp[3], const afsi_grid3& g, PointWeights& pw) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double sf = (xp[a] - g.origin[a]) / g.h[a];
        const double ff = floor(sf - 1.0);
        const double f = sf - (ff + 1.0);
        const bool upper = f >= 0.5;
        pw.bf[a] = (ff > -1e9 && ff < 1e9) ? int(ff) : 0;
        pw.bc[a] = pw.bf[a] - (upper ? 0 : 1);
        peskin_weights<double>(f, pw.wf[a]);
        peskin_weights<double>(upper ? f - 0.5 : f + 0.5, pw.wc[a]);
    }
}

// Inclusive scan of one value per thread over a 1024-thread workgroup
// (wave scans with DPP-free shuffles, then the 16 wave totals).
__device__ inline int32_t block_scan_1024(int32_t v, int32_t* wave_tot) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int32_t up = __shfl_up(v, d, 64);
        if (lane >= d) v += up;
    }
    if (lane == 63) wave_tot[wave] = v;
    __syncthreads();
    int32_t before = 0;
    for (int w = 0; w < wave; ++w) before += wave_tot[w];
    __syncthreads();
    return v + before;
}

// Single workgroup: exclusive scan of ceil(points_in_tile / chunk) over the
// tiles; each thread takes a contiguous run of tiles (its boundary entries
// of cell_ptr are loaded together), one block scan joins the runs.
constexpr int SCAN_RUN = 16;

__device__ inline int32_t chunks_of(int32_t points, int chunk) { return (points + chunk - 1) / chunk; }

__global__ void __launch_bounds__(1024) chunk_scan_kernel(const int32_t* __restrict__ cell_ptr,
                                                          int64_t ntiles, int64_t tvol, int chunk,
                                                          int32_t* __restrict__ chunk_start) {
    __shared__ int32_t wave_tot[16];
    const int64_t per = (ntiles + 1023) / 1024;
    const int64_t t0 = int64_t(threadIdx.x) * per;
    const int64_t t1 = t0 + per < ntiles ? t0 + per : ntiles;
    int32_t sum = 0;
    if (per <= SCAN_RUN) {
        int32_t b[SCAN_RUN + 1];
#pragma unroll
        for (int q = 0; q <= SCAN_RUN; ++q) b[q] = (t0 + q <= t1) ? cell_ptr[(t0 + q) * tvol] : 0;
#pragma unroll
        for (int q = 0; q < SCAN_RUN; ++q)
            if (t0 + q < t1) sum += chunks_of(b[q + 1] - b[q], chunk);
        int32_t run = block_scan_1024(sum, wave_tot) - sum;
#pragma unroll
        for (int q = 0; q < SCAN_RUN; ++q)
            if (t0 + q < t1) {
                chunk_start[t0 + q] = run;
                run += chunks_of(b[q + 1] - b[q], chunk);
            }
        if (threadIdx.x == 1023) chunk_start[ntiles] = run;
        return;
    }
    for (int64_t t = t0; t < t1; ++t) sum += chunks_of(cell_ptr[(t + 1) * tvol] - cell_ptr[t * tvol], chunk);
    int32_t run = block_scan_1024(sum, wave_tot) - sum;
    for (int64_t t = t0; t < t1; ++t) {
        chunk_start[t] = run;
        run += chunks_of(cell_ptr[(t + 1) * tvol] - cell_ptr[t * tvol], chunk);
    }
    if (threadIdx.x == 1023) chunk_start[ntiles] = run;
}

// Work item w -> tile (last t with chunk_start[t] <= w), its first corner
// cell t0 and the point range [pbeg, pend) of its chunk.
struct WorkItem {
    int t0[3];
    int pbeg, pend;
};

__device__ inline WorkItem work_item(int w, const int32_t* __restrict__ chunk_start,
                                     const int32_t* __restrict__ cell_ptr, const TileSpec& ts,
                                     int chunk_points) {
    int lo = 0, hi = int(ts.ntiles);
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (chunk_start[mid] <= w) lo = mid;
        else hi = mid;
    }
    const int tile = lo;
    const int chunk = w - chunk_start[tile];
    const int nt1 = int(ts.nt[1]), nt2 = int(ts.nt[2]);
    WorkItem it;
    it.t0[0] = (tile / (nt2 * nt1)) * ts.t[0];
    it.t0[1] = ((tile / nt2) % nt1) * ts.t[1];
    it.t0[2] = (tile % nt2) * ts.t[2];
    const int tv = int(ts.tvol);
    it.pbeg = cell_ptr[tile * tv] + chunk * chunk_points;
    it.pend = min(cell_ptr[(tile + 1) * tv], it.pbeg + chunk_points);
    return it;
}

// Box entry e of component c -> global face offset, or -1 outside a closed
// box. Box axis a starts at t0[a] - 1 (own axis) or t0[a] - 2, so its
// coordinates lie in [t0 - 2, t0 + tile + 1]. A periodic coordinate wraps
// once when the tile is shorter than the axis; a tile at least as long as
// a periodic axis (e.g. a 4-cell axis under the 8-cell default tile) needs
// the full modulo, taken only on that rare branch.
template <int T>
__host__ __device__ inline int box_face(int e, int c, const int t0[3], const TileSpec& ts,
                                        const afsi_grid3& g) {
    const int d1 = box_dim<T>(ts, c, 1), d2 = box_dim<T>(ts, c, 2);
    const int l[3] = {e / (d1 * d2), (e / d2) % d1, e % d2};
    int idx = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int nf = int(g.n[a]) + ((!g.periodic && a == c) ? 1 : 0);
        int q = t0[a] - (a == c ? 1 : 2) + l[a];
        if (g.periodic) {
            if (q < 0) q += nf;
            if (q >= nf) q -= nf;
            if (q < 0 || q >= nf) q = wrap32(q, nf);
        } else if (q < 0 || q >= nf) {
            return -1;
        }
        idx = idx * nf + q;
    }
    return idx;
}

// Local box base of a point for component c (-1 in lb[0]: outside the box).
template <int T>
__device__ inline bool box_base(const PointWeights& pw, const int shift[3], int c, const int t0[3],
                                const TileSpec& ts, int lb[3]) {
    bool ok = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int b = (a == c ? pw.bf[a] : pw.bc[a]) + shift[a];
        lb[a] = b - (t0[a] - (a == c ? 1 : 2));
        ok = ok && lb[a] >= 0 && lb[a] + 3 < box_dim<T>(ts, c, a);
    }
    return ok;
}

// Periodic frame shift of a point's unwrapped bases onto its binning cell.
__device__ inline void frame_shift(const PointWeights& pw, const afsi_grid3& g, int shift[3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int cell = pw.bf[a] + 1, n = int(g.n[a]);
        // a cell already inside the domain needs no integer remainder
        shift[a] = !g.periodic || (cell >= 0 && cell < n) ? 0 : wrap32(cell, n) - cell;
    }
}

// Fixed-point LDS box. Each work item takes a power-of-two scale s from its
// largest force component: every contribution w F / volume (|w| <= 1/8, the
// Peskin weight peaks at 1/2 per axis) scaled by s stays below 2^50 in
// magnitude, is rounded to an integer by one fused multiply-add onto
// 1.5 * 2^52 (the integer then sits in the low mantissa bits) and is added to
// the box with a 64-bit integer LDS atomic. A box entry sums at most
// SPREAD_CHUNK such integers (< 2^59), so it cannot overflow, and its sum is
// independent of the order of the atomics. The flush converts each entry back
// to FP64 (one rounding) and divides by s (exact). The absolute error per
// contribution is at most 2^-51 of the item's largest contribution bound.
// On gfx950, ds_add_u64 sustains about 1.8x the lane rate of ds_add_f64 for
// scattered addresses. Items whose forces are nonfinite or whose bound lies
// outside [2^-900, 2^900] use FP64 LDS atomics instead (nonfinite forces
// then reach the face arrays as with an FP64 sum).
constexpr double FIX_MAGIC = 0x1.8p52;
constexpr long long FIX_MAGIC_BITS = 0x4338000000000000LL;

// Points per thread of one spread work item (its positions and forces are
// held in registers from the bound pass on).
constexpr int SPREAD_PPT = SPREAD_CHUNK / SPREAD_THREADS;
static_assert(SPREAD_PPT * SPREAD_THREADS == SPREAD_CHUNK, "whole points per thread");

template <int T, bool FIXED>
__device__ __forceinline__ void spread_points(const WorkItem& it, double (&px)[SPREAD_PPT][3],
                                              double (&pf)[SPREAD_PPT][3], const afsi_grid3& g,
                                              const TileSpec& ts, double volume, double scale, int* flags,
                                              double* lds) {
    unsigned long long* acc = reinterpret_cast<unsigned long long*>(lds);
    // Consecutive lanes take consecutive binned points, so a wavefront's
    // points mostly share one or two cells and stencil bases. Each lane
    // walks its point's 64 (i, j, k) links from its own rotation of all
    // three axes: at any instant the 64 lanes of one cell hit 64 different
    // LDS addresses (no same-address serialisation), and at most two share
    // a bank pair. (Measured with an FP64 box at 128^3 with 1.5M clustered
    // points: 255 us against 276 us for contiguous per-thread slices with 16
    // rotations.)
    const int ri = (threadIdx.x >> 4) & 3, rj = (threadIdx.x >> 2) & 3, rk = threadIdx.x & 3;
    // one copy of the loop body: point q is always in slot 0 (the slots
    // shift down after each point; register arrays take static indices only)
#pragma unroll 1
    for (int q = 0; q < SPREAD_PPT; ++q) {
        if (it.pbeg + int(threadIdx.x) + q * SPREAD_THREADS >= it.pend) break;
        const double f0[3] = {pf[0][0], pf[0][1], pf[0][2]};
        PointWeights pw;
        point_weights(px[0], g, pw);
#pragma unroll
        for (int r = 0; r + 1 < SPREAD_PPT; ++r) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                px[r][c] = px[r + 1][c];
                pf[r][c] = pf[r + 1][c];
            }
        }
        int shift[3];
        frame_shift(pw, g, shift);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double fv = FIXED ? (f0[c] / volume) * scale : f0[c] / volume;
            int lb[3];
            if (!box_base<T>(pw, shift, c, it.t0, ts, lb)) {
                raise_flag(flags, F_INDEX);
                continue;
            }
            const double* wx = (c == 0) ? pw.wf[0] : pw.wc[0];
            const double* wy = (c == 1) ? pw.wf[1] : pw.wc[1];
            const double* wz = (c == 2) ? pw.wf[2] : pw.wc[2];
            const int d1 = box_dim<T>(ts, c, 1), d2 = box_dim<T>(ts, c, 2);
            const int b0 = box_off<T>(ts, c) + (lb[0] * d1 + lb[1]) * d2 + lb[2];
            double wxr[4], wyr[4], wzr[4];
            int ox[4], oy[4], oz[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int iq = (r + ri) & 3, jq = (r + rj) & 3, kq = (r + rk) & 3;
                wxr[r] = iq == 0 ? wx[0] : (iq == 1 ? wx[1] : (iq == 2 ? wx[2] : wx[3]));
                wyr[r] = jq == 0 ? wy[0] : (jq == 1 ? wy[1] : (jq == 2 ? wy[2] : wy[3]));
                wzr[r] = kq == 0 ? wz[0] : (kq == 1 ? wz[1] : (kq == 2 ? wz[2] : wz[3]));
                ox[r] = iq * d1 * d2;
                oy[r] = jq * d2;
                oz[r] = kq;
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) {
                    const double wxy = wxr[i] * wyr[jj];
                    const int row = b0 + ox[i] + oy[jj];
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        if (FIXED) {
                            const double t = __builtin_fma(wxy * wzr[k], fv, FIX_MAGIC);
                            atomicAdd(acc + row + oz[k],
                                      static_cast<unsigned long long>(__double_as_longlong(t) - FIX_MAGIC_BITS));
                        } else {
                            atomicAdd(lds + row + oz[k], (wxy * wzr[k]) * fv);
                        }
                    }
                }
            }
        }
    }
}

// Flush: one global FP64 atomic per nonzero box entry.
template <int T, bool FIXED>
__device__ __forceinline__ void spread_flush(const WorkItem& it, const afsi_grid3& g, const TileSpec& ts,
                                             double inv_scale, double* const outs[3], int* flags,
                                             const double* lds) {
    const int nlds = T > 0 ? 3 * (T + 3) * (T + 4) * (T + 4) : ts.lds;
    const long long* acc = reinterpret_cast<const long long*>(lds);
    for (int e = threadIdx.x; e < nlds; e += SPREAD_THREADS) {
        double v;
        if (FIXED) {
            const long long q = acc[e];
            if (q == 0) continue;
            v = double(q) * inv_scale;
        } else {
            v = lds[e];
            if (v == 0.0) continue;
        }
        const int c = e < box_off<T>(ts, 1) ? 0 : (e < box_off<T>(ts, 2) ? 1 : 2);
        const int f = box_face<T>(e - box_off<T>(ts, c), c, it.t0, ts, g);
        if (f < 0) {
            raise_flag(flags, F_INDEX);
            continue;
        }
        unsafeAtomicAdd(outs[c] + f, v);
    }
}

// One spread work item: clear the LDS box and find the item's force bound
// (one block reduction), accumulate the points' 3 x 64 contributions, then
// flush the box to HBM. perm holds point ids offset by `base`.
template <int T>
__device__ __forceinline__ void spread_item(const WorkItem& it, const double* __restrict__ x,
                                            const double* __restrict__ force,
                                            const int32_t* __restrict__ perm, int base,
                                            const afsi_grid3& g, const TileSpec& ts, double volume,
                                            double* const outs[3], int* flags, double* lds) {
    __shared__ double wave_max[SPREAD_THREADS / 64];
    const int nlds = T > 0 ? 3 * (T + 3) * (T + 4) * (T + 4) : ts.lds;
    for (int e = threadIdx.x; e < nlds; e += SPREAD_THREADS) lds[e] = 0.0;
    // the thread's points (gathered through the bin permutation once) and
    // the item's largest |force component| (+inf if any is nonfinite)
    double px[SPREAD_PPT][3], pf[SPREAD_PPT][3];
    double m = 0.0;
#pragma unroll
    for (int q = 0; q < SPREAD_PPT; ++q) {
        const int j = it.pbeg + int(threadIdx.x) + q * SPREAD_THREADS;
        const int idx = j < it.pend ? perm[j] - base : -1;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            px[q][c] = idx >= 0 ? x[3 * idx + c] : 0.0;
            pf[q][c] = idx >= 0 ? force[3 * idx + c] : 0.0;
            const double a = fabs(pf[q][c]);
            m = a <= m ? m : (a == a ? a : __builtin_inf());
        }
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) m = fmax(m, __shfl_xor(m, d, 64));
    if ((threadIdx.x & 63) == 0) wave_max[threadIdx.x >> 6] = m;
    __syncthreads();
    m = wave_max[0];
#pragma unroll
    for (int w = 1; w < SPREAD_THREADS / 64; ++w) m = fmax(m, wave_max[w]);
    // bound of one contribution: |w| <= 1/8 times the largest |F| / volume
    const double bound = (m / volume) * 0.125;
    const bool fixed = bound <= 0x1p900 && (bound >= 0x1p-900 || bound == 0.0);
    if (fixed) {
        // 2^e > bound, so every scaled contribution is below 2^50
        const int e = bound == 0.0 ? 0 : ilogb(bound) + 1;
        spread_points<T, true>(it, px, pf, g, ts, volume, ldexp(1.0, 50 - e), flags, lds);
        __syncthreads();
        spread_flush<T, true>(it, g, ts, ldexp(1.0, e - 50), outs, flags, lds);
    } else {
        spread_points<T, false>(it, px, pf, g, ts, volume, 1.0, flags, lds);
        __syncthreads();
        spread_flush<T, false>(it, g, ts, 1.0, outs, flags, lds);
    }
    __syncthreads();
}

template <int T>
__global__ void __launch_bounds__(SPREAD_THREADS) spread_tiled_kernel(
    const double* __restrict__ x, const double* __restrict__ force,
    const int32_t* __restrict__ perm, const int32_t* __restrict__ cell_ptr,
    const int32_t* __restrict__ chunk_start, afsi_grid3 g, TileSpec ts, double volume,
    double* __restrict__ out0, double* __restrict__ out1, double* __restrict__ out2, int* flags) {
    extern __shared__ double lds[];
    const int total = chunk_start[ts.ntiles];
    double* const outs[3] = {out0, out1, out2};
    for (int w = blockIdx.x; w < total; w += gridDim.x)
        spread_item<T>(work_item(w, chunk_start, cell_ptr, ts, SPREAD_CHUNK), x, force, perm, 0, g, ts, volume,
                       outs, flags, lds);
}

// Interpolation over the same work list: the tile's face box of all three
// components is staged in LDS once per work item, then each point gathers
// its 3 x 64 values from LDS (consecutive threads take consecutive binned
// points; results go to the caller's point order).
// One interpolation work item: stage the box, gather the points' values
// (NT threads per workgroup).
template <int T, int NT>
__device__ __forceinline__ void interp_item(const WorkItem& it, const double* __restrict__ x,
                                            const int32_t* __restrict__ perm, int base,
                                            const afsi_grid3& g, const TileSpec& ts,
                                            const double* const us[3], double* __restrict__ out,
                                            int* flags, double* lds) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int vol = box_dim<T>(ts, c, 0) * box_dim<T>(ts, c, 1) * box_dim<T>(ts, c, 2);
        double* box = lds + box_off<T>(ts, c);
        for (int e = threadIdx.x; e < vol; e += NT) {
            const int f = box_face<T>(e, c, it.t0, ts, g);
            box[e] = f >= 0 ? us[c][f] : 0.0;
        }
    }
    __syncthreads();
    for (int j = it.pbeg + int(threadIdx.x); j < it.pend; j += NT) {
        const int idx = perm[j] - base;
        const double xp[3] = {x[3 * idx], x[3 * idx + 1], x[3 * idx + 2]};
        PointWeights pw;
        point_weights(xp, g, pw);
        int shift[3];
        frame_shift(pw, g, shift);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            int lb[3];
            double sum = 0.0;
            if (!box_base<T>(pw, shift, c, it.t0, ts, lb)) {
                raise_flag(flags, F_INDEX);
            } else {
                const double* wx = (c == 0) ? pw.wf[0] : pw.wc[0];
                const double* wy = (c == 1) ? pw.wf[1] : pw.wc[1];
                const double* wz = (c == 2) ? pw.wf[2] : pw.wc[2];
                const int d1 = box_dim<T>(ts, c, 1), d2 = box_dim<T>(ts, c, 2);
                const double* box = lds + box_off<T>(ts, c) + (lb[0] * d1 + lb[1]) * d2 + lb[2];
                sum = separable_sum(box, d1 * d2, d2, wx, wy, wz);
            }
            out[3 * idx + c] = sum;
        }
    }
    __syncthreads();
}

// Interpolation over the same work list: the tile's face box of all three
// components is staged in LDS once per work item, then each point gathers
// its 3 x 64 values from LDS (consecutive threads take consecutive binned
// points; results go to the caller's point order).
template <int T>
__global__ void __launch_bounds__(INTERP_THREADS) interp_lds_kernel(
    const double* __restrict__ x, const int32_t* __restrict__ perm,
    const int32_t* __restrict__ cell_ptr, const int32_t* __restrict__ chunk_start, afsi_grid3 g,
    TileSpec ts, const double* __restrict__ u0, const double* __restrict__ u1,
    const double* __restrict__ u2, double* __restrict__ out, int* flags) {
    extern __shared__ double lds[];
    const int total = chunk_start[ts.ntiles];
    const double* const us[3] = {u0, u1, u2};
    for (int w = blockIdx.x; w < total; w += gridDim.x)
        interp_item<T, INTERP_THREADS>(work_item(w, chunk_start, cell_ptr, ts, INTERP_CHUNK), x, perm, 0, g, ts, us, out,
                                       flags, lds);
}

// Unbinned interpolation: one thread per point, gathers from HBM/L2.
__global__ void __launch_bounds__(256) interp_points_kernel(
    const double* __restrict__ x, int64_t points, afsi_grid3 g, const double* __restrict__ u0,
    const double* __restrict__ u1, const double* __restrict__ u2, double* __restrict__ out,
    int* flags) {
    const double* us[3] = {u0, u1, u2};
    for (int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < points;
         j += int64_t(gridDim.x) * blockDim.x) {
        const double xp[3] = {x[3 * j], x[3 * j + 1], x[3 * j + 2]};
        PointWeights pw;
        point_weights(xp, g, pw);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            int nf[3], b[3];
            bool ok = true;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                nf[a] = int(g.n[a]) + ((!g.periodic && a == c) ? 1 : 0);
                b[a] = (a == c) ? pw.bf[a] : pw.bc[a];
                if (!g.periodic) ok = ok && b[a] >= 0 && b[a] + 3 < nf[a];
            }
            double sum = 0.0;
            if (!ok) {
                raise_flag(flags, F_INDEX);
            } else {
                const double* wx = (c == 0) ? pw.wf[0] : pw.wc[0];
                const double* wy = (c == 1) ? pw.wf[1] : pw.wc[1];
                const double* wz = (c == 2) ? pw.wf[2] : pw.wc[2];
                int zi[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) zi[k] = g.periodic ? wrap32(b[2] + k, nf[2]) : b[2] + k;
                const double* u = us[c];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int xi = g.periodic ? wrap32(b[0] + i, nf[0]) : b[0] + i;
                    double sy = 0.0;
#pragma unroll
                    for (int jj = 0; jj < 4; ++jj) {
                        const int yj = g.periodic ? wrap32(b[1] + jj, nf[1]) : b[1] + jj;
                        const double* row = u + (xi * nf[1] + yj) * nf[2];
                        double sz = row[zi[0]] * wz[0];
#pragma unroll
                        for (int k = 1; k < 4; ++k) sz = __builtin_fma(row[zi[k]], wz[k], sz);
                        sy = __builtin_fma(wy[jj], sz, sy);
                    }
                    sum = __builtin_fma(wx[i], sy, sum);
                }
            }
            out[3 * j + c] = sum;
        }
    }
}

// ---------------------------------------------------------------------------
// batched transfers: many instances binned into one key space (instance i's
// keys offset by key_offset_i) and one work list of (instance, tile, chunk)

struct XferDesc {
    const double* x;
    const double* force;
    double* face[3];
    double* out;
    afsi_grid3 g;
    TileSpec ts;
    double volume;
    int point_begin;  // first global point id
    int key_offset;   // first global key
    int begin;        // first global tile
};

struct FillDesc {
    double* p;
    int n;
    int begin;  // first workgroup
};

constexpr int FILL_CHUNK = 4096;

template <class D>
__device__ inline int find_begin(const D* __restrict__ desc, int n, int b) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (desc[mid].begin <= b) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ void __launch_bounds__(256) fill_batch_kernel(const FillDesc* __restrict__ desc, int n) {
    const int b = blockIdx.x;
    const FillDesc& d = desc[find_begin(desc, n, b)];
    const int lo = (b - d.begin) * FILL_CHUNK, hi = min(d.n, lo + FILL_CHUNK);
    for (int j = lo + int(threadIdx.x); j < hi; j += 256) d.p[j] = 0.0;
}

__device__ inline int tile_points(const XferDesc& d, const int32_t* __restrict__ cell_ptr, int t,
                                  int& beg) {
    const int tv = int(d.ts.tvol);
    beg = cell_ptr[d.key_offset + t * tv];
    return cell_ptr[d.key_offset + (t + 1) * tv] - beg;
}

// Single workgroup: exclusive scan of the chunk counts of every tile of
// every instance (global tile order).
__global__ void chunk_scan_batch_kernel(const XferDesc* __restrict__ desc, int n, int tiles,
                                        const int32_t* __restrict__ cell_ptr, int chunk,
                                        int32_t* __restrict__ chunk_start) {
    __shared__ int32_t s[1024];
    int32_t running = 0;
    for (int base = 0; base < tiles; base += 1024) {
        const int t = base + int(threadIdx.x);
        int32_t v = 0;
        if (t < tiles) {
            const XferDesc& d = desc[find_begin(desc, n, t)];
            int beg;
            v = (tile_points(d, cell_ptr, t - d.begin, beg) + chunk - 1) / chunk;
        }
        s[threadIdx.x] = v;
        __syncthreads();
        for (int dd = 1; dd < 1024; dd <<= 1) {
            const int32_t add = threadIdx.x >= dd ? s[threadIdx.x - dd] : 0;
            __syncthreads();
            s[threadIdx.x] += add;
            __syncthreads();
        }
        if (t < tiles) chunk_start[t] = running + s[threadIdx.x] - v;
        running += s[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) chunk_start[tiles] = running;
}

__device__ inline WorkItem batch_item(int w, const XferDesc* __restrict__ desc, int n, int tiles,
                                      const int32_t* __restrict__ cell_ptr,
                                      const int32_t* __restrict__ chunk_start, int chunk_points,
                                      int& inst) {
    int lo = 0, hi = tiles;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (chunk_start[mid] <= w) lo = mid;
        else hi = mid;
    }
    inst = find_begin(desc, n, lo);
    const XferDesc& d = desc[inst];
    const int t = lo - d.begin, chunk = w - chunk_start[lo];
    const int nt1 = int(d.ts.nt[1]), nt2 = int(d.ts.nt[2]);
    WorkItem it;
    it.t0[0] = (t / (nt2 * nt1)) * d.ts.t[0];
    it.t0[1] = ((t / nt2) % nt1) * d.ts.t[1];
    it.t0[2] = (t % nt2) * d.ts.t[2];
    int beg;
    const int cnt = tile_points(d, cell_ptr, t, beg);
    it.pbeg = beg + chunk * chunk_points;
    it.pend = min(beg + cnt, it.pbeg + chunk_points);
    return it;
}

template <int T>
__global__ void __launch_bounds__(SPREAD_THREADS) spread_batch_kernel(
    const XferDesc* __restrict__ desc, int n, int tiles, const int32_t* __restrict__ perm,
    const int32_t* __restrict__ cell_ptr, const int32_t* __restrict__ chunk_start, int* flags) {
    extern __shared__ double lds[];
    const int total = chunk_start[tiles];
    for (int w = blockIdx.x; w < total; w += gridDim.x) {
        int i;
        const WorkItem it = batch_item(w, desc, n, tiles, cell_ptr, chunk_start, SPREAD_CHUNK, i);
        const XferDesc& d = desc[i];
        spread_item<T>(it, d.x, d.force, perm, d.point_begin, d.g, d.ts, d.volume, d.face, flags, lds);
    }
}

template <int T>
__global__ void __launch_bounds__(SPREAD_THREADS) interp_batch_kernel(
    const XferDesc* __restrict__ desc, int n, int tiles, const int32_t* __restrict__ perm,
    const int32_t* __restrict__ cell_ptr, const int32_t* __restrict__ chunk_start, int* flags) {
    extern __shared__ double lds[];
    const int total = chunk_start[tiles];
    for (int w = blockIdx.x; w < total; w += gridDim.x) {
        int i;
        const WorkItem it = batch_item(w, desc, n, tiles, cell_ptr, chunk_start, INTERP_CHUNK, i);
        const XferDesc& d = desc[i];
        const double* const us[3] = {d.face[0], d.face[1], d.face[2]};
        interp_item<T, SPREAD_THREADS>(it, d.x, perm, d.point_begin, d.g, d.ts, us, d.out, flags, lds);
    }
}

static size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace afsi


// ---- batched transfer host side ----------------------------------------------

namespace afsi {

int launch_chunk_scan(const int32_t* cell_ptr, int64_t ntiles, int64_t tvol, int chunk, int32_t* chunk_start,
                      hipStream_t s) {
    chunk_scan_kernel<<<1, 1024, 0, s>>>(cell_ptr, ntiles, tvol, chunk, chunk_start);
    AFSI_LAUNCH_CHECK();
    return OK;
}

static int cubic_edge(const int32_t tile[3]) {
    if (tile[0] == tile[1] && tile[1] == tile[2] && (tile[0] == 8 || tile[0] == 4)) return tile[0];
    return 0;
}

static int check_tiled(const afsi_grid3& g, const int32_t tile[3]) {
    AFSI_REQUIRE(tile[0] > 0 && tile[1] > 0 && tile[2] > 0, "tile sizes must be positive");
    AFSI_REQUIRE((g.n[0] + 1) * (g.n[1] + 1) * (g.n[2] + 1) < (int64_t(1) << 31),
                 "grid exceeds 2^31 faces per component");
    return OK;
}

struct XferLayout {
    int n = 0;
    int64_t points = 0, keys = 0;
    int tiles = 0, fill_blocks = 0;
    size_t bin_off = 0, xfer_off = 0, fill_off = 0, chunk_off = 0, sort_off = 0, total = 0;
    size_t lds = 0;
};

static size_t alignb(size_t b) { return (b + 255) & ~size_t(255); }

static int xfer_layout(const afsi_tb_batch* bt, XferLayout& L) {
    AFSI_REQUIRE(bt && bt->items && bt->count >= 1, "invalid transfer batch");
    AFSI_REQUIRE(bt->tile[0] > 0 && bt->tile[1] > 0 && bt->tile[2] > 0, "tile sizes must be positive");
    L.n = bt->count;
    TileSpec ts0{};
    for (int i = 0; i < L.n; ++i) {
        const afsi_tb_item& it = bt->items[i];
        AFSI_REQUIRE(it.points >= 0 && it.x && it.force && it.out && it.face[0] && it.face[1] && it.face[2],
                     "transfer batch item %d: null pointer", i);
        if (int r = check_tiled(it.grid, bt->tile)) return r;
        const TileSpec ts = make_spec(it.grid, bt->tile);
        if (i == 0) ts0 = ts;
        L.points += it.points;
        L.keys += bin_key_space(it.grid, bt->tile);
        L.tiles += int(ts.ntiles);
        for (int c = 0; c < 3; ++c) {
            const int64_t nf = (it.grid.n[0] + (!it.grid.periodic && c == 0)) *
                               (it.grid.n[1] + (!it.grid.periodic && c == 1)) *
                               (it.grid.n[2] + (!it.grid.periodic && c == 2));
            L.fill_blocks += int(cdiv(nf, FILL_CHUNK));
        }
    }
    AFSI_REQUIRE(L.points < (int64_t(1) << 31) && L.keys < (int64_t(1) << 31), "batch too large");
    L.lds = size_t(ts0.lds) * sizeof(double);
    AFSI_REQUIRE(L.lds <= 64 * 1024, "tile too large for LDS");
    size_t off = 0;
    L.bin_off = off; off = alignb(off + size_t(L.n) * sizeof(BinDesc));
    L.xfer_off = off; off = alignb(off + size_t(L.n) * sizeof(XferDesc));
    L.fill_off = off; off = alignb(off + size_t(3 * L.n) * sizeof(FillDesc));
    L.chunk_off = off; off = alignb(off + size_t(L.tiles + 1) * sizeof(int32_t));
    L.sort_off = off; off = alignb(off + bin_workspace_bytes(L.points > 0 ? L.points : 1));
    L.total = off;
    return OK;
}

template <class D>
static D* wsat(void* ws, size_t off) {
    return reinterpret_cast<D*>(static_cast<char*>(ws) + off);
}

static int xfer_ready(const afsi_tb_batch* bt, XferLayout& L) {
    if (int r = xfer_layout(bt, L)) return r;
    AFSI_REQUIRE(bt->workspace && bt->workspace_bytes >= L.total, "transfer batch workspace too small");
    AFSI_REQUIRE(bt->keys_sorted && bt->perm && bt->cell_ptr && bt->flags, "missing batch buffers");
    return OK;
}

// Fill, work-list scan and spread launch of afsi_spread_tiled_f64;
// chunk_start holds ntiles + 1 ints of scratch (unused without points).
// Zero the three face arrays of one spread in one launch (grid-stride over
// their concatenated ranges).
__global__ void __launch_bounds__(256) zero3_kernel(double* __restrict__ a, int64_t na, double* __restrict__ b,
                                                    int64_t nb, double* __restrict__ c, int64_t nc) {
    const int64_t total = na + nb + nc;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += int64_t(gridDim.x) * blockDim.x) {
        if (i < na) a[i] = 0.0;
        else if (i < na + nb) b[i - na] = 0.0;
        else c[i - na - nb] = 0.0;
    }
}

static int spread_tiled(const double* x, const double* force, int64_t points, const afsi_grid3& g,
                        const int32_t tile[3], const int32_t* perm, const int32_t* cell_ptr, double* u,
                        double* v, double* w, int32_t* chunk_start, int* flags, hipStream_t s) {
    AFSI_REQUIRE(points >= 0 && points < (int64_t(1) << 31), "invalid spread arguments");
    AFSI_REQUIRE(u && v && w, "null face array");
    if (int r = check_tiled(g, tile)) return r;
    const TileSpec ts = make_spec(g, tile);
    AFSI_REQUIRE(size_t(ts.lds) * sizeof(double) <= 64 * 1024, "tile too large for LDS");
    const int64_t nf[3][3] = {
        {g.n[0] + !g.periodic, g.n[1], g.n[2]},
        {g.n[0], g.n[1] + !g.periodic, g.n[2]},
        {g.n[0], g.n[1], g.n[2] + !g.periodic}};
    const int64_t counts[3] = {nf[0][0] * nf[0][1] * nf[0][2], nf[1][0] * nf[1][1] * nf[1][2],
                               nf[2][0] * nf[2][1] * nf[2][2]};
    zero3_kernel<<<grid_for(counts[0] + counts[1] + counts[2], 256, 16384), 256, 0, s>>>(
        u, counts[0], v, counts[1], w, counts[2]);
    AFSI_LAUNCH_CHECK();
    if (points == 0) return OK;
    AFSI_REQUIRE(x && force && perm && cell_ptr && chunk_start, "null pointer");
    chunk_scan_kernel<<<1, 1024, 0, s>>>(cell_ptr, ts.ntiles, ts.tvol, SPREAD_CHUNK, chunk_start);
    AFSI_LAUNCH_CHECK();
    const double volume = (g.h[0] * g.h[1]) * g.h[2];
    // one workgroup per work item up to 16384 (measured at 128^3 with 1.5M
    // clustered points, about 7,000 items: 252 us against 270 us when 4096
    // workgroups stride over the items)
    const int64_t upper = ts.ntiles + points / SPREAD_CHUNK + 1;
    const unsigned blocks = unsigned(upper < 16384 ? upper : 16384);
    const size_t shmem = size_t(ts.lds) * sizeof(double);
#define AFSI_SPREAD(E)                                                                          \
    spread_tiled_kernel<E><<<blocks, SPREAD_THREADS, shmem, s>>>(x, force, perm, cell_ptr, chunk_start, \
                                                                g, ts, volume, u, v, w, flags)
    switch (cubic_edge(tile)) {
    case 8: AFSI_SPREAD(8); break;
    case 4: AFSI_SPREAD(4); break;
    default: AFSI_SPREAD(0);
    }
#undef AFSI_SPREAD
    AFSI_LAUNCH_CHECK();
    return OK;
}

// ---- the spread without allocation ------------------------------------------
// For callers that cannot allocate (code under graph capture, serving
// loops): the spread's only scratch, the per-tile work-list prefix
// chunk_start, lives in a fixed device array of NOALLOC_TILES + 1 ints
// (4 MiB). Each GPU holds its own copy from module load until the process
// ends; its address is looked up once per device and kept. Every call
// writes all entries it reads before reading them, so nothing needs
// clearing; calls that overlap on one GPU share the array and race.

constexpr int64_t NOALLOC_TILES = (int64_t(1) << 20) - 1;
static __device__ __align__(256) int32_t noalloc_chunk_start[NOALLOC_TILES + 1];

static std::mutex noalloc_lock;
static std::vector<int32_t*> noalloc_address;

static int noalloc_scratch(int32_t** out) {
    int dev = 0;
    AFSI_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> hold(noalloc_lock);
    if (size_t(dev) >= noalloc_address.size()) noalloc_address.resize(size_t(dev) + 1, nullptr);
    if (!noalloc_address[dev]) {
        void* p = nullptr;
        AFSI_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_chunk_start)));
        AFSI_REQUIRE(p != nullptr, "no device address for the spread scratch array");
        noalloc_address[dev] = static_cast<int32_t*>(p);
    }
    *out = noalloc_address[dev];
    return OK;
}

}  // namespace afsi

using namespace afsi;

extern "C" {

int afsi_tile_box_bounds(const afsi_grid3* grid, const int32_t tile[3], int64_t lo[3], int64_t hi[3]) {
    AFSI_REQUIRE(grid && tile && lo && hi, "invalid box bounds arguments");
    AFSI_REQUIRE(grid->n[0] > 0 && grid->n[1] > 0 && grid->n[2] > 0, "invalid grid");
    if (int r = check_tiled(*grid, tile)) return r;
    const TileSpec ts = make_spec(*grid, tile);
    for (int c = 0; c < 3; ++c) {
        lo[c] = INT64_MAX;
        hi[c] = -1;
        const int vol = box_dim<0>(ts, c, 0) * box_dim<0>(ts, c, 1) * box_dim<0>(ts, c, 2);
        for (int64_t t = 0; t < ts.ntiles; ++t) {
            const int t0[3] = {int(t / (ts.nt[2] * ts.nt[1])) * ts.t[0], int((t / ts.nt[2]) % ts.nt[1]) * ts.t[1],
                               int(t % ts.nt[2]) * ts.t[2]};
            for (int e = 0; e < vol; ++e) {
                const int f = box_face<0>(e, c, t0, ts, *grid);
                if (f < 0) continue;
                lo[c] = f < lo[c] ? f : lo[c];
                hi[c] = f > hi[c] ? f : hi[c];
            }
        }
    }
    return OK;
}

size_t afsi_spread_workspace_bytes(const afsi_grid3* grid, const int32_t tile[3]) {
    if (!grid || !tile || tile[0] <= 0 || tile[1] <= 0 || tile[2] <= 0) return 0;
    const TileSpec ts = make_spec(*grid, tile);
    return align256(size_t(ts.ntiles + 1) * sizeof(int32_t));
}

int afsi_spread_tiled_f64(const double* x, const double* force, int64_t points,
                          const afsi_grid3* grid, const int32_t tile[3], const int32_t* perm,
                          const int32_t* cell_ptr, double* u, double* v, double* w,
                          void* workspace, size_t workspace_bytes, int* flags, void* stream) {
    AFSI_REQUIRE(grid && tile, "invalid spread arguments");
    AFSI_REQUIRE(points == 0 || workspace_bytes >= afsi_spread_workspace_bytes(grid, tile),
                 "spread workspace too small");
    return spread_tiled(x, force, points, *grid, tile, perm, cell_ptr, u, v, w,
                        static_cast<int32_t*>(workspace), flags, hipStream_t(stream));
}

int afsi_spread_tiled_noalloc_f64(const double* x, const double* force, int64_t points,
                                  const afsi_grid3* grid, const int32_t tile[3], const int32_t* perm,
                                  const int32_t* cell_ptr, double* u, double* v, double* w, int* flags,
                                  void* stream) {
    AFSI_REQUIRE(grid && tile, "invalid spread arguments");
    if (int r = check_tiled(*grid, tile)) return r;
    const TileSpec ts = make_spec(*grid, tile);
    AFSI_REQUIRE(ts.ntiles <= NOALLOC_TILES,
                 "grid of %lld tiles exceeds the no-allocation spread's limit of %lld tiles",
                 (long long)ts.ntiles, (long long)NOALLOC_TILES);
    int32_t* chunk_start = nullptr;
    if (points > 0)
        if (int r = noalloc_scratch(&chunk_start)) return r;
    return spread_tiled(x, force, points, *grid, tile, perm, cell_ptr, u, v, w, chunk_start, flags,
                        hipStream_t(stream));
}

int afsi_interp_tiled_f64(const double* x, int64_t points, const afsi_grid3* grid,
                          const int32_t tile[3], const int32_t* perm, const int32_t* cell_ptr,
                          const double* u, const double* v, const double* w, double* out,
                          void* workspace, size_t workspace_bytes, int* flags, void* stream) {
    AFSI_REQUIRE(grid && points >= 0 && points < (int64_t(1) << 31), "invalid interpolation arguments");
    if (points == 0) return OK;
    AFSI_REQUIRE(x && u && v && w && out, "null pointer");
    const afsi_grid3& g = *grid;
    const hipStream_t s = hipStream_t(stream);
    if (!perm) {
        AFSI_REQUIRE((g.n[0] + 1) * (g.n[1] + 1) * (g.n[2] + 1) < (int64_t(1) << 31),
                     "grid exceeds 2^31 faces per component");
        interp_points_kernel<<<grid_for(points, 256, 16384), 256, 0, s>>>(x, points, g, u, v, w, out, flags);
        AFSI_LAUNCH_CHECK();
        return OK;
    }
    AFSI_REQUIRE(tile && cell_ptr && workspace, "binned interpolation needs tile, cell_ptr, workspace");
    if (int r = check_tiled(g, tile)) return r;
    AFSI_REQUIRE(workspace_bytes >= afsi_spread_workspace_bytes(grid, tile),
                 "interpolation workspace too small");
    const TileSpec ts = make_spec(g, tile);
    AFSI_REQUIRE(size_t(ts.lds) * sizeof(double) <= 64 * 1024, "tile too large for LDS");
    int32_t* chunk_start = static_cast<int32_t*>(workspace);
    chunk_scan_kernel<<<1, 1024, 0, s>>>(cell_ptr, ts.ntiles, ts.tvol, INTERP_CHUNK, chunk_start);
    AFSI_LAUNCH_CHECK();
    const int64_t upper = ts.ntiles + points / INTERP_CHUNK + 1;
    const unsigned blocks = unsigned(upper < 4096 ? upper : 4096);
    const size_t shmem = size_t(ts.lds) * sizeof(double);
    switch (cubic_edge(tile)) {
    case 8:
        interp_lds_kernel<8><<<blocks, INTERP_THREADS, shmem, s>>>(x, perm, cell_ptr, chunk_start, g, ts,
                                                                  u, v, w, out, flags);
        break;
    case 4:
        interp_lds_kernel<4><<<blocks, INTERP_THREADS, shmem, s>>>(x, perm, cell_ptr, chunk_start, g, ts,
                                                                  u, v, w, out, flags);
        break;
    default:
        interp_lds_kernel<0><<<blocks, INTERP_THREADS, shmem, s>>>(x, perm, cell_ptr, chunk_start, g, ts,
                                                                  u, v, w, out, flags);
    }
    AFSI_LAUNCH_CHECK();
    return OK;
}


int afsi_tb_sizes(const afsi_tb_batch* batch, int64_t* total_points, int64_t* key_space,
                  size_t* workspace_bytes) {
    XferLayout L;
    if (int r = xfer_layout(batch, L)) return r;
    if (total_points) *total_points = L.points;
    if (key_space) *key_space = L.keys;
    if (workspace_bytes) *workspace_bytes = L.total;
    return OK;
}

int afsi_tb_prepare(const afsi_tb_batch* batch, void* stream) {
    XferLayout L;
    if (int r = xfer_ready(batch, L)) return r;
    std::vector<char> host(L.sort_off, 0);
    void* hw = host.data();
    int64_t pbeg = 0, kbeg = 0;
    int tbeg = 0, fbeg = 0;
    for (int i = 0; i < L.n; ++i) {
        const afsi_tb_item& it = batch->items[i];
        const TileSpec ts = make_spec(it.grid, batch->tile);
        BinDesc b{};
        b.x = it.x;
        b.flags = batch->flags;
        b.g = it.grid;
        b.tg = bin_tile_geom(it.grid, batch->tile);
        b.points = it.points;
        b.begin = pbeg;
        b.key_offset = kbeg;
        wsat<BinDesc>(hw, L.bin_off)[i] = b;
        XferDesc x{};
        x.x = it.x;
        x.force = it.force;
        for (int c = 0; c < 3; ++c) x.face[c] = it.face[c];
        x.out = it.out;
        x.g = it.grid;
        x.ts = ts;
        x.volume = (it.grid.h[0] * it.grid.h[1]) * it.grid.h[2];
        x.point_begin = int(pbeg);
        x.key_offset = int(kbeg);
        x.begin = tbeg;
        wsat<XferDesc>(hw, L.xfer_off)[i] = x;
        for (int c = 0; c < 3; ++c) {
            const int64_t nf = (it.grid.n[0] + (!it.grid.periodic && c == 0)) *
                               (it.grid.n[1] + (!it.grid.periodic && c == 1)) *
                               (it.grid.n[2] + (!it.grid.periodic && c == 2));
            wsat<FillDesc>(hw, L.fill_off)[3 * i + c] = FillDesc{it.face[c], int(nf), fbeg};
            fbeg += int(cdiv(nf, FILL_CHUNK));
        }
        pbeg += it.points;
        kbeg += bin_key_space(it.grid, batch->tile);
        tbeg += int(ts.ntiles);
    }
    const hipStream_t s = hipStream_t(stream);
    AFSI_HIP(hipMemcpyAsync(batch->workspace, host.data(), host.size(), hipMemcpyHostToDevice, s));
    AFSI_HIP(hipStreamSynchronize(s));
    return OK;
}

int afsi_tb_bin(const afsi_tb_batch* batch, void* stream) {
    XferLayout L;
    if (int r = xfer_ready(batch, L)) return r;
    return bin_batch(wsat<BinDesc>(batch->workspace, L.bin_off), L.n, L.points, L.keys, batch->keys_sorted,
                     batch->perm, batch->cell_ptr, wsat<char>(batch->workspace, L.sort_off),
                     hipStream_t(stream));
}

static int tb_transfer(const afsi_tb_batch* batch, bool spread, void* stream) {
    XferLayout L;
    if (int r = xfer_ready(batch, L)) return r;
    const hipStream_t s = hipStream_t(stream);
    const XferDesc* desc = wsat<XferDesc>(batch->workspace, L.xfer_off);
    int32_t* chunk_start = wsat<int32_t>(batch->workspace, L.chunk_off);
    if (spread) {
        fill_batch_kernel<<<unsigned(L.fill_blocks), 256, 0, s>>>(wsat<FillDesc>(batch->workspace, L.fill_off),
                                                                 3 * L.n);
        AFSI_LAUNCH_CHECK();
    }
    if (L.points == 0) return OK;
    const int chunk = spread ? SPREAD_CHUNK : INTERP_CHUNK;
    chunk_scan_batch_kernel<<<1, 1024, 0, s>>>(desc, L.n, L.tiles, batch->cell_ptr, chunk, chunk_start);
    AFSI_LAUNCH_CHECK();
    const int64_t upper = L.tiles + L.points / chunk + 1;
    const unsigned blocks = unsigned(upper < 4096 ? upper : 4096);
    const int edge = cubic_edge(batch->tile);
#define AFSI_TB_ARGS(...) __VA_ARGS__
#define AFSI_TB_LAUNCH(KERNEL, E)                                                                      \
    KERNEL<AFSI_TB_ARGS E><<<blocks, SPREAD_THREADS, L.lds, s>>>(desc, L.n, L.tiles, batch->perm,       \
                                                                 batch->cell_ptr, chunk_start, batch->flags)
    if (spread) {
        if (edge == 8) AFSI_TB_LAUNCH(spread_batch_kernel, (8));
        else if (edge == 4) AFSI_TB_LAUNCH(spread_batch_kernel, (4));
        else AFSI_TB_LAUNCH(spread_batch_kernel, (0));
    } else {
        if (edge == 8) AFSI_TB_LAUNCH(interp_batch_kernel, (8));
        else if (edge == 4) AFSI_TB_LAUNCH(interp_batch_kernel, (4));
        else AFSI_TB_LAUNCH(interp_batch_kernel, (0));
    }
#undef AFSI_TB_LAUNCH
#undef AFSI_TB_ARGS
    AFSI_LAUNCH_CHECK();
    return OK;
}

int afsi_tb_spread(const afsi_tb_batch* batch, void* stream) { return tb_transfer(batch, true, stream); }
int afsi_tb_interp(const afsi_tb_batch* batch, void* stream) { return tb_transfer(batch, false, stream); }

}  // extern "C"
