// Tiled Peskin-4 spread and interpolation over the binned point CSR.
//
// Spread: the work list is (tile, chunk of <= SPREAD_CHUNK points); a
// persistent grid walks it. A work item accumulates the contributions of its
// points for all three components into an LDS box covering every face the
// tile's points can touch (tile + 2 cells on each side; one less on the
// component's own axis), with ds_add_f64, then adds each nonzero LDS entry to
// HBM with one global atomic. Dense tiles therefore collapse up to thousands
// of point contributions into one global update per face, and a point is
// read and its weights evaluated exactly once.
//
// Interp: one thread per binned point (sorted order, so neighbouring lanes
// read neighbouring faces through L1/L2), 64 links per component in the
// link order of the reference's compact kernel, result scattered back to the
// caller's point order.
//
// Weights: per axis f = s - base - 1 and the four Peskin weights from the
// shared radical sqrt(1+4f-4f^2); base = floor(s-1) exactly as the
// reference (mac/compact_transfer.py:161-166) with s=(x-o)/h.
#include "common.hpp"

namespace afsi {

template <class T>
int launch_fill(T* x, int64_t count, T value, hipStream_t s);

constexpr int SPREAD_THREADS = 256;
constexpr int SPREAD_CHUNK = 2048;

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
__device__ inline int box_dim(const TileSpec& ts, int c, int a) {
    return T > 0 ? T + 4 - (a == c ? 1 : 0) : ts.dim[c][a];
}
template <int T>
__device__ inline int box_off(const TileSpec& ts, int c) {
    return T > 0 ? c * (T + 3) * (T + 4) * (T + 4) : ts.off[c];
}
template <int T>
__device__ inline int tile_edge(const TileSpec& ts, int a) {
    return T > 0 ? T : ts.t[a];
}

__device__ inline int wrap32(int i, int n) {
    const int r = i % n;
    return r < 0 ? r + n : r;
}

struct PointWeights {
    int bf[3], bc[3];  // face-lattice and centre-lattice bases
    double wf[3][4], wc[3][4];
};

// Bases and weights of one point (mac/compact_transfer.py:161-166 bases).
__device__ inline void point_weights(const double xp[3], const afsi_grid3& g, PointWeights& pw) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double sf = (xp[a] - g.origin[a]) / g.h[a];
        const double ff = floor(sf - 1.0);
        const double sc = (xp[a] - g.center_origin[a]) / g.h[a];
        const double fc = floor(sc - 1.0);
        pw.bf[a] = (ff > -1e9 && ff < 1e9) ? int(ff) : 0;
        pw.bc[a] = (fc > -1e9 && fc < 1e9) ? int(fc) : 0;
        peskin_weights<double>(sf - (ff + 1.0), pw.wf[a]);
        peskin_weights<double>(sc - (fc + 1.0), pw.wc[a]);
    }
}

__global__ void chunk_scan_kernel(const int32_t* __restrict__ cell_ptr, int64_t ntiles,
                                  int64_t tvol, int32_t* __restrict__ chunk_start) {
    // Single workgroup: exclusive scan of ceil(points_in_tile / CHUNK).
    __shared__ int32_t s[1024];
    int32_t running = 0;
    for (int64_t base = 0; base < ntiles; base += 1024) {
        const int64_t t = base + threadIdx.x;
        int32_t v = 0;
        if (t < ntiles) {
            const int32_t n = cell_ptr[(t + 1) * tvol] - cell_ptr[t * tvol];
            v = (n + SPREAD_CHUNK - 1) / SPREAD_CHUNK;
        }
        s[threadIdx.x] = v;
        __syncthreads();
        for (int d = 1; d < 1024; d <<= 1) {
            const int32_t add = threadIdx.x >= d ? s[threadIdx.x - d] : 0;
            __syncthreads();
            s[threadIdx.x] += add;
            __syncthreads();
        }
        if (t < ntiles) chunk_start[t] = running + s[threadIdx.x] - v;
        running += s[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) chunk_start[ntiles] = running;
}

// Work item w -> tile (last t with chunk_start[t] <= w), its first corner
// cell t0 and the point range [pbeg, pend) of its chunk.
struct WorkItem {
    int t0[3];
    int pbeg, pend;
};

__device__ inline WorkItem work_item(int w, const int32_t* __restrict__ chunk_start,
                                     const int32_t* __restrict__ cell_ptr, const TileSpec& ts) {
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
    it.pbeg = cell_ptr[tile * tv] + chunk * SPREAD_CHUNK;
    it.pend = min(cell_ptr[(tile + 1) * tv], it.pbeg + SPREAD_CHUNK);
    return it;
}

// Box entry e of component c -> global face offset, or -1 outside a closed
// box. Box axis a starts at t0[a] - 1 (own axis) or t0[a] - 2.
template <int T>
__device__ inline int box_face(int e, int c, const int t0[3], const TileSpec& ts,
                               const afsi_grid3& g) {
    const int d1 = box_dim<T>(ts, c, 1), d2 = box_dim<T>(ts, c, 2);
    const int l[3] = {e / (d1 * d2), (e / d2) % d1, e % d2};
    int idx = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int nf = int(g.n[a]) + ((!g.periodic && a == c) ? 1 : 0);
        int q = t0[a] - (a == c ? 1 : 2) + l[a];
        if (g.periodic) q = q < 0 ? q + nf : (q >= nf ? q - nf : q);
        else if (q < 0 || q >= nf) return -1;
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
        const int cell = pw.bf[a] + 1;
        shift[a] = g.periodic ? wrap32(cell, int(g.n[a])) - cell : 0;
    }
}

template <int T>
__global__ void __launch_bounds__(SPREAD_THREADS) spread_tiled_kernel(
    const double* __restrict__ x, const double* __restrict__ force,
    const int32_t* __restrict__ perm, const int32_t* __restrict__ cell_ptr,
    const int32_t* __restrict__ chunk_start, afsi_grid3 g, TileSpec ts, double volume,
    double* __restrict__ out0, double* __restrict__ out1, double* __restrict__ out2, int* flags) {
    extern __shared__ double lds[];
    const int total = chunk_start[ts.ntiles];
    double* outs[3] = {out0, out1, out2};
    const int nlds = T > 0 ? 3 * (T + 3) * (T + 4) * (T + 4) : ts.lds;
    for (int w = blockIdx.x; w < total; w += gridDim.x) {
        const WorkItem it = work_item(w, chunk_start, cell_ptr, ts);
        for (int e = threadIdx.x; e < nlds; e += SPREAD_THREADS) lds[e] = 0.0;
        __syncthreads();
        // Each thread takes a contiguous slice so lanes of one wave sit in
        // different cells and LDS atomics rarely collide.
        const int n = it.pend - it.pbeg;
        const int per = (n + SPREAD_THREADS - 1) / SPREAD_THREADS;
        const int j0 = it.pbeg + int(threadIdx.x) * per;
        const int j1 = min(it.pend, j0 + per);
        for (int j = j0; j < j1; ++j) {
            const int idx = perm[j];
            const double xp[3] = {x[3 * idx], x[3 * idx + 1], x[3 * idx + 2]};
            PointWeights pw;
            point_weights(xp, g, pw);
            int shift[3];
            frame_shift(pw, g, shift);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double fv = force[3 * idx + c] / volume;
                int lb[3];
                if (!box_base<T>(pw, shift, c, it.t0, ts, lb)) {
                    raise_flag(flags, F_INDEX);
                    continue;
                }
                const double* wx = (c == 0) ? pw.wf[0] : pw.wc[0];
                const double* wy = (c == 1) ? pw.wf[1] : pw.wc[1];
                const double* wz = (c == 2) ? pw.wf[2] : pw.wc[2];
                const int d1 = box_dim<T>(ts, c, 1), d2 = box_dim<T>(ts, c, 2);
                double* box = lds + box_off<T>(ts, c);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
#pragma unroll
                    for (int jj = 0; jj < 4; ++jj) {
                        const double wxy = wx[i] * wy[jj];
                        double* row = box + ((lb[0] + i) * d1 + (lb[1] + jj)) * d2 + lb[2];
#pragma unroll
                        for (int k = 0; k < 4; ++k) atomicAdd(row + k, (wxy * wz[k]) * fv);
                    }
                }
            }
        }
        __syncthreads();
        // Flush: one global atomic per nonzero LDS entry.
        for (int e = threadIdx.x; e < nlds; e += SPREAD_THREADS) {
            const double v = lds[e];
            if (v == 0.0) continue;
            const int c = e < box_off<T>(ts, 1) ? 0 : (e < box_off<T>(ts, 2) ? 1 : 2);
            const int f = box_face<T>(e - box_off<T>(ts, c), c, it.t0, ts, g);
            if (f < 0) {
                raise_flag(flags, F_INDEX);
                continue;
            }
            unsafeAtomicAdd(outs[c] + f, v);
        }
        __syncthreads();
    }
}

// Interpolation over the same work list: the tile's face box of all three
// components is staged in LDS once per work item, then each point gathers
// its 3 x 64 values from LDS (consecutive threads take consecutive binned
// points; results go to the caller's point order).
template <int T>
__global__ void __launch_bounds__(SPREAD_THREADS) interp_lds_kernel(
    const double* __restrict__ x, const int32_t* __restrict__ perm,
    const int32_t* __restrict__ cell_ptr, const int32_t* __restrict__ chunk_start, afsi_grid3 g,
    TileSpec ts, const double* __restrict__ u0, const double* __restrict__ u1,
    const double* __restrict__ u2, double* __restrict__ out, int* flags) {
    extern __shared__ double lds[];
    const int total = chunk_start[ts.ntiles];
    const double* us[3] = {u0, u1, u2};
    for (int w = blockIdx.x; w < total; w += gridDim.x) {
        const WorkItem it = work_item(w, chunk_start, cell_ptr, ts);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int vol = box_dim<T>(ts, c, 0) * box_dim<T>(ts, c, 1) * box_dim<T>(ts, c, 2);
            double* box = lds + box_off<T>(ts, c);
            for (int e = threadIdx.x; e < vol; e += SPREAD_THREADS) {
                const int f = box_face<T>(e, c, it.t0, ts, g);
                box[e] = f >= 0 ? us[c][f] : 0.0;
            }
        }
        __syncthreads();
        for (int j = it.pbeg + int(threadIdx.x); j < it.pend; j += SPREAD_THREADS) {
            const int idx = perm[j];
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
                    const double* box = lds + box_off<T>(ts, c);
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
#pragma unroll
                        for (int jj = 0; jj < 4; ++jj) {
                            const double wxy = wx[i] * wy[jj];
                            const double* row = box + ((lb[0] + i) * d1 + (lb[1] + jj)) * d2 + lb[2];
#pragma unroll
                            for (int k = 0; k < 4; ++k) sum += row[k] * (wxy * wz[k]);
                        }
                    }
                }
                out[3 * idx + c] = sum;
            }
        }
        __syncthreads();
    }
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
#pragma unroll
                    for (int jj = 0; jj < 4; ++jj) {
                        const int yj = g.periodic ? wrap32(b[1] + jj, nf[1]) : b[1] + jj;
                        const double wxy = wx[i] * wy[jj];
                        const double* row = u + (xi * nf[1] + yj) * nf[2];
#pragma unroll
                        for (int k = 0; k < 4; ++k) sum += row[zi[k]] * (wxy * wz[k]);
                    }
                }
            }
            out[3 * j + c] = sum;
        }
    }
}

static size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace afsi

using namespace afsi;

extern "C" {

size_t afsi_spread_workspace_bytes(const afsi_grid3* grid, const int32_t tile[3]) {
    if (!grid || !tile || tile[0] <= 0 || tile[1] <= 0 || tile[2] <= 0) return 0;
    const TileSpec ts = make_spec(*grid, tile);
    return align256(size_t(ts.ntiles + 1) * sizeof(int32_t));
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

int afsi_spread_tiled_f64(const double* x, const double* force, int64_t points,
                          const afsi_grid3* grid, const int32_t tile[3], const int32_t* perm,
                          const int32_t* cell_ptr, double* u, double* v, double* w,
                          void* workspace, size_t workspace_bytes, int* flags, void* stream) {
    AFSI_REQUIRE(grid && tile && points >= 0 && points < (int64_t(1) << 31), "invalid spread arguments");
    AFSI_REQUIRE(u && v && w, "null face array");
    const afsi_grid3& g = *grid;
    if (int r = check_tiled(g, tile)) return r;
    const TileSpec ts = make_spec(g, tile);
    AFSI_REQUIRE(size_t(ts.lds) * sizeof(double) <= 64 * 1024, "tile too large for LDS");
    const hipStream_t s = hipStream_t(stream);
    const int64_t nf[3][3] = {
        {g.n[0] + !g.periodic, g.n[1], g.n[2]},
        {g.n[0], g.n[1] + !g.periodic, g.n[2]},
        {g.n[0], g.n[1], g.n[2] + !g.periodic}};
    double* outs[3] = {u, v, w};
    for (int c = 0; c < 3; ++c)
        if (int r = launch_fill<double>(outs[c], nf[c][0] * nf[c][1] * nf[c][2], 0.0, s)) return r;
    if (points == 0) return OK;
    AFSI_REQUIRE(x && force && perm && cell_ptr && workspace, "null pointer");
    AFSI_REQUIRE(workspace_bytes >= afsi_spread_workspace_bytes(grid, tile),
                 "spread workspace too small");
    int32_t* chunk_start = static_cast<int32_t*>(workspace);
    chunk_scan_kernel<<<1, 1024, 0, s>>>(cell_ptr, ts.ntiles, ts.tvol, chunk_start);
    AFSI_LAUNCH_CHECK();
    const double volume = (g.h[0] * g.h[1]) * g.h[2];
    const int64_t upper = ts.ntiles + points / SPREAD_CHUNK + 1;
    const unsigned blocks = unsigned(upper < 4096 ? upper : 4096);
    const size_t shmem = size_t(ts.lds) * sizeof(double);
    switch (cubic_edge(tile)) {
    case 8:
        spread_tiled_kernel<8><<<blocks, SPREAD_THREADS, shmem, s>>>(x, force, perm, cell_ptr, chunk_start,
                                                                    g, ts, volume, u, v, w, flags);
        break;
    case 4:
        spread_tiled_kernel<4><<<blocks, SPREAD_THREADS, shmem, s>>>(x, force, perm, cell_ptr, chunk_start,
                                                                    g, ts, volume, u, v, w, flags);
        break;
    default:
        spread_tiled_kernel<0><<<blocks, SPREAD_THREADS, shmem, s>>>(x, force, perm, cell_ptr, chunk_start,
                                                                    g, ts, volume, u, v, w, flags);
    }
    AFSI_LAUNCH_CHECK();
    return OK;
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
    chunk_scan_kernel<<<1, 1024, 0, s>>>(cell_ptr, ts.ntiles, ts.tvol, chunk_start);
    AFSI_LAUNCH_CHECK();
    const int64_t upper = ts.ntiles + points / SPREAD_CHUNK + 1;
    const unsigned blocks = unsigned(upper < 4096 ? upper : 4096);
    const size_t shmem = size_t(ts.lds) * sizeof(double);
    switch (cubic_edge(tile)) {
    case 8:
        interp_lds_kernel<8><<<blocks, SPREAD_THREADS, shmem, s>>>(x, perm, cell_ptr, chunk_start, g, ts,
                                                                  u, v, w, out, flags);
        break;
    case 4:
        interp_lds_kernel<4><<<blocks, SPREAD_THREADS, shmem, s>>>(x, perm, cell_ptr, chunk_start, g, ts,
                                                                  u, v, w, out, flags);
        break;
    default:
        interp_lds_kernel<0><<<blocks, SPREAD_THREADS, shmem, s>>>(x, perm, cell_ptr, chunk_start, g, ts,
                                                                  u, v, w, out, flags);
    }
    AFSI_LAUNCH_CHECK();
    return OK;
}

}  // extern "C"
