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

struct PointWeights {
    int64_t bf[3], bc[3];  // face-lattice and centre-lattice bases
    double wf[3][4], wc[3][4];
};

// Bases and weights of one point; `shift` moves unwrapped bases into the
// frame of the (wrapped) cell used for binning.
__device__ inline void point_weights(const double xp[3], const afsi_grid3& g,
                                     PointWeights& pw) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double sf = (xp[a] - g.origin[a]) / g.h[a];
        const double ff = floor(sf - 1.0);
        const double sc = (xp[a] - g.center_origin[a]) / g.h[a];
        const double fc = floor(sc - 1.0);
        pw.bf[a] = isfinite(ff) ? int64_t(ff) : 0;
        pw.bc[a] = isfinite(fc) ? int64_t(fc) : 0;
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

__global__ void __launch_bounds__(SPREAD_THREADS) spread_tiled_kernel(
    const double* __restrict__ x, const double* __restrict__ force,
    const int32_t* __restrict__ perm, const int32_t* __restrict__ cell_ptr,
    const int32_t* __restrict__ chunk_start, afsi_grid3 g, TileSpec ts, double volume,
    double* __restrict__ out0, double* __restrict__ out1, double* __restrict__ out2, int* flags) {
    extern __shared__ double lds[];
    const int32_t total = chunk_start[ts.ntiles];
    double* outs[3] = {out0, out1, out2};
    for (int32_t w = blockIdx.x; w < total; w += gridDim.x) {
        // tile = last t with chunk_start[t] <= w
        int64_t lo = 0, hi = ts.ntiles;
        while (hi - lo > 1) {
            const int64_t mid = (lo + hi) >> 1;
            if (chunk_start[mid] <= w) lo = mid;
            else hi = mid;
        }
        const int64_t tile = lo;
        const int32_t chunk = w - chunk_start[tile];
        const int64_t tz = tile % ts.nt[2], ty = (tile / ts.nt[2]) % ts.nt[1],
                      tx = tile / (ts.nt[2] * ts.nt[1]);
        const int64_t t0[3] = {tx * ts.t[0], ty * ts.t[1], tz * ts.t[2]};
        const int32_t pbeg = cell_ptr[tile * ts.tvol] + chunk * SPREAD_CHUNK;
        const int32_t pend = min(cell_ptr[(tile + 1) * ts.tvol], pbeg + SPREAD_CHUNK);
        for (int32_t e = threadIdx.x; e < ts.lds; e += SPREAD_THREADS) lds[e] = 0.0;
        __syncthreads();
        // Each thread takes a contiguous slice so lanes of one wave sit in
        // different cells and LDS atomics rarely collide.
        const int32_t n = pend - pbeg;
        const int32_t per = (n + SPREAD_THREADS - 1) / SPREAD_THREADS;
        const int32_t j0 = pbeg + threadIdx.x * per;
        const int32_t j1 = min(pend, j0 + per);
        for (int32_t j = j0; j < j1; ++j) {
            const int64_t idx = perm[j];
            const double xp[3] = {x[3 * idx], x[3 * idx + 1], x[3 * idx + 2]};
            PointWeights pw;
            point_weights(xp, g, pw);
            // Shift unwrapped bases into this tile's frame (periodic wrap).
            int64_t shift[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int64_t cell = pw.bf[a] + 1;
                shift[a] = g.periodic ? wrap(cell, g.n[a]) - cell : 0;
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double fv = force[3 * idx + c] / volume;
                int64_t lb[3];
                bool ok = true;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const int64_t b = (a == c ? pw.bf[a] : pw.bc[a]) + shift[a];
                    lb[a] = b - (t0[a] - (a == c ? 1 : 2));
                    ok = ok && lb[a] >= 0 && lb[a] + 3 < ts.dim[c][a];
                }
                if (!ok) {
                    raise_flag(flags, F_INDEX);
                    continue;
                }
                const double* wx = (c == 0) ? pw.wf[0] : pw.wc[0];
                const double* wy = (c == 1) ? pw.wf[1] : pw.wc[1];
                const double* wz = (c == 2) ? pw.wf[2] : pw.wc[2];
                const int32_t d1 = ts.dim[c][1], d2 = ts.dim[c][2];
                double* box = lds + ts.off[c];
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
        for (int32_t e = threadIdx.x; e < ts.lds; e += SPREAD_THREADS) {
            const double v = lds[e];
            if (v == 0.0) continue;
            const int c = e < ts.off[1] ? 0 : (e < ts.off[2] ? 1 : 2);
            const int32_t r = e - ts.off[c];
            const int32_t d1 = ts.dim[c][1], d2 = ts.dim[c][2];
            const int64_t l[3] = {r / (d1 * d2), (r / d2) % d1, r % d2};
            int64_t gidx[3], nf[3];
            bool ok = true;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                nf[a] = g.n[a] + ((!g.periodic && a == c) ? 1 : 0);
                int64_t q = t0[a] - (a == c ? 1 : 2) + l[a];
                if (g.periodic) q = wrap(q, g.n[a]);
                ok = ok && q >= 0 && q < nf[a];
                gidx[a] = q;
            }
            if (!ok) {
                raise_flag(flags, F_INDEX);
                continue;
            }
            unsafeAtomicAdd(outs[c] + (gidx[0] * nf[1] + gidx[1]) * nf[2] + gidx[2], v);
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) interp_tiled_kernel(
    const double* __restrict__ x, int64_t points, const int32_t* __restrict__ perm,
    afsi_grid3 g, const double* __restrict__ u0, const double* __restrict__ u1,
    const double* __restrict__ u2, double* __restrict__ out, int* flags) {
    const double* us[3] = {u0, u1, u2};
    for (int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < points;
         j += int64_t(gridDim.x) * blockDim.x) {
        const int64_t idx = perm ? perm[j] : j;
        const double xp[3] = {x[3 * idx], x[3 * idx + 1], x[3 * idx + 2]};
        PointWeights pw;
        point_weights(xp, g, pw);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            int64_t nf[3], b[3];
            bool ok = true;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                nf[a] = g.n[a] + ((!g.periodic && a == c) ? 1 : 0);
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
                int64_t zi[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) zi[k] = g.periodic ? wrap(b[2] + k, nf[2]) : b[2] + k;
                const double* u = us[c];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int64_t xi = g.periodic ? wrap(b[0] + i, nf[0]) : b[0] + i;
#pragma unroll
                    for (int jj = 0; jj < 4; ++jj) {
                        const int64_t yj = g.periodic ? wrap(b[1] + jj, nf[1]) : b[1] + jj;
                        const double wxy = wx[i] * wy[jj];
                        const double* row = u + (xi * nf[1] + yj) * nf[2];
#pragma unroll
                        for (int k = 0; k < 4; ++k) sum += row[zi[k]] * (wxy * wz[k]);
                    }
                }
            }
            out[3 * idx + c] = sum;
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

int afsi_spread_tiled_f64(const double* x, const double* force, int64_t points,
                          const afsi_grid3* grid, const int32_t tile[3], const int32_t* perm,
                          const int32_t* cell_ptr, double* u, double* v, double* w,
                          void* workspace, size_t workspace_bytes, int* flags, void* stream) {
    AFSI_REQUIRE(grid && tile && points >= 0, "invalid spread arguments");
    AFSI_REQUIRE(tile[0] > 0 && tile[1] > 0 && tile[2] > 0, "tile sizes must be positive");
    AFSI_REQUIRE(u && v && w, "null face array");
    const afsi_grid3& g = *grid;
    const TileSpec ts = make_spec(g, tile);
    AFSI_REQUIRE(size_t(ts.lds) * sizeof(double) <= 160 * 1024, "tile too large for LDS");
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
    spread_tiled_kernel<<<blocks, SPREAD_THREADS, shmem, s>>>(x, force, perm, cell_ptr, chunk_start,
                                                             g, ts, volume, u, v, w, flags);
    AFSI_LAUNCH_CHECK();
    return OK;
}

int afsi_interp_tiled_f64(const double* x, int64_t points, const afsi_grid3* grid,
                          const int32_t* perm, const double* u, const double* v, const double* w,
                          double* out, int* flags, void* stream) {
    AFSI_REQUIRE(grid && points >= 0, "invalid interpolation arguments");
    if (points == 0) return OK;
    AFSI_REQUIRE(x && u && v && w && out, "null pointer");
    interp_tiled_kernel<<<grid_for(points, 256, 16384), 256, 0, hipStream_t(stream)>>>(
        x, points, perm, *grid, u, v, w, out, flags);
    AFSI_LAUNCH_CHECK();
    return OK;
}

}  // extern "C"
