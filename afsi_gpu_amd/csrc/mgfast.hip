// Fused V-cycle for gfx950: few launches per cycle, one HBM pass per sweep.
//
// * rb_sweep_kernel: one full red-black Gauss-Seidel sweep per launch. A
//   workgroup stages its 8x8x32 tile plus a 2-cell halo of p in LDS,
//   updates the red cells of the tile and its 1-cell ring in LDS, then the
//   black cells of the tile, and writes the tile (out of place: p_in is
//   never written, so every tile reads start-of-sweep values). Update
//   p = (rhs + sum_nb w_a p_nb) / sum w_a, w_a = 1/h_a^2, over the existing
//   neighbours, which equals p + (rhs - A p)/diag.
// * Mean removal. The reference subtracts the mean of p after each
//   smoothing stage and of every restricted residual (mac/multigrid.py:
//   107-123). The V-cycle commutes with adding a constant (A 1 = 0, both
//   smoothers and the prolongation reproduce constants, and the coarse
//   inverse of a zero-mean residual has zero mean), so on the fine levels
//   those subtractions change nothing but rounding; this driver performs
//   them inside the LDS coarse kernel only; the caller removes the mean of
//   the level-0 iterate once, after the last cycle. No kernel needs a
//   grid-wide reduction, fence or atomic.
// * coarse_cycle_kernel: one 1024-thread workgroup runs the complete
//   V-cycle of every level with <= COARSE_MAX cells (smoothing, restriction,
//   means, the dense coarse solve, prolongation) out of LDS.
// * Tiles are mapped XCD-aware: consecutive workgroup ids are dealt
//   round-robin to the 8 XCDs, so each XCD is given one contiguous run of
//   tiles and the halos it re-reads stay in its own L2.
#include "mgcommon.hpp"
#include <algorithm>
#include <mutex>
#include <vector>

namespace afsi {

template <class T>
int launch_fill(T* x, int64_t count, T value, hipStream_t s);
int launch_reduce_final(const double* partials, int64_t nparts, double* result, hipStream_t s);

constexpr int COARSE_THREADS = 1024;
constexpr int64_t COARSE_MAX = 4096;  // cells of the finest level solved in LDS
constexpr int COARSE_LIST = 4800;     // packed cell descriptors of all LDS levels

struct Dim32 {
    int n0, n1, n2;
};

template <class T>
struct W3 {
    T w[3];          // 1/h^2
    T relax = T(1);  // red-black over-relaxation weight (1: Gauss-Seidel)
};

// Red-black update of a cell from its Gauss-Seidel value gs: gs itself, or
// old + relax * (gs - old) when over-relaxed (relax is uniform per launch).
template <class T>
__device__ inline T rb_relax(const W3<T>& k, T old, T gs) {
    return k.relax == T(1) ? gs : old + k.relax * (gs - old);
}

__device__ inline int wrap32(int g, int n) {
    while (g < 0) g += n;
    while (g >= n) g -= n;
    return g;
}

__device__ inline int64_t xcd_tile(int64_t b, int64_t ntiles) {
    if (ntiles % 8) return b;
    return (b % 8) * (ntiles / 8) + b / 8;
}

// Gauss-Seidel / Jacobi pieces on an array with strides (sx, sy, 1).
// s = sum_nb w_a p_nb and d = sum w_a over the neighbours that exist.
template <class T, bool P>
__device__ inline void neighbour_sums(const T* __restrict__ a, int li, int sx, int sy, int gx,
                                      int gy, int gz, const Dim32& d, const W3<T>& k, T& s, T& dg) {
    s = T(0);
    dg = T(0);
    if (P || gx + 1 < d.n0) { s += k.w[0] * a[li + sx]; dg += k.w[0]; }
    if (P || gx > 0) { s += k.w[0] * a[li - sx]; dg += k.w[0]; }
    if (P || gy + 1 < d.n1) { s += k.w[1] * a[li + sy]; dg += k.w[1]; }
    if (P || gy > 0) { s += k.w[1] * a[li - sy]; dg += k.w[1]; }
    if (P || gz + 1 < d.n2) { s += k.w[2] * a[li + 1]; dg += k.w[2]; }
    if (P || gz > 0) { s += k.w[2] * a[li - 1]; dg += k.w[2]; }
}

// A p = sum_nb w_a (p - p_nb) over existing neighbours (difference form:
// neighbour differences of a smooth field are nearly exact, so the residual
// keeps its accuracy when p carries a constant offset).
template <class T, bool P>
__device__ inline T apply_diff(const T* __restrict__ a, int li, int sx, int sy,
                               int gx, int gy, int gz, const Dim32& d, const W3<T>& k) {
    const T c = a[li];
    T r = T(0);
    if (P) {
        r += k.w[0] * (c - a[li + (gx + 1 < d.n0 ? sx : -(d.n0 - 1) * sx)]);
        r += k.w[0] * (c - a[li + (gx > 0 ? -sx : (d.n0 - 1) * sx)]);
        r += k.w[1] * (c - a[li + (gy + 1 < d.n1 ? sy : -(d.n1 - 1) * sy)]);
        r += k.w[1] * (c - a[li + (gy > 0 ? -sy : (d.n1 - 1) * sy)]);
        r += k.w[2] * (c - a[li + (gz + 1 < d.n2 ? 1 : -(d.n2 - 1))]);
        r += k.w[2] * (c - a[li + (gz > 0 ? -1 : (d.n2 - 1))]);
    } else {
        if (gx + 1 < d.n0) r += k.w[0] * (c - a[li + sx]);
        if (gx > 0) r += k.w[0] * (c - a[li - sx]);
        if (gy + 1 < d.n1) r += k.w[1] * (c - a[li + sy]);
        if (gy > 0) r += k.w[1] * (c - a[li - sy]);
        if (gz + 1 < d.n2) r += k.w[2] * (c - a[li + 1]);
        if (gz > 0) r += k.w[2] * (c - a[li - 1]);
    }
    return r;
}

// ---------------------------------------------------------------------------
// fine levels

// Tile TX x TY x TZ cells, NT threads. All global loads of a thread (its
// share of the p box and the rhs of its red-ring and black-tile cells) are
// issued into registers before the first LDS write, so their latencies
// overlap instead of serialising one loop iteration at a time.
template <class T, bool P, int TX, int TY, int TZ, int NT>
__global__ void __launch_bounds__(NT) rb_sweep_kernel(const T* __restrict__ pin,
                                                      const T* __restrict__ rhs,
                                                      T* __restrict__ pout, Dim32 d, W3<T> k,
                                                      int zero_in, int ntx, int nty, int ntz) {
    constexpr int BX = TX + 4, BY = TY + 4, BZ = TZ + 4, BOX = BX * BY * BZ;
    constexpr int RX = TX + 2, RY = TY + 2, RZ = TZ + 2, RING = RX * RY * RZ;
    constexpr int TILE = TX * TY * TZ;
    constexpr int NB = (BOX + NT - 1) / NT, NR = (RING + NT - 1) / NT, NL = (TILE + NT - 1) / NT;
    __shared__ T box[BOX];
    const int tile = int(xcd_tile(blockIdx.x, int64_t(ntx) * nty * ntz));
    const int tz = tile % ntz, ty = (tile / ntz) % nty, tx = tile / (ntz * nty);
    const int x0 = tx * TX - 2, y0 = ty * TY - 2, z0 = tz * TZ - 2;
    const int sxg = d.n1 * d.n2;
    const int t = threadIdx.x;
    T pv[NB], rr[NR], rt[NL];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        const int e = t + j * NT;
        pv[j] = T(0);
        if (e < BOX && !zero_in) {
            const int lz = e % BZ, ly = (e / BZ) % BY, lx = e / (BZ * BY);
            int gx = x0 + lx, gy = y0 + ly, gz = z0 + lz;
            if (P) {
                gx = wrap32(gx, d.n0); gy = wrap32(gy, d.n1); gz = wrap32(gz, d.n2);
                pv[j] = pin[gx * sxg + gy * d.n2 + gz];
            } else if (gx >= 0 && gx < d.n0 && gy >= 0 && gy < d.n1 && gz >= 0 && gz < d.n2) {
                pv[j] = pin[gx * sxg + gy * d.n2 + gz];
            }
        }
    }
#pragma unroll
    for (int j = 0; j < NR; ++j) {
        const int e = t + j * NT;
        rr[j] = T(0);
        if (e < RING) {
            const int lz = e % RZ + 1, ly = (e / RZ) % RY + 1, lx = e / (RZ * RY) + 1;
            int gx = x0 + lx, gy = y0 + ly, gz = z0 + lz;
            bool in = true;
            if (P) { gx = wrap32(gx, d.n0); gy = wrap32(gy, d.n1); gz = wrap32(gz, d.n2); }
            else in = gx >= 0 && gx < d.n0 && gy >= 0 && gy < d.n1 && gz >= 0 && gz < d.n2;
            if (in && !((gx + gy + gz) & 1)) rr[j] = rhs[gx * sxg + gy * d.n2 + gz];
        }
    }
#pragma unroll
    for (int j = 0; j < NL; ++j) {
        const int e = t + j * NT;
        rt[j] = T(0);
        if (e < TILE) {
            const int lz = e % TZ + 2, ly = (e / TZ) % TY + 2, lx = e / (TZ * TY) + 2;
            const int gx = x0 + lx, gy = y0 + ly, gz = z0 + lz;
            if (gx < d.n0 && gy < d.n1 && gz < d.n2 && ((gx + gy + gz) & 1))
                rt[j] = rhs[gx * sxg + gy * d.n2 + gz];
        }
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        const int e = t + j * NT;
        if (e < BOX) box[e] = pv[j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NR; ++j) {
        const int e = t + j * NT;
        if (e >= RING) continue;
        const int lz = e % RZ + 1, ly = (e / RZ) % RY + 1, lx = e / (RZ * RY) + 1;
        int gx = x0 + lx, gy = y0 + ly, gz = z0 + lz;
        if (P) {
            gx = wrap32(gx, d.n0); gy = wrap32(gy, d.n1); gz = wrap32(gz, d.n2);
        } else if (gx < 0 || gx >= d.n0 || gy < 0 || gy >= d.n1 || gz < 0 || gz >= d.n2) {
            continue;
        }
        if ((gx + gy + gz) & 1) continue;  // red = even parity
        const int li = (lx * BY + ly) * BZ + lz;
        T sum, dg;
        neighbour_sums<T, P>(box, li, BY * BZ, BZ, gx, gy, gz, d, k, sum, dg);
        box[li] = rb_relax(k, box[li], (rr[j] + sum) / dg);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NL; ++j) {
        const int e = t + j * NT;
        if (e >= TILE) continue;
        const int lz = e % TZ + 2, ly = (e / TZ) % TY + 2, lx = e / (TZ * TY) + 2;
        const int gx = x0 + lx, gy = y0 + ly, gz = z0 + lz;
        if (gx >= d.n0 || gy >= d.n1 || gz >= d.n2) continue;
        const int li = (lx * BY + ly) * BZ + lz;
        T v = box[li];
        if ((gx + gy + gz) & 1) {
            T sum, dg;
            neighbour_sums<T, P>(box, li, BY * BZ, BZ, gx, gy, gz, d, k, sum, dg);
            v = rb_relax(k, v, (rt[j] + sum) / dg);
        }
        pout[gx * sxg + gy * d.n2 + gz] = v;
    }
}

// weighted Jacobi: out = p + omega*(rhs - A p)/diag
template <class T, bool P>
__global__ void __launch_bounds__(256) jacobi_fast_kernel(const T* __restrict__ pin,
                                                          const T* __restrict__ rhs,
                                                          T* __restrict__ pout, Dim32 d, W3<T> k,
                                                          T omega, int zero_in) {
    const int sxg = d.n1 * d.n2;
    const int total = sxg * d.n0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        const int x = i / sxg;
        const int r = i - x * sxg;
        const int y = r / d.n2, z = r - y * d.n2;
        const T dg = (P || x + 1 < d.n0 ? k.w[0] : T(0)) + (P || x > 0 ? k.w[0] : T(0)) +
                     (P || y + 1 < d.n1 ? k.w[1] : T(0)) + (P || y > 0 ? k.w[1] : T(0)) +
                     (P || z + 1 < d.n2 ? k.w[2] : T(0)) + (P || z > 0 ? k.w[2] : T(0));
        const T c = zero_in ? T(0) : pin[i];
        const T a = zero_in ? T(0) : apply_diff<T, P>(pin, i, sxg, d.n2, x, y, z, d, k);
        pout[i] = c + omega * (rhs[i] - a) / dg;
    }
}

// coarse = avg8(rhs - A p)
template <class T, bool P>
__global__ void __launch_bounds__(256) restrict_fast_kernel(const T* __restrict__ p,
                                                            const T* __restrict__ rhs,
                                                            T* __restrict__ coarse, Dim32 d,
                                                            W3<T> k) {
    const int c1 = d.n1 / 2, c2 = d.n2 / 2;
    const int cs = c1 * c2;
    const int total = cs * (d.n0 / 2);
    const int sxg = d.n1 * d.n2;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        const int x = i / cs;
        const int r = i - x * cs;
        const int y = r / c2, z = r - y * c2;
        T v = T(0);
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int cc = 0; cc < 2; ++cc) {
                    const int fx = 2 * x + a, fy = 2 * y + b, fz = 2 * z + cc;
                    const int j = fx * sxg + fy * d.n2 + fz;
                    v += rhs[j] - apply_diff<T, P>(p, j, sxg, d.n2, fx, fy, fz, d, k);
                }
        coarse[i] = v * T(0.125);
    }
}

template <class T, bool P>
__device__ inline T prolong32(const T* __restrict__ coarse, int x, int y, int z, int C0, int C1,
                              int C2) {
    const int ax = x / 2 - 1 + x % 2, ay = y / 2 - 1 + y % 2, az = z / 2 - 1 + z % 2;
    int x0, x1, y0, y1, z0, z1;
    if (P) {  // -1 <= ax and ax + 1 <= C0: one conditional wrap each
        x0 = ax < 0 ? ax + C0 : ax; x1 = ax + 1 >= C0 ? ax + 1 - C0 : ax + 1;
        y0 = ay < 0 ? ay + C1 : ay; y1 = ay + 1 >= C1 ? ay + 1 - C1 : ay + 1;
        z0 = az < 0 ? az + C2 : az; z1 = az + 1 >= C2 ? az + 1 - C2 : az + 1;
    } else {
        x0 = ax > 0 ? ax : 0; x1 = ax + 1 < C0 - 1 ? ax + 1 : C0 - 1;
        y0 = ay > 0 ? ay : 0; y1 = ay + 1 < C1 - 1 ? ay + 1 : C1 - 1;
        z0 = az > 0 ? az : 0; z1 = az + 1 < C2 - 1 ? az + 1 : C2 - 1;
    }
    const T wx = (x % 2 == 0) ? T(0.75) : T(0.25);
    const T wy = (y % 2 == 0) ? T(0.75) : T(0.25);
    const T wz = (z % 2 == 0) ? T(0.75) : T(0.25);
    const int s1 = C2, s0 = C1 * C2;
    const T* b00 = coarse + x0 * s0 + y0 * s1;
    const T* b01 = coarse + x0 * s0 + y1 * s1;
    const T* b10 = coarse + x1 * s0 + y0 * s1;
    const T* b11 = coarse + x1 * s0 + y1 * s1;
    const T c00 = b00[z0] * (T(1) - wz) + b00[z1] * wz;
    const T c01 = b01[z0] * (T(1) - wz) + b01[z1] * wz;
    const T c10 = b10[z0] * (T(1) - wz) + b10[z1] * wz;
    const T c11 = b11[z0] * (T(1) - wz) + b11[z1] * wz;
    return (c00 * (T(1) - wy) + c01 * wy) * (T(1) - wx) + (c10 * (T(1) - wy) + c11 * wy) * wx;
}

template <class T, bool P>
__global__ void __launch_bounds__(256) prolong_fast_kernel(const T* __restrict__ coarse,
                                                           T* __restrict__ fine, Dim32 d) {
    const int sxg = d.n1 * d.n2;
    const int total = sxg * d.n0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        const int x = i / sxg;
        const int r = i - x * sxg;
        const int y = r / d.n2, z = r - y * d.n2;
        fine[i] = fine[i] + prolong32<T, P>(coarse, x, y, z, d.n0 / 2, d.n1 / 2, d.n2 / 2);
    }
}

// per-workgroup [sum r, sum r^2, sum |r|, non-finite] of r = rhs - A p
template <class T, bool P>
__global__ void __launch_bounds__(256) residual_partials_kernel(const T* __restrict__ p,
                                                                const T* __restrict__ rhs, Dim32 d,
                                                                W3<T> k, double* __restrict__ partials) {
    __shared__ double sh[4][256 / WAVE];
    const int sxg = d.n1 * d.n2;
    const int total = sxg * d.n0;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        const int x = i / sxg;
        const int rr = i - x * sxg;
        const int y = rr / d.n2, z = rr - y * d.n2;
        const double r = double(rhs[i] - apply_diff<T, P>(p, i, sxg, d.n2, x, y, z, d, k));
        if (isfinite(r)) {
            v[0] += r;
            v[1] += r * r;
            v[2] += fabs(r);
        } else {
            v[3] += 1.0;
        }
    }
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    for (int q = 0; q < 4; ++q) {
        double t = v[q];
#pragma unroll
        for (int dd = 32; dd >= 1; dd >>= 1) t += __shfl_xor(t, dd, WAVE);
        if (lane == 0) sh[q][wave] = t;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double t = 0.0;
        for (int w = 0; w < 256 / WAVE; ++w) t += sh[threadIdx.x][w];
        partials[4 * blockIdx.x + threadIdx.x] = t;
    }
}

// ---------------------------------------------------------------------------
// Row kernels (levels with an even z extent <= ZROW): one wavefront per
// (x, y) row, lanes along z. Row decode, wrap and boundary tests are
// wave-uniform, lanes map to z (or to the red / black cells of a row by
// parity), so no per-element integer division remains.

constexpr int ZROW = 128;

__device__ inline int face_class(int g, int n) { return g == 0 ? 0 : (g == n - 1 ? 2 : 1); }

// 1/diag of the Neumann operator by boundary class (0 low, 1 interior,
// 2 high) per axis; entry t of a 27-table.
template <class T>
__device__ inline T inv_diag_entry(int t, const W3<T>& k) {
    const int cls[3] = {t / 9, (t / 3) % 3, t % 3};
    T dg = T(0);
    for (int a = 0; a < 3; ++a) dg += k.w[a] * T(cls[a] == 1 ? 2 : 1);
    return T(1) / dg;
}

template <class T, bool P>
__device__ inline int wrap_or(int g, int n, bool& in) {
    if (P) return wrap32(g, n);
    in = in && g >= 0 && g < n;
    return g;
}

// One red-black Gauss-Seidel sweep of TX x TY columns of rows spanning all
// of z. LDS holds the (TX+4) x (TY+4) rows of p of one tile; red is updated
// on the tile and its 1-row ring, then black on the tile. Lane r < ROWS
// writes the record of box row r (wrapped global offset, coordinates,
// in-domain flag) to LDS, so per-row bookkeeping is done once per row.
// The kernel is persistent: workgroup b sweeps tiles b, b + G, ... (XCD-
// aware: each XCD owns a contiguous run of tiles, so halo re-reads stay in
// its L2) and software-pipelined: once tile j is staged in LDS, the global
// loads of tile j+1 are issued before the red and black phases of tile j,
// so HBM traffic overlaps compute. G = ntiles gives one tile per workgroup.
struct RowRec {
    int off, gx, gy, in;
};

// Tiles first, first + slots, ... (count of them) of one level.
template <class T, bool P, int TX, int TY, int N2>
__device__ __forceinline__ void rb_rows_tiles(const T* __restrict__ pin, const T* __restrict__ rhs,
                                              T* __restrict__ pout, const Dim32 d, const W3<T> k,
                                              int zero_in, int ntx, int nty, int first, int count,
                                              int slots) {
    constexpr int NW = 4;
    constexpr int BYr = TY + 4, ROWS = (TX + 4) * BYr;
    constexpr int RYr = TY + 2, RING = (TX + 2) * RYr;
    constexpr int TILE = TX * TY;
    constexpr int LR = (ROWS + NW - 1) / NW, RR = (RING + NW - 1) / NW, TR = (TILE + NW - 1) / NW;
    static_assert(ROWS <= 256, "one thread per box row");
    extern __shared__ unsigned char raw[];
    T* box = reinterpret_cast<T*>(raw);
    __shared__ T tab[27];
    __shared__ RowRec rec[2][ROWS];
    // N2 > 0: the z extent is N2 and the x/y extents are multiples of TX/TY
    // (compile-time lane tests and tile bounds); N2 = 0: any extents.
    constexpr bool EXACT = N2 > 0;
    const int n2 = N2 > 0 ? N2 : d.n2, h2 = n2 >> 1;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (wave >= NW) __builtin_unreachable();  // 256 threads = NW wavefronts
    const int sxg = d.n1 * n2;
    auto build = [&](int buf, int tile) {
        if (threadIdx.x < ROWS) {
            const int r = threadIdx.x;
            int gx = (tile / nty) * TX - 2 + r / BYr, gy = (tile % nty) * TY - 2 + r % BYr;
            int in = 1;
            if (P) {
                gx = gx < 0 ? gx + d.n0 : (gx >= d.n0 ? gx - d.n0 : gx);
                gy = gy < 0 ? gy + d.n1 : (gy >= d.n1 ? gy - d.n1 : gy);
            } else {
                in = gx >= 0 && gx < d.n0 && gy >= 0 && gy < d.n1;
            }
            rec[buf][r] = RowRec{in ? gx * sxg + gy * n2 : 0, gx, gy, in};
        }
    };
    auto load_rec = [&](int buf, int r) {
        RowRec q = rec[buf][r];
        q.off = __builtin_amdgcn_readfirstlane(q.off);
        q.gx = __builtin_amdgcn_readfirstlane(q.gx);
        q.gy = __builtin_amdgcn_readfirstlane(q.gy);
        q.in = __builtin_amdgcn_readfirstlane(q.in);
        return q;
    };
    T pv[LR][2], rr[RR], rt[TR];
    // global loads of one tile into registers (records in rec[buf])
    auto fetch = [&](int buf, int tile) {
        const int x0 = (tile / nty) * TX - 2, y0 = (tile % nty) * TY - 2;
#pragma unroll
        for (int q = 0; q < LR; ++q) pv[q][0] = pv[q][1] = T(0);
        if (!zero_in) {
#pragma unroll
            for (int q = 0; q < LR; ++q) {
                const int r = wave + q * NW;
                if (r < ROWS) {
                    const RowRec rc = load_rec(buf, r);
                    if (P || rc.in) {
                        const T* src = pin + rc.off;
                        if (lane < n2) pv[q][0] = src[lane];
                        if (lane + 64 < n2) pv[q][1] = src[lane + 64];
                    }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < RR; ++q) {
            const int rw = wave + q * NW;
            rr[q] = T(0);
            if (rw < RING && lane < h2) {
                const RowRec rc = load_rec(buf, (1 + rw / RYr) * BYr + 1 + rw % RYr);
                if (P || rc.in) rr[q] = rhs[rc.off + 2 * lane + ((rc.gx + rc.gy) & 1)];
            }
        }
#pragma unroll
        for (int q = 0; q < TR; ++q) {
            const int tw = wave + q * NW;
            rt[q] = T(0);
            if (tw < TILE && lane < h2) {
                const RowRec rc = load_rec(buf, (2 + tw / TY) * BYr + 2 + tw % TY);
                if (EXACT || (x0 + 2 + tw / TY < d.n0 && y0 + 2 + tw % TY < d.n1))
                    rt[q] = rhs[rc.off + 2 * lane + ((rc.gx + rc.gy + 1) & 1)];
            }
        }
    };
    build(0, first);
    if (!P && threadIdx.x < 27) tab[threadIdx.x] = inv_diag_entry<T>(threadIdx.x, k);
    __syncthreads();
    fetch(0, first);
    const T inv_p = T(1) / (T(2) * (k.w[0] + k.w[1] + k.w[2]));
    auto update = [&](const T* row, int r, int gx, int gy, int z, T b) -> T {
        T s = T(0);
        if (P) {
            s += k.w[2] * row[z + 1 < n2 ? z + 1 : 0];
            s += k.w[2] * row[z > 0 ? z - 1 : n2 - 1];
        } else {
            if (z + 1 < n2) s += k.w[2] * row[z + 1];
            if (z > 0) s += k.w[2] * row[z - 1];
        }
        if (P || gy + 1 < d.n1) s += k.w[1] * box[(r + 1) * n2 + z];
        if (P || gy > 0) s += k.w[1] * box[(r - 1) * n2 + z];
        if (P || gx + 1 < d.n0) s += k.w[0] * box[(r + BYr) * n2 + z];
        if (P || gx > 0) s += k.w[0] * box[(r - BYr) * n2 + z];
        const T inv = P ? inv_p : tab[face_class(gx, d.n0) * 9 + face_class(gy, d.n1) * 3 + face_class(z, n2)];
        return rb_relax(k, row[z], (b + s) * inv);
    };
    for (int j = 0; j < count; ++j) {
        const int cur = j & 1;
        const int tile = first + j * slots;
        const int x0 = (tile / nty) * TX - 2, y0 = (tile % nty) * TY - 2;
#pragma unroll
        for (int q = 0; q < LR; ++q) {
            const int r = wave + q * NW;
            if (r < ROWS) {
                if (lane < n2) box[r * n2 + lane] = pv[q][0];
                if (lane + 64 < n2) box[r * n2 + lane + 64] = pv[q][1];
            }
        }
        T rrc[RR], rtc[TR];
#pragma unroll
        for (int q = 0; q < RR; ++q) rrc[q] = rr[q];
#pragma unroll
        for (int q = 0; q < TR; ++q) rtc[q] = rt[q];
        const bool more = j + 1 < count;
        if (more) build(cur ^ 1, tile + slots);
        __syncthreads();
        if (more) fetch(cur ^ 1, tile + slots);
#pragma unroll
        for (int q = 0; q < RR; ++q) {
            const int rw = wave + q * NW;
            if (rw >= RING || lane >= h2) continue;
            const int r = (1 + rw / RYr) * BYr + 1 + rw % RYr;
            const RowRec rc = load_rec(cur, r);
            if (!P && !rc.in) continue;
            T* row = box + r * n2;
            const int z = 2 * lane + ((rc.gx + rc.gy) & 1);
            row[z] = update(row, r, rc.gx, rc.gy, z, rrc[q]);
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < TR; ++q) {
            const int tw = wave + q * NW;
            if (tw >= TILE || lane >= h2) continue;
            if (!EXACT && (x0 + 2 + tw / TY >= d.n0 || y0 + 2 + tw % TY >= d.n1)) continue;
            const int r = (2 + tw / TY) * BYr + 2 + tw % TY;
            const RowRec rc = load_rec(cur, r);
            const T* row = box + r * n2;
            const int zb = 2 * lane + ((rc.gx + rc.gy + 1) & 1), zr = 2 * lane + ((rc.gx + rc.gy) & 1);
            T* dst = pout + rc.off;
            dst[zb] = update(row, r, rc.gx, rc.gy, zb, rtc[q]);
            dst[zr] = row[zr];
        }
        __syncthreads();
    }
}

template <class T, bool P, int TX, int TY, int N2>
__global__ void __launch_bounds__(256) rb_rows_kernel(const T* __restrict__ pin,
                                                      const T* __restrict__ rhs,
                                                      T* __restrict__ pout, Dim32 d, W3<T> k,
                                                      int zero_in, int ntx, int nty) {
    // tile sequence of this workgroup
    const int ntiles = ntx * nty;
    const bool by_xcd = (gridDim.x % 8 == 0) && (ntiles % 8 == 0);
    const int per = by_xcd ? ntiles / 8 : ntiles;
    const int slots = by_xcd ? int(gridDim.x) / 8 : int(gridDim.x);
    const int slot = by_xcd ? int(blockIdx.x) / 8 : int(blockIdx.x);
    const int first = (by_xcd ? int(blockIdx.x % 8) * per : 0) + slot;
    const int count = slot < per ? (per - slot + slots - 1) / slots : 0;
    if (count == 0) return;  // uniform over the workgroup
    rb_rows_tiles<T, P, TX, TY, N2>(pin, rhs, pout, d, k, zero_in, ntx, nty, first, count, slots);
}

// Pointers to a row and its four x/y neighbour rows (null where a Neumann
// neighbour does not exist).
template <class T, bool P>
struct Rows5 {
    const T *c, *xm, *xp, *ym, *yp;
};

template <class T, bool P>
__device__ inline Rows5<T, P> rows5(const T* base, int x, int y, const Dim32& d) {
    const int sx = d.n1 * d.n2, sy = d.n2;
    Rows5<T, P> r;
    r.c = base + x * sx + y * sy;
    if (P) {
        r.xm = base + (x > 0 ? x - 1 : d.n0 - 1) * sx + y * sy;
        r.xp = base + (x + 1 < d.n0 ? x + 1 : 0) * sx + y * sy;
        r.ym = base + x * sx + (y > 0 ? y - 1 : d.n1 - 1) * sy;
        r.yp = base + x * sx + (y + 1 < d.n1 ? y + 1 : 0) * sy;
    } else {
        r.xm = x > 0 ? r.c - sx : nullptr;
        r.xp = x + 1 < d.n0 ? r.c + sx : nullptr;
        r.ym = y > 0 ? r.c - sy : nullptr;
        r.yp = y + 1 < d.n1 ? r.c + sy : nullptr;
    }
    return r;
}

// (A p)_z on a row, difference form.
template <class T, bool P>
__device__ inline T apply_row(const Rows5<T, P>& r, int z, int n2, const W3<T>& k) {
    const T c = r.c[z];
    T a = T(0);
    if (P) {
        a += k.w[0] * (c - r.xp[z]);
        a += k.w[0] * (c - r.xm[z]);
        a += k.w[1] * (c - r.yp[z]);
        a += k.w[1] * (c - r.ym[z]);
        a += k.w[2] * (c - r.c[z + 1 < n2 ? z + 1 : 0]);
        a += k.w[2] * (c - r.c[z > 0 ? z - 1 : n2 - 1]);
    } else {
        if (r.xp) a += k.w[0] * (c - r.xp[z]);
        if (r.xm) a += k.w[0] * (c - r.xm[z]);
        if (r.yp) a += k.w[1] * (c - r.yp[z]);
        if (r.ym) a += k.w[1] * (c - r.ym[z]);
        if (z + 1 < n2) a += k.w[2] * (c - r.c[z + 1]);
        if (z > 0) a += k.w[2] * (c - r.c[z - 1]);
    }
    return a;
}

// coarse = avg8(rhs - A p); one wavefront per coarse row.
template <class T, bool P>
__device__ __forceinline__ void restrict_row(const T* __restrict__ p, const T* __restrict__ rhs,
                                             T* __restrict__ coarse, const Dim32 d, const W3<T> k,
                                             int row) {
    const int lane = threadIdx.x & 63;
    const int C0 = d.n0 / 2, C1 = d.n1 / 2, C2 = d.n2 / 2;
    if (row >= C0 * C1) return;
    const int cx = row / C1, cy = row - cx * C1;
    Rows5<T, P> pr[2][2];
    const T* rr[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            pr[a][b] = rows5<T, P>(p, 2 * cx + a, 2 * cy + b, d);
            rr[a][b] = rhs + (2 * cx + a) * d.n1 * d.n2 + (2 * cy + b) * d.n2;
        }
    T* out = coarse + (cx * C1 + cy) * C2;
    for (int cz = lane; cz < C2; cz += 64) {
        T v = T(0);
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    const int z = 2 * cz + c;
                    v += rr[a][b][z] - apply_row<T, P>(pr[a][b], z, d.n2, k);
                }
        out[cz] = v * T(0.125);
    }
}

template <class T, bool P>
__global__ void __launch_bounds__(256) restrict_rows_kernel(const T* __restrict__ p,
                                                            const T* __restrict__ rhs,
                                                            T* __restrict__ coarse, Dim32 d,
                                                            W3<T> k) {
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
    restrict_row<T, P>(p, rhs, coarse, d, k, int(blockIdx.x) * 4 + wave);
}

// fine += trilinear(coarse); one wavefront per fine row.
template <class T, bool P>
__device__ __forceinline__ void prolong_row(const T* __restrict__ coarse, T* __restrict__ fine,
                                            const Dim32 d, int row) {
    const int lane = threadIdx.x & 63;
    if (row >= d.n0 * d.n1) return;
    const int x = row / d.n1, y = row - x * d.n1;
    const int C0 = d.n0 / 2, C1 = d.n1 / 2, C2 = d.n2 / 2;
    const int ax = x / 2 - 1 + (x & 1), ay = y / 2 - 1 + (y & 1);
    int x0, x1, y0, y1;
    if (P) {  // -1 <= ax and ax + 1 <= C0
        x0 = ax < 0 ? ax + C0 : ax; x1 = ax + 1 >= C0 ? ax + 1 - C0 : ax + 1;
        y0 = ay < 0 ? ay + C1 : ay; y1 = ay + 1 >= C1 ? ay + 1 - C1 : ay + 1;
    } else {
        x0 = ax > 0 ? ax : 0; x1 = ax + 1 < C0 - 1 ? ax + 1 : C0 - 1;
        y0 = ay > 0 ? ay : 0; y1 = ay + 1 < C1 - 1 ? ay + 1 : C1 - 1;
    }
    const T wx = (x & 1) ? T(0.25) : T(0.75);
    const T wy = (y & 1) ? T(0.25) : T(0.75);
    const T* c00 = coarse + (x0 * C1 + y0) * C2;
    const T* c01 = coarse + (x0 * C1 + y1) * C2;
    const T* c10 = coarse + (x1 * C1 + y0) * C2;
    const T* c11 = coarse + (x1 * C1 + y1) * C2;
    T* f = fine + (x * d.n1 + y) * d.n2;
    for (int z = lane; z < d.n2; z += 64) {
        const int az = z / 2 - 1 + (z & 1);
        int z0, z1;
        if (P) { z0 = az < 0 ? az + C2 : az; z1 = az + 1 >= C2 ? az + 1 - C2 : az + 1; }
        else { z0 = az > 0 ? az : 0; z1 = az + 1 < C2 - 1 ? az + 1 : C2 - 1; }
        const T wz = (z & 1) ? T(0.25) : T(0.75);
        const T v00 = c00[z0] * (T(1) - wz) + c00[z1] * wz;
        const T v01 = c01[z0] * (T(1) - wz) + c01[z1] * wz;
        const T v10 = c10[z0] * (T(1) - wz) + c10[z1] * wz;
        const T v11 = c11[z0] * (T(1) - wz) + c11[z1] * wz;
        f[z] = f[z] + ((v00 * (T(1) - wy) + v01 * wy) * (T(1) - wx) + (v10 * (T(1) - wy) + v11 * wy) * wx);
    }
}

template <class T, bool P>
__global__ void __launch_bounds__(256) prolong_rows_kernel(const T* __restrict__ coarse,
                                                           T* __restrict__ fine, Dim32 d) {
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
    prolong_row<T, P>(coarse, fine, d, int(blockIdx.x) * 4 + wave);
}

// The same prolongation for a 2x2 group of fine rows: fine rows 2k+1 and
// 2k+2 of an axis interpolate between the same coarse rows k and k+1 (only
// their weights differ), so one wave reads the four coarse rows once and
// updates up to four fine rows. Group (kx, ky), kx in [-1, C0), ky in
// [-1, C1); each fine value is formed exactly as prolong_row forms it.
template <class T, bool P>
__global__ void __launch_bounds__(256) prolong_groups_kernel(const T* __restrict__ coarse,
                                                             T* __restrict__ fine, Dim32 d) {
    const int g = int(blockIdx.x) * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int C0 = d.n0 / 2, C1 = d.n1 / 2, C2 = d.n2 / 2;
    if (g >= (C0 + 1) * (C1 + 1)) return;
    const int ax = g / (C1 + 1) - 1, ay = g - (ax + 1) * (C1 + 1) - 1;
    int x0, x1, y0, y1;
    if (P) {  // -1 <= ax and ax + 1 <= C0
        x0 = ax < 0 ? ax + C0 : ax; x1 = ax + 1 >= C0 ? ax + 1 - C0 : ax + 1;
        y0 = ay < 0 ? ay + C1 : ay; y1 = ay + 1 >= C1 ? ay + 1 - C1 : ay + 1;
    } else {
        x0 = ax > 0 ? ax : 0; x1 = ax + 1 < C0 - 1 ? ax + 1 : C0 - 1;
        y0 = ay > 0 ? ay : 0; y1 = ay + 1 < C1 - 1 ? ay + 1 : C1 - 1;
    }
    const T* c00 = coarse + (x0 * C1 + y0) * C2;
    const T* c01 = coarse + (x0 * C1 + y1) * C2;
    const T* c10 = coarse + (x1 * C1 + y0) * C2;
    const T* c11 = coarse + (x1 * C1 + y1) * C2;
    const int fx[2] = {2 * ax + 1, 2 * ax + 2}, fy[2] = {2 * ay + 1, 2 * ay + 2};
    // fine rows of the group (wave-uniform; null where outside the level)
    T* rows[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
            rows[i][j] = fx[i] >= 0 && fx[i] < d.n0 && fy[j] >= 0 && fy[j] < d.n1
                             ? fine + (fx[i] * d.n1 + fy[j]) * d.n2 : nullptr;
    // Two z per lane per pass (z and z + 64): every load of a pass is issued
    // before its stores, so a 128-wide row takes one memory round trip.
    for (int zs = 0; zs < d.n2; zs += 128) {
        T v[2][4], f[2][2][2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int z = zs + lane + 64 * h;
            if (z >= d.n2) continue;
            const int az = z / 2 - 1 + (z & 1);
            int z0, z1;
            if (P) { z0 = az < 0 ? az + C2 : az; z1 = az + 1 >= C2 ? az + 1 - C2 : az + 1; }
            else { z0 = az > 0 ? az : 0; z1 = az + 1 < C2 - 1 ? az + 1 : C2 - 1; }
            const T wz = (z & 1) ? T(0.25) : T(0.75);
            v[h][0] = c00[z0] * (T(1) - wz) + c00[z1] * wz;
            v[h][1] = c01[z0] * (T(1) - wz) + c01[z1] * wz;
            v[h][2] = c10[z0] * (T(1) - wz) + c10[z1] * wz;
            v[h][3] = c11[z0] * (T(1) - wz) + c11[z1] * wz;
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    if (rows[i][j]) f[h][i][j] = rows[i][j][z];
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int z = zs + lane + 64 * h;
            if (z >= d.n2) continue;
            const T v00 = v[h][0], v01 = v[h][1], v10 = v[h][2], v11 = v[h][3];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const T wx = (fx[i] & 1) ? T(0.25) : T(0.75);
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    if (!rows[i][j]) continue;
                    const T wy = (fy[j] & 1) ? T(0.25) : T(0.75);
                    rows[i][j][z] = f[h][i][j] +
                                    ((v00 * (T(1) - wy) + v01 * wy) * (T(1) - wx) + (v10 * (T(1) - wy) + v11 * wy) * wx);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Fused small levels. On levels of up to 64^3 cells a launch costs a few
// microseconds whatever its size (global-memory round trips and launch
// gaps, not bytes), so the 2 nu + 2 launches a level takes per V-cycle
// (nu sweeps, restriction, prolongation, nu sweeps) dominate its time.
// Two kernels do a level's work in two launches, for nu = 2:
// * vdown_kernel: the pre-smoothing sweeps, the residual and its
//   restriction. A workgroup stages its TX x TY tile of full z rows in LDS
//   with a halo of 2 nu + 1 rows of p and 2 nu rows of rhs, then runs the
//   2 nu half-sweeps (red, black, red, black) on shrinking regions:
//   half-sweep s updates the tile and a ring of 2 nu - s rows, whose
//   neighbours were all updated by half-sweep s - 1. After the last one the
//   tile and its 1-row ring hold exactly what nu global sweeps produce, so
//   the residual of the tile and its 2x2x2 averages follow from LDS. The
//   workgroup writes its tile of p and the coarse cells under the tile.
// * vup_kernel: the prolongation of the coarse correction and the
//   post-smoothing sweeps. The coarse rows under the tile and a 2 nu halo
//   are z-interpolated once each into LDS (the v terms of
//   prolong_groups_kernel), p + P(e) is formed on the tile and its 2 nu halo
//   (p read from the pre-smoothed buffer), then 2 nu half-sweeps shrink to
//   the tile, which is written to the output buffer.
// Every tile recomputes the halo rows it needs. Each value is formed by the
// same expressions in the same order as in rb_rows_tiles, restrict_row and
// prolong_groups_kernel (sweep update, difference-form residual, trilinear
// weights), so the fused levels reproduce the unfused ones.

constexpr int FUSE_NU = 2;
constexpr int FUSE_NT = 512;  // threads per fused workgroup

// Box row (bx, by) of a box with halo H around the tile at (x0, y0): the
// global row (wrapped on periodic axes); false where it lies outside a
// closed domain. The halo never exceeds an axis extent (checked on the
// host), so one wrap suffices.
template <bool P, int H>
__device__ __forceinline__ bool box_row(int x0, int y0, int bx, int by, const Dim32& d, int& gx, int& gy) {
    gx = x0 - H + bx;
    gy = y0 - H + by;
    if (P) {
        gx = gx < 0 ? gx + d.n0 : (gx >= d.n0 ? gx - d.n0 : gx);
        gy = gy < 0 ? gy + d.n1 : (gy >= d.n1 ? gy - d.n1 : gy);
        return true;
    }
    return gx >= 0 && gx < d.n0 && gy >= 0 && gy < d.n1;
}

// Red-black half-sweep `colour` (0 red: gx + gy + z even) over the tile and
// a ring of W rows, in place in the p box (BY rows per box x, N2 per row);
// the rhs box has a halo of H - 1. Same update as rb_rows_tiles. A thread
// takes its cells in groups of FUSE_ILP: all LDS reads of a group are
// issued before any of its writes (cells of one colour never read each
// other), so their latencies overlap.
constexpr int FUSE_ILP = 4;

template <class T, bool P, int TX, int TY, int N2, int H, int W>
__device__ __forceinline__ void fused_half_sweep(T* box, const T* rbx, int colour, int x0, int y0,
                                                 const Dim32& d, const W3<T>& k, const T* tab, T inv_p) {
    constexpr int BY = TY + 2 * H, RY = BY - 2, H2 = N2 / 2;
    constexpr int RWY = TY + 2 * W, ITEMS = (TX + 2 * W) * RWY * H2;
    constexpr int PER = (ITEMS + FUSE_NT - 1) / FUSE_NT;
#pragma unroll
    for (int g0 = 0; g0 < PER; g0 += FUSE_ILP) {
        T val[FUSE_ILP];
        int at[FUSE_ILP];
#pragma unroll
        for (int g = 0; g < FUSE_ILP; ++g) {
            at[g] = -1;
            const int i = int(threadIdx.x) + FUSE_NT * (g0 + g);
            if (g0 + g >= PER || i >= ITEMS) continue;
            const int rr = i / H2, zl = i % H2;
            const int bx = H - W + rr / RWY, by = H - W + rr % RWY;
            int gx, gy;
            if (!box_row<P, H>(x0, y0, bx, by, d, gx, gy)) continue;
            const int r = bx * BY + by;
            const int z = 2 * zl + ((gx + gy + colour) & 1);
            const T* row = box + r * N2;
            const T b = rbx[((bx - 1) * RY + (by - 1)) * N2 + z];
            T s = T(0);
            if (P) {
                s += k.w[2] * row[z + 1 < N2 ? z + 1 : 0];
                s += k.w[2] * row[z > 0 ? z - 1 : N2 - 1];
            } else {
                if (z + 1 < N2) s += k.w[2] * row[z + 1];
                if (z > 0) s += k.w[2] * row[z - 1];
            }
            if (P || gy + 1 < d.n1) s += k.w[1] * box[(r + 1) * N2 + z];
            if (P || gy > 0) s += k.w[1] * box[(r - 1) * N2 + z];
            if (P || gx + 1 < d.n0) s += k.w[0] * box[(r + BY) * N2 + z];
            if (P || gx > 0) s += k.w[0] * box[(r - BY) * N2 + z];
            const T inv = P ? inv_p : tab[face_class(gx, d.n0) * 9 + face_class(gy, d.n1) * 3 + face_class(z, N2)];
            val[g] = rb_relax(k, row[z], (b + s) * inv);
            at[g] = r * N2 + z;
        }
#pragma unroll
        for (int g = 0; g < FUSE_ILP; ++g)
            if (at[g] >= 0) box[at[g]] = val[g];
    }
}

// The 2 nu half-sweeps of one smoothing stage; the first updates a ring of
// W0 rows, each later one a ring one row narrower.
template <class T, bool P, int TX, int TY, int N2, int H, int W0>
__device__ __forceinline__ void fused_sweeps(T* box, const T* rbx, int x0, int y0, const Dim32& d,
                                             const W3<T>& k, const T* tab, T inv_p) {
    static_assert(2 * FUSE_NU == 4, "four half-sweeps");
    fused_half_sweep<T, P, TX, TY, N2, H, W0>(box, rbx, 0, x0, y0, d, k, tab, inv_p);
    __syncthreads();
    fused_half_sweep<T, P, TX, TY, N2, H, W0 - 1>(box, rbx, 1, x0, y0, d, k, tab, inv_p);
    __syncthreads();
    fused_half_sweep<T, P, TX, TY, N2, H, W0 - 2>(box, rbx, 0, x0, y0, d, k, tab, inv_p);
    __syncthreads();
    fused_half_sweep<T, P, TX, TY, N2, H, W0 - 3>(box, rbx, 1, x0, y0, d, k, tab, inv_p);
    __syncthreads();
}

// Global rows of the p box (halo H) and of the rhs box (halo H - 1) into
// registers, 16 bytes per load; rows outside a closed domain read as 0.
template <class T, bool P, int TX, int TY, int N2, int H, int NL, bool RHS>
__device__ __forceinline__ void fused_fetch(const T* __restrict__ src, uint4 (&v)[NL], int x0, int y0,
                                            const Dim32& d) {
    constexpr int CPR = N2 * int(sizeof(T)) / 16;
    constexpr int BY = TY + 2 * H;
    constexpr int OX = RHS ? TX + 2 * H - 2 : TX + 2 * H, OY = RHS ? BY - 2 : BY;
    constexpr int CHUNKS = OX * OY * CPR;
    static_assert(NL * FUSE_NT >= CHUNKS, "register staging too small");
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        const int c = int(threadIdx.x) + FUSE_NT * i;
        v[i] = make_uint4(0u, 0u, 0u, 0u);
        if (c < CHUNKS) {
            const int row = c / CPR, part = c % CPR;
            const int bx = row / OY + (RHS ? 1 : 0), by = row % OY + (RHS ? 1 : 0);
            int gx, gy;
            if (box_row<P, H>(x0, y0, bx, by, d, gx, gy))
                v[i] = reinterpret_cast<const uint4*>(src + (int64_t(gx) * d.n1 + gy) * N2)[part];
        }
    }
}

template <class T, int NL, int CHUNKS>
__device__ __forceinline__ void fused_store(T* lds, const uint4 (&v)[NL]) {
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        const int c = int(threadIdx.x) + FUSE_NT * i;
        if (c < CHUNKS) reinterpret_cast<uint4*>(lds)[c] = v[i];
    }
}

// Tile rows of the box to the output level (16-byte stores).
template <class T, int TX, int TY, int N2, int H>
__device__ __forceinline__ void fused_write_tile(const T* box, T* __restrict__ out, int x0, int y0,
                                                 const Dim32& d) {
    constexpr int CPR = N2 * int(sizeof(T)) / 16, BY = TY + 2 * H;
    for (int c = threadIdx.x; c < TX * TY * CPR; c += FUSE_NT) {
        const int row = c / CPR, part = c % CPR;
        const int lx = row / TY, ly = row % TY;
        reinterpret_cast<uint4*>(out + (int64_t(x0 + lx) * d.n1 + (y0 + ly)) * N2)[part] =
            reinterpret_cast<const uint4*>(box + ((H + lx) * BY + (H + ly)) * N2)[part];
    }
}

template <class T, bool P, int TX, int TY, int N2>
struct FusedDown {
    static constexpr int H = 2 * FUSE_NU + 1;
    static constexpr int BX = TX + 2 * H, BY = TY + 2 * H, RX = BX - 2, RY = BY - 2;
    static constexpr int CPR = N2 * int(sizeof(T)) / 16;
    static constexpr int PCH = BX * BY * CPR, RCH = RX * RY * CPR;
    static constexpr int NP = (PCH + FUSE_NT - 1) / FUSE_NT, NR = (RCH + FUSE_NT - 1) / FUSE_NT;
    static constexpr size_t LDS = size_t(BX * BY + RX * RY) * N2 * sizeof(T);
};

template <class T, bool P, int TX, int TY, int N2>
struct FusedUp {
    static constexpr int H = 2 * FUSE_NU;
    static constexpr int BX = TX + 2 * H, BY = TY + 2 * H, RX = BX - 2, RY = BY - 2;
    static constexpr int CX = TX / 2 + H + 2, CY = TY / 2 + H + 2, C2 = N2 / 2;
    static constexpr int CPR = N2 * int(sizeof(T)) / 16, CCPR = C2 * int(sizeof(T)) / 16;
    static constexpr int PCH = BX * BY * CPR, RCH = RX * RY * CPR, CCH = CX * CY * CCPR;
    static constexpr int NP = (PCH + FUSE_NT - 1) / FUSE_NT, NR = (RCH + FUSE_NT - 1) / FUSE_NT, NC = (CCH + FUSE_NT - 1) / FUSE_NT;
    static constexpr size_t LDS = size_t(BX * BY + RX * RY) * N2 * sizeof(T) +
                                  size_t(CX * CY) * (C2 + N2) * sizeof(T);
};

// nu = 2 pre-smoothing sweeps of pin (zero_in: of a zero iterate) into the
// tile of pout, and coarse = avg8(rhs - A p) under the tile.
template <class T, bool P, int TX, int TY, int N2>
__global__ void __launch_bounds__(FUSE_NT) vdown_kernel(const T* __restrict__ pin, const T* __restrict__ rhs,
                                                    T* __restrict__ pout, T* __restrict__ coarse, Dim32 d,
                                                    W3<T> k, int zero_in, int ntx, int nty) {
    using G = FusedDown<T, P, TX, TY, N2>;
    constexpr int H = G::H, BY = G::BY, RY = G::RY, C2 = N2 / 2;
    static_assert(TX % 2 == 0 && TY % 2 == 0 && N2 % 32 == 0 && N2 <= 128, "fused tile shape");
    extern __shared__ unsigned char raw[];
    T* box = reinterpret_cast<T*>(raw);
    T* rbx = box + G::BX * BY * N2;
    __shared__ T tab[27];
    const int tile = int(xcd_tile(blockIdx.x, int64_t(ntx) * nty));
    const int x0 = (tile / nty) * TX, y0 = (tile % nty) * TY;
    {
        uint4 vp[G::NP], vr[G::NR];
        if (zero_in) {
#pragma unroll
            for (int i = 0; i < G::NP; ++i) vp[i] = make_uint4(0u, 0u, 0u, 0u);
        } else {
            fused_fetch<T, P, TX, TY, N2, H, G::NP, false>(pin, vp, x0, y0, d);
        }
        fused_fetch<T, P, TX, TY, N2, H, G::NR, true>(rhs, vr, x0, y0, d);
        fused_store<T, G::NP, G::PCH>(box, vp);
        fused_store<T, G::NR, G::RCH>(rbx, vr);
    }
    if (!P && threadIdx.x < 27) tab[threadIdx.x] = inv_diag_entry<T>(threadIdx.x, k);
    __syncthreads();
    const T inv_p = T(1) / (T(2) * (k.w[0] + k.w[1] + k.w[2]));
    fused_sweeps<T, P, TX, TY, N2, H, H - 1>(box, rbx, x0, y0, d, k, tab, inv_p);
    fused_write_tile<T, TX, TY, N2, H>(box, pout, x0, y0, d);
    // residual on the tile (difference form of apply_row) and its 2x2x2
    // averages, summed in restrict_row's order
    const int C1 = d.n1 / 2;
    for (int i = threadIdx.x; i < (TX / 2) * (TY / 2) * C2; i += FUSE_NT) {
        const int cr = i / C2, cz = i % C2;
        const int cxl = cr / (TY / 2), cyl = cr % (TY / 2);
        T v = T(0);
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int lx = 2 * cxl + a, ly = 2 * cyl + b;
                const int gx = x0 + lx, gy = y0 + ly;
                const int r = (H + lx) * BY + (H + ly);
                const T* row = box + r * N2;
                const T* rr = rbx + ((H - 1 + lx) * RY + (H - 1 + ly)) * N2;
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    const int z = 2 * cz + c;
                    const T cv = row[z];
                    T ap = T(0);
                    if (P) {
                        ap += k.w[0] * (cv - box[(r + BY) * N2 + z]);
                        ap += k.w[0] * (cv - box[(r - BY) * N2 + z]);
                        ap += k.w[1] * (cv - box[(r + 1) * N2 + z]);
                        ap += k.w[1] * (cv - box[(r - 1) * N2 + z]);
                        ap += k.w[2] * (cv - row[z + 1 < N2 ? z + 1 : 0]);
                        ap += k.w[2] * (cv - row[z > 0 ? z - 1 : N2 - 1]);
                    } else {
                        if (gx + 1 < d.n0) ap += k.w[0] * (cv - box[(r + BY) * N2 + z]);
                        if (gx > 0) ap += k.w[0] * (cv - box[(r - BY) * N2 + z]);
                        if (gy + 1 < d.n1) ap += k.w[1] * (cv - box[(r + 1) * N2 + z]);
                        if (gy > 0) ap += k.w[1] * (cv - box[(r - 1) * N2 + z]);
                        if (z + 1 < N2) ap += k.w[2] * (cv - row[z + 1]);
                        if (z > 0) ap += k.w[2] * (cv - row[z - 1]);
                    }
                    v += rr[z] - ap;
                }
            }
        coarse[(int64_t(x0 / 2 + cxl) * C1 + (y0 / 2 + cyl)) * C2 + cz] = v * T(0.125);
    }
}

// pout = nu = 2 post-smoothing sweeps of pin + trilinear(coarse), tile by
// tile.
template <class T, bool P, int TX, int TY, int N2>
__global__ void __launch_bounds__(FUSE_NT) vup_kernel(const T* __restrict__ pin, const T* __restrict__ coarse,
                                                  const T* __restrict__ rhs, T* __restrict__ pout, Dim32 d,
                                                  W3<T> k, int ntx, int nty) {
    using G = FusedUp<T, P, TX, TY, N2>;
    constexpr int H = G::H, BY = G::BY, CX = G::CX, CY = G::CY, C2 = G::C2;
    static_assert(TX % 2 == 0 && TY % 2 == 0 && N2 % 32 == 0 && N2 <= 128, "fused tile shape");
    extern __shared__ unsigned char raw[];
    T* box = reinterpret_cast<T*>(raw);
    T* rbx = box + G::BX * BY * N2;
    T* craw = rbx + G::RX * G::RY * N2;  // CX*CY coarse rows of C2
    T* vz = craw + CX * CY * C2;         // the same rows z-interpolated to N2
    __shared__ T tab[27];
    const int tile = int(xcd_tile(blockIdx.x, int64_t(ntx) * nty));
    const int x0 = (tile / nty) * TX, y0 = (tile % nty) * TY;
    const int C0 = d.n0 / 2, C1 = d.n1 / 2;
    // coarse row ia of the staged block is the unwrapped coarse x index
    // ax_lo + ia (wrapped, or clamped on a closed axis as prolong_groups
    // clamps it)
    const int ax_lo = x0 / 2 - FUSE_NU - 1, ay_lo = y0 / 2 - FUSE_NU - 1;
    uint4 vp[G::NP], vr[G::NR];
    fused_fetch<T, P, TX, TY, N2, H, G::NP, false>(pin, vp, x0, y0, d);
    fused_fetch<T, P, TX, TY, N2, H, G::NR, true>(rhs, vr, x0, y0, d);
    {
        uint4 vc[G::NC];
#pragma unroll
        for (int i = 0; i < G::NC; ++i) {
            const int c = int(threadIdx.x) + FUSE_NT * i;
            vc[i] = make_uint4(0u, 0u, 0u, 0u);
            if (c < G::CCH) {
                const int row = c / G::CCPR, part = c % G::CCPR;
                int cx = ax_lo + row / CY, cy = ay_lo + row % CY;
                if (P) {
                    cx = cx < 0 ? cx + C0 : (cx >= C0 ? cx - C0 : cx);
                    cy = cy < 0 ? cy + C1 : (cy >= C1 ? cy - C1 : cy);
                } else {
                    cx = cx < 0 ? 0 : (cx > C0 - 1 ? C0 - 1 : cx);
                    cy = cy < 0 ? 0 : (cy > C1 - 1 ? C1 - 1 : cy);
                }
                vc[i] = reinterpret_cast<const uint4*>(coarse + (int64_t(cx) * C1 + cy) * C2)[part];
            }
        }
        fused_store<T, G::NC, G::CCH>(craw, vc);
    }
    if (!P && threadIdx.x < 27) tab[threadIdx.x] = inv_diag_entry<T>(threadIdx.x, k);
    __syncthreads();
    for (int i = threadIdx.x; i < CX * CY * N2; i += FUSE_NT) {
        const int row = i / N2, z = i % N2;
        const int az = z / 2 - 1 + (z & 1);
        int z0, z1;
        if (P) { z0 = az < 0 ? az + C2 : az; z1 = az + 1 >= C2 ? az + 1 - C2 : az + 1; }
        else { z0 = az > 0 ? az : 0; z1 = az + 1 < C2 - 1 ? az + 1 : C2 - 1; }
        const T wz = (z & 1) ? T(0.25) : T(0.75);
        const T* c = craw + row * C2;
        vz[i] = c[z0] * (T(1) - wz) + c[z1] * wz;
    }
    fused_store<T, G::NR, G::RCH>(rbx, vr);
    __syncthreads();
    // p + P(e) on the box (rows outside a closed domain stay unused)
    constexpr int VEC = 16 / int(sizeof(T));
#pragma unroll
    for (int i = 0; i < G::NP; ++i) {
        const int c = int(threadIdx.x) + FUSE_NT * i;
        if (c >= G::PCH) continue;
        const int row = c / G::CPR, part = c % G::CPR;
        const int bx = row / BY, by = row % BY;
        int gx, gy;
        union { uint4 u; T t[VEC]; } val;
        val.u = vp[i];
        if (box_row<P, H>(x0, y0, bx, by, d, gx, gy)) {  // reads only: vz and registers
            const int ux = x0 - H + bx, uy = y0 - H + by;  // unwrapped fine row
            const int ia = ((ux - 1) >> 1) - ax_lo, ib = ((uy - 1) >> 1) - ay_lo;
            const T* v00 = vz + (ia * CY + ib) * N2;
            const T* v01 = vz + (ia * CY + ib + 1) * N2;
            const T* v10 = vz + ((ia + 1) * CY + ib) * N2;
            const T* v11 = vz + ((ia + 1) * CY + ib + 1) * N2;
            const T wx = (ux & 1) ? T(0.25) : T(0.75);
            const T wy = (uy & 1) ? T(0.25) : T(0.75);
#pragma unroll
            for (int e = 0; e < VEC; ++e) {
                const int z = part * VEC + e;
                val.t[e] = val.t[e] + ((v00[z] * (T(1) - wy) + v01[z] * wy) * (T(1) - wx) +
                                       (v10[z] * (T(1) - wy) + v11[z] * wy) * wx);
            }
        }
        vp[i] = val.u;
    }
    // all vz reads above are issued before the box writes below
    fused_store<T, G::NP, G::PCH>(box, vp);
    __syncthreads();
    const T inv_p = T(1) / (T(2) * (k.w[0] + k.w[1] + k.w[2]));
    fused_sweeps<T, P, TX, TY, N2, H, H - 1>(box, rbx, x0, y0, d, k, tab, inv_p);
    fused_write_tile<T, TX, TY, N2, H>(box, pout, x0, y0, d);
}

// per-workgroup [sum r, sum r^2, sum |r|, non-finite] of r = rhs - A p,
// one wavefront per row, grid-stride over rows.
template <class T, bool P>
__global__ void __launch_bounds__(256) residual_rows_kernel(const T* __restrict__ p,
                                                            const T* __restrict__ rhs, Dim32 d,
                                                            W3<T> k, double* __restrict__ partials) {
    __shared__ double sh[4][4];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int row = blockIdx.x * 4 + wave; row < d.n0 * d.n1; row += gridDim.x * 4) {
        const int x = row / d.n1, y = row - x * d.n1;
        const Rows5<T, P> pr = rows5<T, P>(p, x, y, d);
        const T* rr = rhs + x * d.n1 * d.n2 + y * d.n2;
        for (int z = lane; z < d.n2; z += 64) {
            const double r = double(rr[z] - apply_row<T, P>(pr, z, d.n2, k));
            if (isfinite(r)) {
                v[0] += r;
                v[1] += r * r;
                v[2] += fabs(r);
            } else {
                v[3] += 1.0;
            }
        }
    }
    for (int q = 0; q < 4; ++q) {
        double t = v[q];
#pragma unroll
        for (int dd = 32; dd >= 1; dd >>= 1) t += __shfl_xor(t, dd, WAVE);
        if (lane == 0) sh[q][wave] = t;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double t = 0.0;
        for (int w = 0; w < 4; ++w) t += sh[threadIdx.x][w];
        partials[4 * blockIdx.x + threadIdx.x] = t;
    }
}

// ---------------------------------------------------------------------------
// FP64 iterative refinement around FP32 V-cycles: r = b - A x in FP64 with
// the FP32 copy of r written for the inner solve and [sum, sum r^2,
// sum |r|, non-finite] partials per workgroup; and x += e.

template <bool P>
__global__ void __launch_bounds__(256) refine_residual_kernel(const double* __restrict__ x,
                                                              const double* __restrict__ b,
                                                              float* __restrict__ r32, Dim32 d,
                                                              W3<double> k,
                                                              double* __restrict__ partials) {
    __shared__ double sh[4][4];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int row = blockIdx.x * 4 + wave; row < d.n0 * d.n1; row += gridDim.x * 4) {
        const int xx = row / d.n1, yy = row - xx * d.n1;
        const Rows5<double, P> pr = rows5<double, P>(x, xx, yy, d);
        const int off = row * d.n2;
        for (int z = lane; z < d.n2; z += 64) {
            const double r = b[off + z] - apply_row<double, P>(pr, z, d.n2, k);
            r32[off + z] = float(r);
            if (isfinite(r)) {
                v[0] += r;
                v[1] += r * r;
                v[2] += fabs(r);
            } else {
                v[3] += 1.0;
            }
        }
    }
    for (int q = 0; q < 4; ++q) {
        double t = v[q];
#pragma unroll
        for (int dd = 32; dd >= 1; dd >>= 1) t += __shfl_xor(t, dd, WAVE);
        if (lane == 0) sh[q][wave] = t;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double t = 0.0;
        for (int w = 0; w < 4; ++w) t += sh[threadIdx.x][w];
        partials[4 * blockIdx.x + threadIdx.x] = t;
    }
}

// One refinement step in one pass: x_out = x + e (FP64 + FP32 correction)
// and r = b - A x_out in FP64, with the neighbours' x + e formed on the fly
// from the same expression that writes them, so the result is identical to
// an update pass followed by a residual pass. Reduce: also the residual's
// block partials [sum, sum of squares, sum |.|, non-finite].
template <bool P, bool Reduce>
__global__ void __launch_bounds__(256) refine_step_kernel(const double* __restrict__ x,
                                                          const float* __restrict__ e,
                                                          const double* __restrict__ b,
                                                          double* __restrict__ xo,
                                                          float* __restrict__ r32, Dim32 d,
                                                          W3<double> k,
                                                          double* __restrict__ partials) {
    __shared__ double sh[4][4];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int row = blockIdx.x * 4 + wave; row < d.n0 * d.n1; row += gridDim.x * 4) {
        const int xx = row / d.n1, yy = row - xx * d.n1;
        const Rows5<double, P> px = rows5<double, P>(x, xx, yy, d);
        const Rows5<float, P> pe = rows5<float, P>(e, xx, yy, d);
        const int off = row * d.n2;
        for (int z = lane; z < d.n2; z += 64) {
            const double c = px.c[z] + double(pe.c[z]);
            double a = 0.0;
            if (P) {
                a += k.w[0] * (c - (px.xp[z] + double(pe.xp[z])));
                a += k.w[0] * (c - (px.xm[z] + double(pe.xm[z])));
                a += k.w[1] * (c - (px.yp[z] + double(pe.yp[z])));
                a += k.w[1] * (c - (px.ym[z] + double(pe.ym[z])));
                const int zp = z + 1 < d.n2 ? z + 1 : 0, zm = z > 0 ? z - 1 : d.n2 - 1;
                a += k.w[2] * (c - (px.c[zp] + double(pe.c[zp])));
                a += k.w[2] * (c - (px.c[zm] + double(pe.c[zm])));
            } else {
                if (px.xp) a += k.w[0] * (c - (px.xp[z] + double(pe.xp[z])));
                if (px.xm) a += k.w[0] * (c - (px.xm[z] + double(pe.xm[z])));
                if (px.yp) a += k.w[1] * (c - (px.yp[z] + double(pe.yp[z])));
                if (px.ym) a += k.w[1] * (c - (px.ym[z] + double(pe.ym[z])));
                if (z + 1 < d.n2) a += k.w[2] * (c - (px.c[z + 1] + double(pe.c[z + 1])));
                if (z > 0) a += k.w[2] * (c - (px.c[z - 1] + double(pe.c[z - 1])));
            }
            const double r = b[off + z] - a;
            xo[off + z] = c;
            r32[off + z] = float(r);
            if (Reduce) {
                if (isfinite(r)) {
                    v[0] += r;
                    v[1] += r * r;
                    v[2] += fabs(r);
                } else {
                    v[3] += 1.0;
                }
            }
        }
    }
    if (!Reduce) return;
    for (int q = 0; q < 4; ++q) {
        double t = v[q];
#pragma unroll
        for (int dd = 32; dd >= 1; dd >>= 1) t += __shfl_xor(t, dd, WAVE);
        if (lane == 0) sh[q][wave] = t;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double t = 0.0;
        for (int w = 0; w < 4; ++w) t += sh[threadIdx.x][w];
        partials[4 * blockIdx.x + threadIdx.x] = t;
    }
}

__global__ void refine_update_kernel(double* __restrict__ x, const float* __restrict__ e,
                                     int64_t count) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
         i += int64_t(gridDim.x) * blockDim.x)
        x[i] = x[i] + double(e[i]);
}

// ---------------------------------------------------------------------------
// single-workgroup V-cycle over the coarse levels, all arrays in LDS. Like
// the fine levels it removes no means: the operator annihilates constants,
// smoothing and prolongation carry a constant unchanged, restricted
// residuals of a zero-mean right-hand side have zero mean, and the coarsest
// inv(A + 11^T/n) solve returns the mean of its right-hand side, so the
// reference's mean subtractions change only rounding; the caller removes
// the level-0 mean once after the last cycle.
// Work items are enumerated per colour (lanes map to the red or the black
// cells of rows directly) and decoded with float-reciprocal division plus
// an exact integer correction; 1/diag comes from a per-level table.

template <class T>
struct CoarseArgs {
    int32_t levels;  // number of levels handled (>= 1)
    int32_t smoother, nu;
    T omega;
    Dim32 d[AFSI_MG_MAX_LEVELS];
    W3<T> k[AFSI_MG_MAX_LEVELS];
    int32_t off_p[AFSI_MG_MAX_LEVELS], off_r[AFSI_MG_MAX_LEVELS], off_w[AFSI_MG_MAX_LEVELS];
    const T* rhs_in;   // first handled level's rhs (global)
    const T* p_in;     // its initial iterate, or null (zero)
    T* p_out;          // mean-free result of the first handled level
    const T* coarse_inverse;
    int32_t coarse_cells;
    const int32_t* active;  // batched launches: skip when *active == 0 (null: always run)
};

// floor(e / d) for 0 <= e < 2^22, d >= 1
__device__ inline int qdiv(int e, int d) {
    int q = int((float(e) + 0.5f) / float(d));
    if (q * d > e) --q;
    else if ((q + 1) * d <= e) ++q;
    return q;
}

// sum_nb w_a p_nb (and, for Neumann, the boundary class index) at (x,y,z).
template <class T, bool P>
__device__ inline T lds_nsum(const T* a, int i, int x, int y, int z, const Dim32& d,
                             const W3<T>& k) {
    const int sx = d.n1 * d.n2, sy = d.n2;
    T s = T(0);
    if (P) {
        s += k.w[0] * a[i + (x + 1 < d.n0 ? sx : -(d.n0 - 1) * sx)];
        s += k.w[0] * a[i + (x > 0 ? -sx : (d.n0 - 1) * sx)];
        s += k.w[1] * a[i + (y + 1 < d.n1 ? sy : -(d.n1 - 1) * sy)];
        s += k.w[1] * a[i + (y > 0 ? -sy : (d.n1 - 1) * sy)];
        s += k.w[2] * a[i + (z + 1 < d.n2 ? 1 : -(d.n2 - 1))];
        s += k.w[2] * a[i + (z > 0 ? -1 : (d.n2 - 1))];
    } else {
        if (x + 1 < d.n0) s += k.w[0] * a[i + sx];
        if (x > 0) s += k.w[0] * a[i - sx];
        if (y + 1 < d.n1) s += k.w[1] * a[i + sy];
        if (y > 0) s += k.w[1] * a[i - sy];
        if (z + 1 < d.n2) s += k.w[2] * a[i + 1];
        if (z > 0) s += k.w[2] * a[i - 1];
    }
    return s;
}

template <class T, bool P>
__device__ inline T lds_inv(const T* tab, T inv_p, int x, int y, int z, const Dim32& d) {
    if (P) return inv_p;
    return tab[face_class(x, d.n0) * 9 + face_class(y, d.n1) * 3 + face_class(z, d.n2)];
}

// Packed cell descriptor: x (bits 24-29), y (18-23), z (12-17), index (0-11).
__device__ inline uint32_t pack_cell(int x, int y, int z, int i) {
    return (uint32_t(x) << 24) | (uint32_t(y) << 18) | (uint32_t(z) << 12) | uint32_t(i);
}
__device__ inline void unpack_cell(uint32_t c, int& x, int& y, int& z, int& i) {
    x = int(c >> 24);
    y = int((c >> 18) & 63);
    z = int((c >> 12) & 63);
    i = int(c & 4095);
}

// Lists: for a level with even n2, cells of colour c (x+y+z even: red) at
// list[c*n/2 + row*n2/2 + z/2]; otherwise all cells in index order.
__device__ inline void build_list(uint32_t* list, const Dim32& d) {
    const int n = d.n0 * d.n1 * d.n2;
    const bool split = (d.n2 & 1) == 0;
    for (int e = threadIdx.x; e < n; e += COARSE_THREADS) {
        const int row = qdiv(e, d.n2), z = e - row * d.n2;
        const int x = qdiv(row, d.n1), y = row - x * d.n1;
        const uint32_t c = pack_cell(x, y, z, e);
        if (split) list[((x + y + z) & 1) * (n >> 1) + row * (d.n2 >> 1) + (z >> 1)] = c;
        else list[e] = c;
    }
}

template <class T, bool P>
__device__ inline void lds_smooth(T* p, T* w, const T* r, const uint32_t* list, const Dim32& d,
                                  const W3<T>& k, const T* tab, T inv_p, int smoother, int nu,
                                  T omega) {
    const int n = d.n0 * d.n1 * d.n2;
    const int sx = d.n1 * d.n2, sy = d.n2;
    if (smoother == 1) {
        const int half = n >> 1;
        for (int it = 0; it < nu; ++it) {
            for (int color = 0; color < 2; ++color) {
                for (int e = threadIdx.x; e < half; e += COARSE_THREADS) {
                    int x, y, z, i;
                    unpack_cell(list[color * half + e], x, y, z, i);
                    T s;
                    if (P) {
                        s = k.w[0] * p[i + (x + 1 < d.n0 ? sx : -(d.n0 - 1) * sx)];
                        s += k.w[0] * p[i + (x > 0 ? -sx : (d.n0 - 1) * sx)];
                        s += k.w[1] * p[i + (y + 1 < d.n1 ? sy : -(d.n1 - 1) * sy)];
                        s += k.w[1] * p[i + (y > 0 ? -sy : (d.n1 - 1) * sy)];
                        s += k.w[2] * p[i + (z + 1 < d.n2 ? 1 : -(d.n2 - 1))];
                        s += k.w[2] * p[i + (z > 0 ? -1 : (d.n2 - 1))];
                    } else {
                        s = T(0);
                        if (x + 1 < d.n0) s += k.w[0] * p[i + sx];
                        if (x > 0) s += k.w[0] * p[i - sx];
                        if (y + 1 < d.n1) s += k.w[1] * p[i + sy];
                        if (y > 0) s += k.w[1] * p[i - sy];
                        if (z + 1 < d.n2) s += k.w[2] * p[i + 1];
                        if (z > 0) s += k.w[2] * p[i - 1];
                    }
                    p[i] = rb_relax(k, p[i], (r[i] + s) * lds_inv<T, P>(tab, inv_p, x, y, z, d));
                }
                __syncthreads();
            }
        }
    } else {
        for (int it = 0; it < nu; ++it) {
            for (int e = threadIdx.x; e < n; e += COARSE_THREADS) {
                int x, y, z, i;
                unpack_cell(list[e], x, y, z, i);
                const T a = apply_diff<T, P>(p, i, sx, sy, x, y, z, d, k);
                w[i] = p[i] + omega * (r[i] - a) * lds_inv<T, P>(tab, inv_p, x, y, z, d);
            }
            __syncthreads();
            for (int i = threadIdx.x; i < n; i += COARSE_THREADS) p[i] = w[i];
            __syncthreads();
        }
    }
}

template <class T, bool P>
__device__ __forceinline__ void coarse_cycle_body(const CoarseArgs<T>& a) {
    extern __shared__ unsigned char smem_raw[];
    T* smem = reinterpret_cast<T*>(smem_raw);
    __shared__ T tabs[AFSI_MG_MAX_LEVELS][27];
    __shared__ uint32_t lists[COARSE_LIST];
    __shared__ int list_off[AFSI_MG_MAX_LEVELS];
    const int L = a.levels;
    // The first level's rhs and initial iterate (at most COARSE_MAX cells)
    // are loaded into registers first, so the loads' latency overlaps the
    // cell-list build; they reach LDS after it.
    constexpr int LOADS = int(COARSE_MAX / COARSE_THREADS);
    static_assert(LOADS * COARSE_THREADS == COARSE_MAX, "whole loads per thread");
    T rv[LOADS], pv[LOADS];
    {
        const int n = a.d[0].n0 * a.d[0].n1 * a.d[0].n2;
#pragma unroll
        for (int q = 0; q < LOADS; ++q) {
            const int i = int(threadIdx.x) + q * COARSE_THREADS;
            rv[q] = i < n ? a.rhs_in[i] : T(0);
            pv[q] = i < n && a.p_in ? a.p_in[i] : T(0);
        }
    }
    if (threadIdx.x == 0) {
        int off = 0;
        for (int l = 0; l < L; ++l) {
            list_off[l] = off;
            off += a.d[l].n0 * a.d[l].n1 * a.d[l].n2;
        }
    }
    if (!P && threadIdx.x < 27 * L) {
        const int l = threadIdx.x / 27;
        tabs[l][threadIdx.x - 27 * l] = inv_diag_entry<T>(threadIdx.x - 27 * l, a.k[l]);
    }
    __syncthreads();
    for (int l = 0; l < L; ++l) build_list(lists + list_off[l], a.d[l]);
    {
        const int n = a.d[0].n0 * a.d[0].n1 * a.d[0].n2;
        T* p = smem + a.off_p[0];
        T* r = smem + a.off_r[0];
#pragma unroll
        for (int q = 0; q < LOADS; ++q) {
            const int i = int(threadIdx.x) + q * COARSE_THREADS;
            if (i < n) {
                r[i] = rv[q];
                p[i] = pv[q];
            }
        }
        __syncthreads();
    }
    for (int l = 0; l < L - 1; ++l) {
        const Dim32& d = a.d[l];
        const W3<T>& k = a.k[l];
        T* p = smem + a.off_p[l];
        T* r = smem + a.off_r[l];
        T* w = smem + a.off_w[l];
        const T inv_p = T(1) / (T(2) * (k.w[0] + k.w[1] + k.w[2]));
        lds_smooth<T, P>(p, w, r, lists + list_off[l], d, k, tabs[l], inv_p, a.smoother, a.nu, a.omega);
        const Dim32& c = a.d[l + 1];
        T* rc = smem + a.off_r[l + 1];
        T* pc = smem + a.off_p[l + 1];
        const uint32_t* clist = lists + list_off[l + 1];
        const int nc = c.n0 * c.n1 * c.n2;
        const int sx = d.n1 * d.n2, sy = d.n2;
        for (int e = threadIdx.x; e < nc; e += COARSE_THREADS) {
            int x, y, z, i;
            unpack_cell(clist[e], x, y, z, i);
            T v = T(0);
            for (int aa = 0; aa < 2; ++aa)
                for (int b = 0; b < 2; ++b)
                    for (int cc = 0; cc < 2; ++cc) {
                        const int fx = 2 * x + aa, fy = 2 * y + b, fz = 2 * z + cc;
                        const int j = fx * sx + fy * sy + fz;
                        v += r[j] - apply_diff<T, P>(p, j, sx, sy, fx, fy, fz, d, k);
                    }
            rc[i] = v * T(0.125);
            pc[i] = T(0);
        }
        __syncthreads();
    }
    {
        const Dim32& d = a.d[L - 1];
        T* p = smem + a.off_p[L - 1];
        const T* r = smem + a.off_r[L - 1];
        const int m = a.coarse_cells;
        // tpr threads per row (a power of two <= 64, inside one wave)
        int tpr = 1;
        while (tpr < 64 && tpr * 2 * m <= COARSE_THREADS) tpr *= 2;
        const int lane = threadIdx.x % tpr;
        for (int row0 = 0; row0 < m; row0 += COARSE_THREADS / tpr) {
            const int row = row0 + threadIdx.x / tpr;
            T s = T(0);
            if (row < m)
                for (int col = lane; col < m; col += tpr) s += a.coarse_inverse[row * m + col] * r[col];
            for (int dd = tpr / 2; dd >= 1; dd >>= 1) s += __shfl_xor(s, dd, WAVE);
            if (row < m && lane == 0) p[row] = s;
        }
        __syncthreads();
    }
    for (int l = L - 2; l >= 0; --l) {
        const Dim32& d = a.d[l];
        const W3<T>& k = a.k[l];
        T* p = smem + a.off_p[l];
        T* r = smem + a.off_r[l];
        T* w = smem + a.off_w[l];
        const T* pc = smem + a.off_p[l + 1];
        const uint32_t* list = lists + list_off[l];
        const int n = d.n0 * d.n1 * d.n2;
        for (int e = threadIdx.x; e < n; e += COARSE_THREADS) {
            int x, y, z, i;
            unpack_cell(list[e], x, y, z, i);
            p[i] = p[i] + prolong32<T, P>(pc, x, y, z, d.n0 / 2, d.n1 / 2, d.n2 / 2);
        }
        __syncthreads();
        const T inv_p = T(1) / (T(2) * (k.w[0] + k.w[1] + k.w[2]));
        lds_smooth<T, P>(p, w, r, list, d, k, tabs[l], inv_p, a.smoother, a.nu, a.omega);
    }
    {
        const Dim32& d = a.d[0];
        const int n = d.n0 * d.n1 * d.n2;
        const T* p = smem + a.off_p[0];
        for (int i = threadIdx.x; i < n; i += COARSE_THREADS) a.p_out[i] = p[i];
    }
}

template <class T, bool P>
__global__ void __launch_bounds__(COARSE_THREADS) coarse_cycle_kernel(CoarseArgs<T> a) {
    coarse_cycle_body<T, P>(a);
}

// ---------------------------------------------------------------------------
// batched kernels: entry i of a launch's descriptor array covers workgroups
// [begin_i, begin_{i+1}); workgroup b finds its entry by binary search over
// the (few) entries with wave-uniform scalar loads.

template <class D>
__device__ inline int find_entry(const D* __restrict__ desc, int n, int b) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (desc[mid].begin <= b) lo = mid;
        else hi = mid;
    }
    return __builtin_amdgcn_readfirstlane(lo);
}

template <class T>
struct SweepDesc {
    const T* pin;
    const T* rhs;
    T* pout;
    const int32_t* active;
    Dim32 d;
    W3<T> k;
    int ntx, nty;
    int begin;
    int zero_in;
};

template <class T>
struct RowDesc {
    const T* a;  // restriction: p; prolongation: coarse p
    const T* b;  // restriction: rhs
    T* c;        // restriction: coarse rhs; prolongation: fine p
    const int32_t* active;
    Dim32 d;     // fine level
    W3<T> k;
    int begin;
};

struct RefDesc {
    const double* x;
    const float* e;
    const double* b;
    double* xo;
    float* r32;
    const int32_t* active;
    Dim32 d;
    W3<double> k;
    int begin;
};

struct VecDesc {
    double* x;
    int n;
    int begin;
};

constexpr int VEC_CHUNK = 4096;  // elements per workgroup in batched vector passes

template <class T, bool P>
__global__ void __launch_bounds__(256) rb_rows_batch_kernel(const SweepDesc<T>* __restrict__ desc,
                                                            int n) {
    const int b = blockIdx.x;
    const SweepDesc<T>& s = desc[find_entry(desc, n, b)];
    if (s.active && s.active[0] == 0) return;  // uniform over the workgroup
    rb_rows_tiles<T, P, 2, 2, 0>(s.pin, s.rhs, s.pout, s.d, s.k, s.zero_in, s.ntx, s.nty,
                                 b - s.begin, 1, 1);
}

template <class T, bool P>
__global__ void __launch_bounds__(256) restrict_batch_kernel(const RowDesc<T>* __restrict__ desc,
                                                             int n) {
    const int b = blockIdx.x;
    const RowDesc<T>& s = desc[find_entry(desc, n, b)];
    if (s.active && s.active[0] == 0) return;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    restrict_row<T, P>(s.a, s.b, s.c, s.d, s.k, (b - s.begin) * 4 + wave);
}

template <class T, bool P>
__global__ void __launch_bounds__(256) prolong_batch_kernel(const RowDesc<T>* __restrict__ desc,
                                                            int n) {
    const int b = blockIdx.x;
    const RowDesc<T>& s = desc[find_entry(desc, n, b)];
    if (s.active && s.active[0] == 0) return;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    prolong_row<T, P>(s.a, s.c, s.d, (b - s.begin) * 4 + wave);
}

template <class T, bool P>
__global__ void __launch_bounds__(COARSE_THREADS) coarse_batch_kernel(const CoarseArgs<T>* __restrict__ args) {
    const CoarseArgs<T>& a = args[blockIdx.x];
    if (a.active && a.active[0] == 0) return;
    coarse_cycle_body<T, P>(a);
}

__device__ inline void block_stats(double v[4], double* __restrict__ out) {
    __shared__ double sh[4][4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int q = 0; q < 4; ++q) {
        double t = v[q];
#pragma unroll
        for (int dd = 32; dd >= 1; dd >>= 1) t += __shfl_xor(t, dd, WAVE);
        if (lane == 0) sh[q][wave] = t;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double t = 0.0;
        for (int w = 0; w < 4; ++w) t += sh[threadIdx.x][w];
        out[threadIdx.x] = t;
    }
}

__device__ inline void add_stats(double v[4], double r) {
    if (isfinite(r)) {
        v[0] += r;
        v[1] += r * r;
        v[2] += fabs(r);
    } else {
        v[3] += 1.0;
    }
}

// Step: x_out = x + e and r = b - A x_out (one wavefront per row);
// otherwise r = b - A x. Inactive instances copy x to x_out (Step).
template <bool P, bool Step, bool Reduce>
__global__ void __launch_bounds__(256) refine_batch_kernel(const RefDesc* __restrict__ desc, int n,
                                                           double* __restrict__ partials) {
    const int b = blockIdx.x;
    const RefDesc& s = desc[find_entry(desc, n, b)];
    const bool live = !(s.active && s.active[0] == 0);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const Dim32 d = s.d;
    const W3<double> k = s.k;
    const int row = (b - s.begin) * 4 + wave;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (row < d.n0 * d.n1) {
        const int off = row * d.n2;
        if (!live) {
            if (Step)
                for (int z = lane; z < d.n2; z += 64) s.xo[off + z] = s.x[off + z];
        } else {
            const int xx = row / d.n1, yy = row - xx * d.n1;
            const Rows5<double, P> px = rows5<double, P>(s.x, xx, yy, d);
            const Rows5<float, P> pe = rows5<float, P>(s.e, xx, yy, d);
            auto val = [&](const double* xr, const float* er, int z) -> double {
                return Step ? xr[z] + double(er[z]) : xr[z];
            };
            for (int z = lane; z < d.n2; z += 64) {
                const double c = val(px.c, pe.c, z);
                double a = 0.0;
                if (P) {
                    a += k.w[0] * (c - val(px.xp, pe.xp, z));
                    a += k.w[0] * (c - val(px.xm, pe.xm, z));
                    a += k.w[1] * (c - val(px.yp, pe.yp, z));
                    a += k.w[1] * (c - val(px.ym, pe.ym, z));
                    const int zp = z + 1 < d.n2 ? z + 1 : 0, zm = z > 0 ? z - 1 : d.n2 - 1;
                    a += k.w[2] * (c - val(px.c, pe.c, zp));
                    a += k.w[2] * (c - val(px.c, pe.c, zm));
                } else {
                    if (px.xp) a += k.w[0] * (c - val(px.xp, pe.xp, z));
                    if (px.xm) a += k.w[0] * (c - val(px.xm, pe.xm, z));
                    if (px.yp) a += k.w[1] * (c - val(px.yp, pe.yp, z));
                    if (px.ym) a += k.w[1] * (c - val(px.ym, pe.ym, z));
                    if (z + 1 < d.n2) a += k.w[2] * (c - val(px.c, pe.c, z + 1));
                    if (z > 0) a += k.w[2] * (c - val(px.c, pe.c, z - 1));
                }
                const double r = s.b[off + z] - a;
                if (Step) s.xo[off + z] = c;
                s.r32[off + z] = float(r);
                if (Reduce) add_stats(v, r);
            }
        }
    }
    if (Reduce) block_stats(v, partials + 4 * b);
}

// Per instance: sum its workgroups' partials in a fixed order, store the
// stats at results[16 i + slot]; slot 0 (the residual) also updates
// active[i] (0 once the residual meets tol[i] or is non-finite). Frozen
// (inactive) instances are left untouched.
template <class D>
__global__ void __launch_bounds__(64) batch_reduce_kernel(const D* __restrict__ desc, int n,
                                                          int total_blocks,
                                                          const double* __restrict__ partials,
                                                          double* __restrict__ results, int slot,
                                                          const double* __restrict__ tol,
                                                          int32_t* __restrict__ active) {
    const int i = blockIdx.x;
    if (slot == 0 && active[i] == 0) return;
    const int beg = desc[i].begin, end = i + 1 < n ? desc[i + 1].begin : total_blocks;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = beg + int(threadIdx.x); b < end; b += 64)
        for (int q = 0; q < 4; ++q) v[q] += partials[4 * b + q];
    for (int q = 0; q < 4; ++q) {
#pragma unroll
        for (int dd = 32; dd >= 1; dd >>= 1) v[q] += __shfl_xor(v[q], dd, WAVE);
    }
    if (threadIdx.x == 0) {
        for (int q = 0; q < 4; ++q) results[16 * i + slot + q] = v[q];
        if (slot == 0 && (v[3] > 0.0 || !(sqrt(v[1]) > tol[i]))) active[i] = 0;
    }
}

__global__ void __launch_bounds__(256) vec_partials_kernel(const VecDesc* __restrict__ desc, int n,
                                                           double* __restrict__ partials) {
    const int b = blockIdx.x;
    const VecDesc& s = desc[find_entry(desc, n, b)];
    const int lo = (b - s.begin) * VEC_CHUNK, hi = min(s.n, lo + VEC_CHUNK);
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int j = lo + int(threadIdx.x); j < hi; j += 256) add_stats(v, s.x[j]);
    block_stats(v, partials + 4 * b);
}

__global__ void __launch_bounds__(256) vec_sub_mean_kernel(const VecDesc* __restrict__ desc, int n,
                                                           const double* __restrict__ results,
                                                           int slot) {
    const int b = blockIdx.x;
    const int i = find_entry(desc, n, b);
    const VecDesc& s = desc[i];
    const double m = results[16 * i + slot] / double(s.n);
    const int lo = (b - s.begin) * VEC_CHUNK, hi = min(s.n, lo + VEC_CHUNK);
    for (int j = lo + int(threadIdx.x); j < hi; j += 256) s.x[j] = s.x[j] - m;
}

__global__ void batch_begin_kernel(int n, const double* __restrict__ results, double* __restrict__ tol,
                                   int32_t* __restrict__ active, double rtol, double atol) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // b statistics after the mean removal sit in slot 8
    tol[i] = fmax(atol, rtol * sqrt(results[16 * i + 8 + 1]));
    active[i] = 1;
}


// ---------------------------------------------------------------------------
// host driver

// Dynamic LDS a kernel may request: the device's per-workgroup LDS minus the
// kernel's static LDS. The first call per (kernel, device) raises the
// kernel's dynamic-LDS attribute to that amount; `need` is checked against it.
int dynamic_lds(const void* fn, size_t need) {
    struct Entry { const void* fn; int dev; size_t avail; };
    static std::mutex lock;
    static Entry cache[64];
    static int used = 0;
    int dev = 0;
    AFSI_HIP(hipGetDevice(&dev));
    size_t avail = 0;
    bool found = false;
    {
        std::lock_guard<std::mutex> g(lock);
        for (int i = 0; i < used && !found; ++i)
            if (cache[i].fn == fn && cache[i].dev == dev) { avail = cache[i].avail; found = true; }
    }
    if (!found) {
        int per_block = 0, optin = 0;
        AFSI_HIP(hipDeviceGetAttribute(&per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        AFSI_HIP(hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev));
        const size_t limit = size_t(per_block > optin ? per_block : optin);
        hipFuncAttributes attr{};
        AFSI_HIP(hipFuncGetAttributes(&attr, fn));
        avail = limit > attr.sharedSizeBytes ? limit - attr.sharedSizeBytes : 0;
        AFSI_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(avail)));
        std::lock_guard<std::mutex> g(lock);
        if (used < 64) cache[used++] = Entry{fn, dev, avail};
    }
    if (need > avail) {
        set_error("kernel needs %zu bytes of dynamic LDS; the device allows %zu", need, avail);
        return E_LIMIT;
    }
    return OK;
}

template <class T>
struct FastRunner {
    const afsi_mg3d_plan& pl;
    hipStream_t s;
    int coarse_first;  // first level solved in LDS

    T* P(int l) const { return static_cast<T*>(pl.p[l]); }
    T* R(int l) const { return static_cast<T*>(pl.rhs[l]); }
    T* W(int l) const { return static_cast<T*>(pl.tmp[l]); }
    Dim32 D(int l) const { return Dim32{int(pl.n[l][0]), int(pl.n[l][1]), int(pl.n[l][2])}; }
    int64_t cells(int l) const { return pl.n[l][0] * pl.n[l][1] * pl.n[l][2]; }
    W3<T> K(int l) const {
        W3<T> k;
        for (int a = 0; a < 3; ++a) k.w[a] = T(pl.inv_h2[l][a]);
        k.relax = T(pl.relax > 0.0 ? pl.relax : 1.0);
        return k;
    }

    bool rows_ok(int l) const { return pl.n[l][2] <= ZROW && pl.n[l][2] % 2 == 0; }

    template <int TXr, int TYr, int N2>
    int rb_rows_n(int l, const T* in, T* out, bool zero_in) const {
        const Dim32 d = D(l);
        const int ntx = int(cdiv(d.n0, TXr)), nty = int(cdiv(d.n1, TYr));
        const int ntiles = ntx * nty;
        const size_t bytes = size_t((TXr + 4) * (TYr + 4)) * d.n2 * sizeof(T);
        const void* fn = pl.periodic ? reinterpret_cast<const void*>(rb_rows_kernel<T, true, TXr, TYr, N2>)
                                     : reinterpret_cast<const void*>(rb_rows_kernel<T, false, TXr, TYr, N2>);
        if (int r = dynamic_lds(fn, bytes)) return r;
        int g = pl.sweep_blocks[l] > 0 ? pl.sweep_blocks[l] : ntiles;
        if (g > ntiles) g = ntiles;
        if (pl.periodic)
            rb_rows_kernel<T, true, TXr, TYr, N2><<<unsigned(g), 256, bytes, s>>>(in, R(l), out, d, K(l), zero_in, ntx, nty);
        else
            rb_rows_kernel<T, false, TXr, TYr, N2><<<unsigned(g), 256, bytes, s>>>(in, R(l), out, d, K(l), zero_in, ntx, nty);
        AFSI_LAUNCH_CHECK();
        return OK;
    }

    // z extent as a compile-time constant where the level tiles exactly
    template <int TXr, int TYr>
    int rb_rows(int l, const T* in, T* out, bool zero_in) const {
        const Dim32 d = D(l);
        if (d.n0 % TXr == 0 && d.n1 % TYr == 0) {
            if (d.n2 == 128) return rb_rows_n<TXr, TYr, 128>(l, in, out, zero_in);
            if (d.n2 == 64) return rb_rows_n<TXr, TYr, 64>(l, in, out, zero_in);
            if (d.n2 == 32) return rb_rows_n<TXr, TYr, 32>(l, in, out, zero_in);
        }
        return rb_rows_n<TXr, TYr, 0>(l, in, out, zero_in);
    }

    int sweep(int l, const T* in, T* out, bool zero_in) const {
        const Dim32 d = D(l);
        const W3<T> k = K(l);
        if (pl.smoother == 1 && rows_ok(l)) {
            int tile = pl.sweep_tile[l];
            if (tile != 2 && tile != 4 && tile != 8 && tile != 84)
                tile = cdiv(d.n0, 4) * cdiv(d.n1, 4) >= 1024 ? 4 : 2;
            if (tile == 8) return rb_rows<8, 8>(l, in, out, zero_in);
            if (tile == 84) return rb_rows<8, 4>(l, in, out, zero_in);
            if (tile == 4) return rb_rows<4, 4>(l, in, out, zero_in);
            return rb_rows<2, 2>(l, in, out, zero_in);
        }
        if (pl.smoother == 1) {
            // 8x8x32 tiles on large levels, 4x4x32 where that leaves < 512 tiles
            const bool big = cdiv(d.n0, 8) * cdiv(d.n1, 8) * cdiv(d.n2, 32) >= 512;
            const int tx = big ? 8 : 4, ty = big ? 8 : 4, tz = 32;
            const int ntx = int(cdiv(d.n0, tx)), nty = int(cdiv(d.n1, ty)), ntz = int(cdiv(d.n2, tz));
            const unsigned g = unsigned(ntx * nty * ntz);
            if (big) {
                if (pl.periodic)
                    rb_sweep_kernel<T, true, 8, 8, 32, 512><<<g, 512, 0, s>>>(in, R(l), out, d, k, zero_in, ntx, nty, ntz);
                else
                    rb_sweep_kernel<T, false, 8, 8, 32, 512><<<g, 512, 0, s>>>(in, R(l), out, d, k, zero_in, ntx, nty, ntz);
            } else {
                if (pl.periodic)
                    rb_sweep_kernel<T, true, 4, 4, 32, 256><<<g, 256, 0, s>>>(in, R(l), out, d, k, zero_in, ntx, nty, ntz);
                else
                    rb_sweep_kernel<T, false, 4, 4, 32, 256><<<g, 256, 0, s>>>(in, R(l), out, d, k, zero_in, ntx, nty, ntz);
            }
        } else {
            const unsigned g = grid_for(cells(l), 256, 16384);
            if (pl.periodic)
                jacobi_fast_kernel<T, true><<<g, 256, 0, s>>>(in, R(l), out, d, k, T(pl.omega), zero_in);
            else
                jacobi_fast_kernel<T, false><<<g, 256, 0, s>>>(in, R(l), out, d, k, T(pl.omega), zero_in);
        }
        AFSI_LAUNCH_CHECK();
        return OK;
    }

    // nu sweeps ping-ponging between p and tmp; `cur` tracks the result.
    int smooth(int l, T*& cur, bool zero_first) const {
        for (int it = 0; it < pl.nu; ++it) {
            T* out = (cur == P(l)) ? W(l) : P(l);
            if (int r = sweep(l, cur, out, it == 0 && zero_first)) return r;
            cur = out;
        }
        return OK;
    }

    int coarse(int first, const T* p_in, T* p_out) const {
        CoarseArgs<T> a{};
        a.levels = pl.levels - first;
        a.smoother = pl.smoother;
        a.nu = pl.nu;
        a.omega = T(pl.omega);
        int32_t off = 0;
        for (int j = 0; j < a.levels; ++j) {
            const int l = first + j;
            a.d[j] = D(l);
            a.k[j] = K(l);
            const int32_t n = int32_t(cells(l));
            a.off_p[j] = off; off += n;
            a.off_r[j] = off; off += n;
            a.off_w[j] = off; off += (pl.smoother == 1 ? 0 : n);
        }
        const size_t bytes = size_t(off) * sizeof(T);
        a.rhs_in = R(first);
        a.p_in = p_in;
        a.p_out = p_out;
        a.coarse_inverse = static_cast<const T*>(pl.coarse_inverse);
        a.coarse_cells = int32_t(pl.coarse_cells);
        if (int r = dynamic_lds(pl.periodic ? reinterpret_cast<const void*>(coarse_cycle_kernel<T, true>)
                                            : reinterpret_cast<const void*>(coarse_cycle_kernel<T, false>),
                                bytes))
            return r;
        if (pl.periodic) coarse_cycle_kernel<T, true><<<1, COARSE_THREADS, bytes, s>>>(a);
        else coarse_cycle_kernel<T, false><<<1, COARSE_THREADS, bytes, s>>>(a);
        AFSI_LAUNCH_CHECK();
        return OK;
    }

    // Tile of the fused down/up launches at level l: 2 for 2x2, 4 for 4x4
    // (FP32) or 4x2 (FP64) rows; 0 where the level runs unfused.
    int fuse_code(int l) const {
        if (pl.smoother != 1 || pl.nu != FUSE_NU || l >= coarse_first || pl.fuse_cells < 0) return 0;
        const int64_t limit = pl.fuse_cells > 0 ? int64_t(pl.fuse_cells) : (int64_t(1) << 18);
        const int64_t n0 = pl.n[l][0], n1 = pl.n[l][1], n2 = pl.n[l][2];
        if (cells(l) > limit || (n2 != 32 && n2 != 64)) return 0;
        const int code = n2 == 64 ? 4 : 2;
        const int tx = code, ty = (code == 4 && sizeof(T) == 8) ? 2 : code;
        if (n0 % tx || n1 % ty || n0 < 2 * FUSE_NU + 1 || n1 < 2 * FUSE_NU + 1) return 0;
        const void* ptrs[4] = {pl.p[l], pl.tmp[l], pl.rhs[l], pl.p[l + 1]};
        for (const void* q : ptrs)
            if (reinterpret_cast<uintptr_t>(q) % 16) return 0;
        return code;
    }

    template <int TX, int TY, int N2>
    int vdown_n(int l, const T* in, T* out, bool zero_in) const {
        const Dim32 d = D(l);
        const int ntx = d.n0 / TX, nty = d.n1 / TY;
        const size_t bytes = FusedDown<T, true, TX, TY, N2>::LDS;
        const void* fn = pl.periodic ? reinterpret_cast<const void*>(vdown_kernel<T, true, TX, TY, N2>)
                                     : reinterpret_cast<const void*>(vdown_kernel<T, false, TX, TY, N2>);
        if (int r = dynamic_lds(fn, bytes)) return r;
        const unsigned g = unsigned(ntx * nty);
        if (pl.periodic)
            vdown_kernel<T, true, TX, TY, N2><<<g, FUSE_NT, bytes, s>>>(in, R(l), out, R(l + 1), d, K(l), zero_in, ntx, nty);
        else
            vdown_kernel<T, false, TX, TY, N2><<<g, FUSE_NT, bytes, s>>>(in, R(l), out, R(l + 1), d, K(l), zero_in, ntx, nty);
        AFSI_LAUNCH_CHECK();
        return OK;
    }

    template <int TX, int TY, int N2>
    int vup_n(int l, const T* in, const T* cc, T* out) const {
        const Dim32 d = D(l);
        const int ntx = d.n0 / TX, nty = d.n1 / TY;
        const size_t bytes = FusedUp<T, true, TX, TY, N2>::LDS;
        const void* fn = pl.periodic ? reinterpret_cast<const void*>(vup_kernel<T, true, TX, TY, N2>)
                                     : reinterpret_cast<const void*>(vup_kernel<T, false, TX, TY, N2>);
        if (int r = dynamic_lds(fn, bytes)) return r;
        const unsigned g = unsigned(ntx * nty);
        if (pl.periodic)
            vup_kernel<T, true, TX, TY, N2><<<g, FUSE_NT, bytes, s>>>(in, cc, R(l), out, d, K(l), ntx, nty);
        else
            vup_kernel<T, false, TX, TY, N2><<<g, FUSE_NT, bytes, s>>>(in, cc, R(l), out, d, K(l), ntx, nty);
        AFSI_LAUNCH_CHECK();
        return OK;
    }

    int vdown(int code, int l, const T* in, T* out, bool zero_in) const {
        constexpr int TY4 = sizeof(T) == 8 ? 2 : 4;
        if (code == 2)
            return D(l).n2 == 64 ? vdown_n<2, 2, 64>(l, in, out, zero_in) : vdown_n<2, 2, 32>(l, in, out, zero_in);
        return D(l).n2 == 64 ? vdown_n<4, TY4, 64>(l, in, out, zero_in) : vdown_n<4, TY4, 32>(l, in, out, zero_in);
    }

    int vup(int code, int l, const T* in, const T* cc, T* out) const {
        constexpr int TY4 = sizeof(T) == 8 ? 2 : 4;
        if (code == 2) return D(l).n2 == 64 ? vup_n<2, 2, 64>(l, in, cc, out) : vup_n<2, 2, 32>(l, in, cc, out);
        return D(l).n2 == 64 ? vup_n<4, TY4, 64>(l, in, cc, out) : vup_n<4, TY4, 32>(l, in, cc, out);
    }

    int cycle(int l, T*& cur, bool zero_initial = false) const {
        if (l == coarse_first) return coarse(l, (l == 0 && !zero_initial) ? cur : nullptr, cur);
        if (const int code = fuse_code(l)) {
            T* out = cur == P(l) ? W(l) : P(l);
            if (int r = vdown(code, l, cur, out, l > 0 || zero_initial)) return r;
            cur = out;
            T* cc = P(l + 1);
            if (int r = cycle(l + 1, cc)) return r;
            out = cur == P(l) ? W(l) : P(l);
            if (int r = vup(code, l, cur, cc, out)) return r;
            cur = out;
            return OK;
        }
        // coarse iterates (and a zero initial level-0 iterate) start at zero
        if (int r = smooth(l, cur, l > 0 || zero_initial)) return r;
        const Dim32 d = D(l);
        const W3<T> k = K(l);
        if (rows_ok(l)) {
            const unsigned g = unsigned(cdiv(int64_t(d.n0 / 2) * (d.n1 / 2), 4));
            if (pl.periodic) restrict_rows_kernel<T, true><<<g, 256, 0, s>>>(cur, R(l), R(l + 1), d, k);
            else restrict_rows_kernel<T, false><<<g, 256, 0, s>>>(cur, R(l), R(l + 1), d, k);
        } else {
            const unsigned g = grid_for(cells(l + 1), 256, 16384);
            if (pl.periodic) restrict_fast_kernel<T, true><<<g, 256, 0, s>>>(cur, R(l), R(l + 1), d, k);
            else restrict_fast_kernel<T, false><<<g, 256, 0, s>>>(cur, R(l), R(l + 1), d, k);
        }
        AFSI_LAUNCH_CHECK();
        T* cc = P(l + 1);
        if (int r = cycle(l + 1, cc)) return r;
        if (rows_ok(l)) {
            const unsigned gf = unsigned(cdiv(int64_t(d.n0 / 2 + 1) * (d.n1 / 2 + 1), 4));
            if (pl.periodic) prolong_groups_kernel<T, true><<<gf, 256, 0, s>>>(cc, cur, d);
            else prolong_groups_kernel<T, false><<<gf, 256, 0, s>>>(cc, cur, d);
        } else {
            const unsigned gf = grid_for(cells(l), 256, 16384);
            if (pl.periodic) prolong_fast_kernel<T, true><<<gf, 256, 0, s>>>(cc, cur, d);
            else prolong_fast_kernel<T, false><<<gf, 256, 0, s>>>(cc, cur, d);
        }
        AFSI_LAUNCH_CHECK();
        return smooth(l, cur, false);
    }
};

// First level solved by the LDS kernel: at most COARSE_MAX cells, every
// extent <= 64, and all of its levels' cell lists within COARSE_LIST.
static int choose_coarse(const afsi_mg3d_plan& pl) {
    for (int l = 0; l < pl.levels; ++l) {
        int64_t total = 0;
        bool ok = pl.n[l][0] * pl.n[l][1] * pl.n[l][2] <= COARSE_MAX;
        for (int j = l; j < pl.levels && ok; ++j) {
            ok = pl.n[j][0] <= 64 && pl.n[j][1] <= 64 && pl.n[j][2] <= 64;
            total += pl.n[j][0] * pl.n[j][1] * pl.n[j][2];
        }
        if (ok && total <= COARSE_LIST) return l;
    }
    return -1;
}

template <class T>
static int fast_check(const afsi_mg3d_plan* pl) {
    AFSI_REQUIRE(pl && pl->levels >= 1 && pl->levels <= AFSI_MG_MAX_LEVELS, "invalid multigrid plan");
    AFSI_REQUIRE(pl->nu >= 1 && (pl->smoother == 0 || pl->smoother == 1), "invalid smoother");
    AFSI_REQUIRE(pl->reduce_partials && pl->reduce_result, "missing reduction scratch");
    for (int l = 0; l < pl->levels; ++l) {
        AFSI_REQUIRE(pl->n[l][0] > 0 && pl->n[l][1] > 0 && pl->n[l][2] > 0 && pl->p[l] && pl->rhs[l],
                     "invalid level %d", l);
        AFSI_REQUIRE(pl->n[l][0] * pl->n[l][1] * pl->n[l][2] < (int64_t(1) << 31),
                     "level %d exceeds 2^31 cells", l);
        AFSI_REQUIRE(pl->tmp[l], "missing smoothing buffer at level %d", l);
        if (l + 1 < pl->levels)
            for (int a = 0; a < 3; ++a)
                AFSI_REQUIRE(pl->n[l][a] == 2 * pl->n[l + 1][a], "levels %d/%d are not 2:1", l, l + 1);
    }
    const int last = pl->levels - 1;
    AFSI_REQUIRE(pl->coarse_inverse &&
                     pl->coarse_cells == pl->n[last][0] * pl->n[last][1] * pl->n[last][2],
                 "coarse inverse does not match the coarsest level");
    AFSI_REQUIRE(pl->coarse_cells <= 512, "coarsest level exceeds 512 cells");
    AFSI_REQUIRE(choose_coarse(*pl) >= 0,
                 "fused driver needs coarse levels with extents <= 64 (use the unfused driver)");
    return OK;
}

template <class T>
static int fast_cycles(const afsi_mg3d_plan* pl, int32_t cycles, void* stream) {
    if (int r = fast_check<T>(pl)) return r;
    AFSI_REQUIRE(cycles >= 0, "negative cycle count");
    const FastRunner<T> run{*pl, hipStream_t(stream), choose_coarse(*pl)};
    for (int c = 0; c < cycles; ++c) {
        T* cur = run.P(0);
        if (int r = run.cycle(0, cur, c == 0 && pl->zero_initial)) return r;
        AFSI_REQUIRE(cur == run.P(0), "level-0 iterate left the canonical buffer");
    }
    return OK;
}

template <class T>
static int fast_residual(const afsi_mg3d_plan* pl, void* stream) {
    if (int r = fast_check<T>(pl)) return r;
    const FastRunner<T> run{*pl, hipStream_t(stream), choose_coarse(*pl)};
    const hipStream_t s = hipStream_t(stream);
    const int64_t nb = grid_for(run.cells(0), 256, pl->partials_capacity / 4);
    AFSI_REQUIRE(4 * nb <= pl->partials_capacity, "reduction scratch too small");
    if (run.rows_ok(0)) {
        const int64_t rows = pl->n[0][0] * pl->n[0][1];
        int64_t nr = cdiv(rows, 4);
        if (nr > pl->partials_capacity / 4) nr = pl->partials_capacity / 4;
        if (pl->periodic)
            residual_rows_kernel<T, true><<<unsigned(nr), 256, 0, s>>>(run.P(0), run.R(0), run.D(0), run.K(0),
                                                                      pl->reduce_partials);
        else
            residual_rows_kernel<T, false><<<unsigned(nr), 256, 0, s>>>(run.P(0), run.R(0), run.D(0), run.K(0),
                                                                       pl->reduce_partials);
        AFSI_LAUNCH_CHECK();
        return launch_reduce_final(pl->reduce_partials, nr, pl->reduce_result, s);
    }
    if (pl->periodic)
        residual_partials_kernel<T, true><<<unsigned(nb), 256, 0, s>>>(run.P(0), run.R(0), run.D(0), run.K(0),
                                                                      pl->reduce_partials);
    else
        residual_partials_kernel<T, false><<<unsigned(nb), 256, 0, s>>>(run.P(0), run.R(0), run.D(0), run.K(0),
                                                                       pl->reduce_partials);
    AFSI_LAUNCH_CHECK();
    return launch_reduce_final(pl->reduce_partials, nb, pl->reduce_result, s);
}

// ---------------------------------------------------------------------------
// batched mixed-precision solve (host side)

struct BatchLayout {
    int n = 0, levels = 0, nu = 0, periodic = 0;
    int cf[1024];                      // first LDS-solved level per instance
    std::vector<int> members[AFSI_MG_MAX_LEVELS];
    size_t sweep_off[AFSI_MG_MAX_LEVELS][3];  // k = 0 (zero input), even k, odd k
    int sweep_blocks[AFSI_MG_MAX_LEVELS];
    size_t sweep_lds[AFSI_MG_MAX_LEVELS];
    size_t restrict_off[AFSI_MG_MAX_LEVELS], prolong_off[AFSI_MG_MAX_LEVELS];
    int restrict_blocks[AFSI_MG_MAX_LEVELS], prolong_blocks[AFSI_MG_MAX_LEVELS];
    size_t coarse_off[2] = {0, 0}, coarse_lds = 0;  // first cycle of an iteration (zero start); later ones
    size_t ref_off[4] = {0, 0, 0, 0};  // residual of x[0]; step from x[0]; step from x[1]; residual of x[1]
    int ref_blocks = 0;
    size_t vec_off[3] = {0, 0, 0};     // b, x[0], x[1]
    int vec_blocks[3] = {0, 0, 0};
    size_t total = 0;
    int64_t partials = 0;              // doubles
};

static size_t align64(size_t b) { return (b + 63) & ~size_t(63); }

static int batch_layout(const afsi_mgb_batch* bt, BatchLayout& L) {
    AFSI_REQUIRE(bt && bt->items && bt->count >= 1 && bt->count <= 1024, "invalid batch (1..1024 instances)");
    L.n = bt->count;
    L.periodic = bt->items[0].plan.periodic;
    L.nu = bt->items[0].plan.nu;
    for (int i = 0; i < L.n; ++i) {
        const afsi_mgb_item& it = bt->items[i];
        const afsi_mg3d_plan& pl = it.plan;
        if (int r = fast_check<float>(&pl)) return r;
        AFSI_REQUIRE(pl.periodic == L.periodic, "instances of one batch share the boundary type");
        AFSI_REQUIRE(pl.smoother == 1 && pl.nu == L.nu, "batched instances use red-black GS with one nu");
        AFSI_REQUIRE(it.x[0] && it.x[1] && it.b && it.x[0] != it.x[1], "invalid FP64 iterate buffers");
        L.cf[i] = choose_coarse(pl);
        for (int l = 0; l < L.cf[i]; ++l)
            AFSI_REQUIRE(pl.n[l][2] <= ZROW && pl.n[l][2] % 2 == 0 && pl.n[l][0] % 2 == 0 &&
                             pl.n[l][1] % 2 == 0,
                         "instance %d level %d: batched sweeps need even extents and z <= %d", i, l, ZROW);
        AFSI_REQUIRE(pl.n[0][0] * pl.n[0][1] * pl.n[0][2] < (int64_t(1) << 31), "instance too large");
        if (L.cf[i] > L.levels) L.levels = L.cf[i];
    }
    size_t off = 0;
    int64_t partials = 0;
    for (int l = 0; l < L.levels; ++l) {
        L.members[l].clear();
        int sb = 0, rb = 0, pb = 0;
        size_t lds = 0;
        for (int i = 0; i < L.n; ++i) {
            if (l >= L.cf[i]) continue;
            L.members[l].push_back(i);
            const int64_t* nn = bt->items[i].plan.n[l];
            sb += int((nn[0] / 2) * (nn[1] / 2));
            rb += int(cdiv((nn[0] / 2) * (nn[1] / 2), 4));
            pb += int(cdiv(nn[0] * nn[1], 4));
            const size_t box = size_t(6 * 6) * size_t(nn[2]) * sizeof(float);
            if (box > lds) lds = box;
        }
        L.sweep_blocks[l] = sb;
        L.restrict_blocks[l] = rb;
        L.prolong_blocks[l] = pb;
        L.sweep_lds[l] = lds;
        const size_t m = L.members[l].size();
        for (int v = 0; v < 3; ++v) {
            L.sweep_off[l][v] = off;
            off = align64(off + m * sizeof(SweepDesc<float>));
        }
        L.restrict_off[l] = off;
        off = align64(off + m * sizeof(RowDesc<float>));
        L.prolong_off[l] = off;
        off = align64(off + m * sizeof(RowDesc<float>));
    }
    for (int c = 0; c < 2; ++c) {
        L.coarse_off[c] = off;
        off = align64(off + size_t(L.n) * sizeof(CoarseArgs<float>));
    }
    L.coarse_lds = 0;
    L.ref_blocks = 0;
    for (int i = 0; i < L.n; ++i) {
        const afsi_mg3d_plan& pl = bt->items[i].plan;
        size_t cells = 0;
        for (int l = L.cf[i]; l < pl.levels; ++l) cells += 2 * size_t(pl.n[l][0] * pl.n[l][1] * pl.n[l][2]);
        if (cells * sizeof(float) > L.coarse_lds) L.coarse_lds = cells * sizeof(float);
        L.ref_blocks += int(cdiv(pl.n[0][0] * pl.n[0][1], 4));
    }
    for (int v = 0; v < 4; ++v) {
        L.ref_off[v] = off;
        off = align64(off + size_t(L.n) * sizeof(RefDesc));
    }
    for (int v = 0; v < 3; ++v) {
        L.vec_off[v] = off;
        off = align64(off + size_t(L.n) * sizeof(VecDesc));
        int nb = 0;
        for (int i = 0; i < L.n; ++i) {
            const int64_t* nn = bt->items[i].plan.n[0];
            nb += int(cdiv(nn[0] * nn[1] * nn[2], VEC_CHUNK));
        }
        L.vec_blocks[v] = nb;
    }
    L.total = off;
    partials = std::max<int64_t>(L.ref_blocks, L.vec_blocks[0]);
    L.partials = 4 * partials;
    return OK;
}

template <class D>
static D* at(void* ws, size_t off) {
    return reinterpret_cast<D*>(static_cast<char*>(ws) + off);
}

static int batch_prepare(const afsi_mgb_batch* bt, hipStream_t s) {
    BatchLayout L;
    if (int r = batch_layout(bt, L)) return r;
    AFSI_REQUIRE(bt->workspace && bt->workspace_bytes >= L.total, "batch workspace too small");
    AFSI_REQUIRE(bt->active && bt->tol && bt->results && bt->partials, "missing batch state buffers");
    std::vector<char> host(L.total, 0);
    void* hw = host.data();
    for (int l = 0; l < L.levels; ++l) {
        int sb = 0, rb = 0, pb = 0;
        for (size_t j = 0; j < L.members[l].size(); ++j) {
            const int i = L.members[l][j];
            const afsi_mg3d_plan& pl = bt->items[i].plan;
            const Dim32 d{int(pl.n[l][0]), int(pl.n[l][1]), int(pl.n[l][2])};
            W3<float> k;
            for (int a = 0; a < 3; ++a) k.w[a] = float(pl.inv_h2[l][a]);
            k.relax = float(pl.relax > 0.0 ? pl.relax : 1.0);
            float* P = static_cast<float*>(pl.p[l]);
            float* W = static_cast<float*>(pl.tmp[l]);
            const float* R = static_cast<const float*>(pl.rhs[l]);
            const int ntx = d.n0 / 2, nty = d.n1 / 2;
            const int32_t* act = bt->active + i;
            at<SweepDesc<float>>(hw, L.sweep_off[l][0])[j] = SweepDesc<float>{P, R, W, act, d, k, ntx, nty, sb, 1};
            at<SweepDesc<float>>(hw, L.sweep_off[l][1])[j] = SweepDesc<float>{P, R, W, act, d, k, ntx, nty, sb, 0};
            at<SweepDesc<float>>(hw, L.sweep_off[l][2])[j] = SweepDesc<float>{W, R, P, act, d, k, ntx, nty, sb, 0};
            float* cur = (L.nu % 2) ? W : P;  // iterate after the pre-smoothing sweeps
            at<RowDesc<float>>(hw, L.restrict_off[l])[j] =
                RowDesc<float>{cur, R, static_cast<float*>(pl.rhs[l + 1]), act, d, k, rb};
            at<RowDesc<float>>(hw, L.prolong_off[l])[j] =
                RowDesc<float>{static_cast<const float*>(pl.p[l + 1]), nullptr, cur, act, d, k, pb};
            sb += ntx * nty;
            rb += int(cdiv(int64_t(d.n0 / 2) * (d.n1 / 2), 4));
            pb += int(cdiv(int64_t(d.n0) * d.n1, 4));
        }
    }
    int rbeg = 0, vbeg = 0;
    for (int i = 0; i < L.n; ++i) {
        const afsi_mgb_item& it = bt->items[i];
        const afsi_mg3d_plan& pl = it.plan;
        CoarseArgs<float> a{};
        const int first = L.cf[i];
        a.levels = pl.levels - first;
        a.smoother = pl.smoother;
        a.nu = pl.nu;
        a.omega = float(pl.omega);
        int32_t o = 0;
        for (int j = 0; j < a.levels; ++j) {
            const int l = first + j;
            a.d[j] = Dim32{int(pl.n[l][0]), int(pl.n[l][1]), int(pl.n[l][2])};
            for (int q = 0; q < 3; ++q) a.k[j].w[q] = float(pl.inv_h2[l][q]);
            a.k[j].relax = float(pl.relax > 0.0 ? pl.relax : 1.0);
            const int32_t c = int32_t(pl.n[l][0] * pl.n[l][1] * pl.n[l][2]);
            a.off_p[j] = o; o += c;
            a.off_r[j] = o; o += c;
            a.off_w[j] = o;
        }
        a.rhs_in = static_cast<const float*>(pl.rhs[first]);
        a.p_in = nullptr;  // zero initial guess on every level
        a.p_out = static_cast<float*>(pl.p[first]);
        a.coarse_inverse = static_cast<const float*>(pl.coarse_inverse);
        a.coarse_cells = int32_t(pl.coarse_cells);
        a.active = bt->active + i;
        at<CoarseArgs<float>>(hw, L.coarse_off[0])[i] = a;
        // a later cycle of one refinement iteration continues from the
        // correction so far when the coarse kernel holds level 0 (it loads
        // p_in before it writes p_out, so the two may alias)
        if (first == 0) a.p_in = a.p_out;
        at<CoarseArgs<float>>(hw, L.coarse_off[1])[i] = a;
        const Dim32 d0{int(pl.n[0][0]), int(pl.n[0][1]), int(pl.n[0][2])};
        W3<double> k0;
        for (int q = 0; q < 3; ++q) k0.w[q] = pl.inv_h2[0][q];
        const float* e = static_cast<const float*>(pl.p[0]);
        float* r32 = static_cast<float*>(pl.rhs[0]);
        const int32_t* act = bt->active + i;
        at<RefDesc>(hw, L.ref_off[0])[i] = RefDesc{it.x[0], e, it.b, it.x[0], r32, nullptr, d0, k0, rbeg};
        at<RefDesc>(hw, L.ref_off[1])[i] = RefDesc{it.x[0], e, it.b, it.x[1], r32, act, d0, k0, rbeg};
        at<RefDesc>(hw, L.ref_off[2])[i] = RefDesc{it.x[1], e, it.b, it.x[0], r32, act, d0, k0, rbeg};
        at<RefDesc>(hw, L.ref_off[3])[i] = RefDesc{it.x[1], e, it.b, it.x[1], r32, nullptr, d0, k0, rbeg};
        rbeg += int(cdiv(int64_t(d0.n0) * d0.n1, 4));
        const int cells = d0.n0 * d0.n1 * d0.n2;
        at<VecDesc>(hw, L.vec_off[0])[i] = VecDesc{it.b, cells, vbeg};
        at<VecDesc>(hw, L.vec_off[1])[i] = VecDesc{it.x[0], cells, vbeg};
        at<VecDesc>(hw, L.vec_off[2])[i] = VecDesc{it.x[1], cells, vbeg};
        vbeg += int(cdiv(cells, VEC_CHUNK));
    }
    AFSI_HIP(hipMemcpyAsync(bt->workspace, host.data(), L.total, hipMemcpyHostToDevice, s));
    AFSI_HIP(hipStreamSynchronize(s));
    return OK;
}

template <bool P>
static int batch_vcycle(const afsi_mgb_batch* bt, const BatchLayout& L, hipStream_t s, bool first = true) {
    void* ws = bt->workspace;
    const void* sweep_fn = reinterpret_cast<const void*>(rb_rows_batch_kernel<float, P>);
    for (int l = 0; l < L.levels; ++l) {
        const int m = int(L.members[l].size());
        if (int r = dynamic_lds(sweep_fn, L.sweep_lds[l])) return r;
        for (int it = 0; it < L.nu; ++it) {
            // the first sweep of a cycle starts from zero, except on level 0
            // of a later cycle of one refinement iteration (which continues
            // from the correction so far)
            const int v = it == 0 ? (l == 0 && !first ? 1 : 0) : (it % 2 ? 2 : 1);
            rb_rows_batch_kernel<float, P><<<unsigned(L.sweep_blocks[l]), 256, L.sweep_lds[l], s>>>(
                at<SweepDesc<float>>(ws, L.sweep_off[l][v]), m);
        }
        restrict_batch_kernel<float, P><<<unsigned(L.restrict_blocks[l]), 256, 0, s>>>(
            at<RowDesc<float>>(ws, L.restrict_off[l]), m);
        AFSI_LAUNCH_CHECK();
    }
    const void* coarse_fn = reinterpret_cast<const void*>(coarse_batch_kernel<float, P>);
    if (int r = dynamic_lds(coarse_fn, L.coarse_lds)) return r;
    coarse_batch_kernel<float, P><<<unsigned(L.n), COARSE_THREADS, L.coarse_lds, s>>>(
        at<CoarseArgs<float>>(ws, L.coarse_off[first ? 0 : 1]));
    AFSI_LAUNCH_CHECK();
    for (int l = L.levels - 1; l >= 0; --l) {
        const int m = int(L.members[l].size());
        prolong_batch_kernel<float, P><<<unsigned(L.prolong_blocks[l]), 256, 0, s>>>(
            at<RowDesc<float>>(ws, L.prolong_off[l]), m);
        for (int it = L.nu; it < 2 * L.nu; ++it) {
            const int v = it % 2 ? 2 : 1;
            rb_rows_batch_kernel<float, P><<<unsigned(L.sweep_blocks[l]), 256, L.sweep_lds[l], s>>>(
                at<SweepDesc<float>>(ws, L.sweep_off[l][v]), m);
        }
        AFSI_LAUNCH_CHECK();
    }
    return OK;
}

template <bool P>
static int batch_refine(const afsi_mgb_batch* bt, const BatchLayout& L, int which, bool reduce,
                        hipStream_t s) {
    const RefDesc* desc = at<RefDesc>(bt->workspace, L.ref_off[which]);
    const unsigned g = unsigned(L.ref_blocks);
    const bool residual_only = which == 0 || which == 3;
    if (residual_only) refine_batch_kernel<P, false, true><<<g, 256, 0, s>>>(desc, L.n, bt->partials);
    else if (reduce) refine_batch_kernel<P, true, true><<<g, 256, 0, s>>>(desc, L.n, bt->partials);
    else refine_batch_kernel<P, true, false><<<g, 256, 0, s>>>(desc, L.n, bt->partials);
    AFSI_LAUNCH_CHECK();
    if (residual_only || reduce) {
        batch_reduce_kernel<RefDesc><<<unsigned(L.n), 64, 0, s>>>(desc, L.n, L.ref_blocks, bt->partials,
                                                                  bt->results, 0, bt->tol, bt->active);
        AFSI_LAUNCH_CHECK();
    }
    return OK;
}

// stats of vector v into results slot, optionally followed by mean removal
static int batch_vec(const afsi_mgb_batch* bt, const BatchLayout& L, int v, int slot, bool sub,
                     hipStream_t s) {
    const VecDesc* desc = at<VecDesc>(bt->workspace, L.vec_off[v]);
    vec_partials_kernel<<<unsigned(L.vec_blocks[v]), 256, 0, s>>>(desc, L.n, bt->partials);
    batch_reduce_kernel<VecDesc><<<unsigned(L.n), 64, 0, s>>>(desc, L.n, L.vec_blocks[v], bt->partials,
                                                              bt->results, slot, bt->tol, bt->active);
    if (sub) vec_sub_mean_kernel<<<unsigned(L.vec_blocks[v]), 256, 0, s>>>(desc, L.n, bt->results, slot);
    AFSI_LAUNCH_CHECK();
    return OK;
}

static int batch_check_ws(const afsi_mgb_batch* bt, BatchLayout& L) {
    if (int r = batch_layout(bt, L)) return r;
    AFSI_REQUIRE(bt->workspace && bt->workspace_bytes >= L.total, "batch workspace too small");
    AFSI_REQUIRE(bt->active && bt->tol && bt->results && bt->partials, "missing batch state buffers");
    return OK;
}

}  // namespace afsi

extern "C" {
int64_t afsi_mg3d_partials_doubles(const int64_t n[3]) {
    if (!n || n[0] <= 0 || n[1] <= 0 || n[2] <= 0) return -1;
    int64_t blocks = afsi::cdiv(n[0] * n[1] * n[2], 256);
    if (blocks > 4096) blocks = 4096;
    return 4 * blocks;
}
int afsi_refine_residual_3d(const double* x, const double* b, float* r32, const int64_t n[3],
                            const double inv_h2[3], int periodic, double* partials,
                            int64_t partials_capacity, double* result, void* stream) {
    using namespace afsi;
    AFSI_REQUIRE(n && inv_h2 && x && b && r32 && partials && result, "invalid refinement arguments");
    AFSI_REQUIRE(n[0] > 0 && n[1] > 0 && n[2] > 0 && n[0] * n[1] * n[2] < (int64_t(1) << 31),
                 "invalid grid");
    const Dim32 d{int(n[0]), int(n[1]), int(n[2])};
    W3<double> k;
    for (int a = 0; a < 3; ++a) k.w[a] = inv_h2[a];
    int64_t nb = cdiv(n[0] * n[1], 4);
    if (nb > partials_capacity / 4) nb = partials_capacity / 4;
    AFSI_REQUIRE(nb >= 1, "reduction scratch too small");
    const hipStream_t s = hipStream_t(stream);
    if (periodic) refine_residual_kernel<true><<<unsigned(nb), 256, 0, s>>>(x, b, r32, d, k, partials);
    else refine_residual_kernel<false><<<unsigned(nb), 256, 0, s>>>(x, b, r32, d, k, partials);
    AFSI_LAUNCH_CHECK();
    return launch_reduce_final(partials, nb, result, s);
}
int afsi_refine_step_3d(const double* x, const float* e, const double* b, double* x_out,
                        float* r32, const int64_t n[3], const double inv_h2[3], int periodic,
                        double* partials, int64_t partials_capacity, double* result, void* stream) {
    using namespace afsi;
    AFSI_REQUIRE(n && inv_h2 && x && e && b && x_out && r32, "invalid refinement arguments");
    AFSI_REQUIRE(x != x_out, "the refinement step needs distinct input and output iterates");
    AFSI_REQUIRE(n[0] > 0 && n[1] > 0 && n[2] > 0 && n[0] * n[1] * n[2] < (int64_t(1) << 31),
                 "invalid grid");
    const Dim32 d{int(n[0]), int(n[1]), int(n[2])};
    W3<double> k;
    for (int a = 0; a < 3; ++a) k.w[a] = inv_h2[a];
    const hipStream_t s = hipStream_t(stream);
    int64_t nb = cdiv(n[0] * n[1], 4);
    if (result) {
        AFSI_REQUIRE(partials, "missing reduction scratch");
        if (nb > partials_capacity / 4) nb = partials_capacity / 4;
        AFSI_REQUIRE(nb >= 1, "reduction scratch too small");
        if (periodic) refine_step_kernel<true, true><<<unsigned(nb), 256, 0, s>>>(x, e, b, x_out, r32, d, k, partials);
        else refine_step_kernel<false, true><<<unsigned(nb), 256, 0, s>>>(x, e, b, x_out, r32, d, k, partials);
        AFSI_LAUNCH_CHECK();
        return launch_reduce_final(partials, nb, result, s);
    }
    if (nb > 16384) nb = 16384;
    if (periodic) refine_step_kernel<true, false><<<unsigned(nb), 256, 0, s>>>(x, e, b, x_out, r32, d, k, nullptr);
    else refine_step_kernel<false, false><<<unsigned(nb), 256, 0, s>>>(x, e, b, x_out, r32, d, k, nullptr);
    AFSI_LAUNCH_CHECK();
    return OK;
}
int afsi_refine_update(double* x, const float* e, int64_t count, void* stream) {
    using namespace afsi;
    AFSI_REQUIRE(count >= 0 && (count == 0 || (x && e)), "invalid refinement update");
    if (count == 0) return OK;
    refine_update_kernel<<<grid_for(count, 256, 16384), 256, 0, hipStream_t(stream)>>>(x, e, count);
    AFSI_LAUNCH_CHECK();
    return OK;
}
size_t afsi_mgb_workspace_bytes(const afsi_mgb_batch* batch) {
    afsi::BatchLayout L;
    return afsi::batch_layout(batch, L) == afsi::OK ? L.total : 0;
}
int64_t afsi_mgb_partials_doubles(const afsi_mgb_batch* batch) {
    afsi::BatchLayout L;
    return afsi::batch_layout(batch, L) == afsi::OK ? L.partials : -1;
}
int afsi_mgb_prepare(const afsi_mgb_batch* batch, void* stream) {
    return afsi::batch_prepare(batch, hipStream_t(stream));
}
int afsi_mgb_begin_at(const afsi_mgb_batch* batch, int32_t start, double rtol, double atol, void* stream) {
    using namespace afsi;
    BatchLayout L;
    if (int r = batch_check_ws(batch, L)) return r;
    AFSI_REQUIRE(start == 0 || start == 1, "invalid iterate index");
    const hipStream_t s = hipStream_t(stream);
    if (int r = batch_vec(batch, L, 0, 4, true, s)) return r;          // b: stats, mean removal
    if (int r = batch_vec(batch, L, 1 + start, 12, true, s)) return r;  // x[start]: stats, mean removal
    if (int r = batch_vec(batch, L, 0, 8, false, s)) return r;         // ||b|| after the removal
    batch_begin_kernel<<<unsigned(cdiv(L.n, 256)), 256, 0, s>>>(L.n, batch->results, batch->tol,
                                                               batch->active, rtol, atol);
    AFSI_LAUNCH_CHECK();
    const int which = start == 0 ? 0 : 3;  // residual of x[start]
    return L.periodic ? batch_refine<true>(batch, L, which, true, s) : batch_refine<false>(batch, L, which, true, s);
}
int afsi_mgb_begin(const afsi_mgb_batch* batch, double rtol, double atol, void* stream) {
    return afsi_mgb_begin_at(batch, 0, rtol, atol, stream);
}
int afsi_mgb_iterate(const afsi_mgb_batch* batch, int32_t start, int32_t count, void* stream) {
    using namespace afsi;
    BatchLayout L;
    if (int r = batch_check_ws(batch, L)) return r;
    AFSI_REQUIRE((start == 0 || start == 1) && count >= 1, "invalid batch iteration range");
    const hipStream_t s = hipStream_t(stream);
    const int inner = batch->inner_cycles > 0 ? batch->inner_cycles : 1;
    for (int it = 0; it < count; ++it) {
        const int from = (start + it) % 2;
        const bool last = it == count - 1;
        for (int c = 0; c < inner; ++c) {
            const int r = L.periodic ? batch_vcycle<true>(batch, L, s, c == 0) : batch_vcycle<false>(batch, L, s, c == 0);
            if (r) return r;
        }
        if (L.periodic) {
            if (int r = batch_refine<true>(batch, L, 1 + from, last, s)) return r;
        } else {
            if (int r = batch_refine<false>(batch, L, 1 + from, last, s)) return r;
        }
    }
    return OK;
}
int afsi_mgb_finish(const afsi_mgb_batch* batch, int32_t cur, void* stream) {
    using namespace afsi;
    BatchLayout L;
    if (int r = batch_check_ws(batch, L)) return r;
    AFSI_REQUIRE(cur == 0 || cur == 1, "invalid iterate index");
    return batch_vec(batch, L, 1 + cur, 12, true, hipStream_t(stream));
}
int afsi_mg3d_sweep_f64(const afsi_mg3d_plan* plan, int32_t level, void* stream) {
    using namespace afsi;
    if (int r = fast_check<double>(plan)) return r;
    AFSI_REQUIRE(level >= 0 && level < plan->levels, "invalid level");
    const FastRunner<double> run{*plan, hipStream_t(stream), choose_coarse(*plan)};
    return run.sweep(level, run.P(level), run.W(level), false);
}
int afsi_mg3d_sweep_f32(const afsi_mg3d_plan* plan, int32_t level, void* stream) {
    using namespace afsi;
    if (int r = fast_check<float>(plan)) return r;
    AFSI_REQUIRE(level >= 0 && level < plan->levels, "invalid level");
    const FastRunner<float> run{*plan, hipStream_t(stream), choose_coarse(*plan)};
    return run.sweep(level, run.P(level), run.W(level), false);
}
int afsi_mg3d_supported(const afsi_mg3d_plan* plan) {
    return afsi::fast_check<double>(plan) == afsi::OK ? 1 : 0;
}
int afsi_mg3d_cycles_f64(const afsi_mg3d_plan* plan, int32_t cycles, void* stream) {
    return afsi::fast_cycles<double>(plan, cycles, stream);
}
int afsi_mg3d_cycles_f32(const afsi_mg3d_plan* plan, int32_t cycles, void* stream) {
    return afsi::fast_cycles<float>(plan, cycles, stream);
}
int afsi_mg3d_residual_f64(const afsi_mg3d_plan* plan, void* stream) {
    return afsi::fast_residual<double>(plan, stream);
}
int afsi_mg3d_residual_f32(const afsi_mg3d_plan* plan, void* stream) {
    return afsi::fast_residual<float>(plan, stream);
}
}  // extern "C"
