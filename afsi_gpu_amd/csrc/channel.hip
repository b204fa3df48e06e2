// 2D staggered channel on gfx950: wall-reflected 16-link stencil, 5-point
// pressure operator, multigrid kernels and a native V-cycle driver.
//
// Two boundary variants (template parameter P):
//   P = false  the reference channel (mac2d/grid.py:61-71): zero normal
//              derivative at the inlet (x = 0) and the walls (y = 0, Ly),
//              p = 0 half a cell beyond the outlet (+2p/hx^2 on the last
//              row); non-singular, no mean removal (mac2d/multigrid.py).
//   P = true   periodic in x, walls in y (build-own); singular, so the
//              driver removes means as the 3D closed box does
//              (mac/multigrid.py:107-123).
// Cell (x, y) of an nx x ny level is element x*ny + y. Every elementwise
// kernel evaluates the reference's expression in the reference's order
// (mac2d/_triton_mg.py:7-75), so results are bit-identical to the
// restatement in FP64 (compiled with -ffp-contract=off).
#include "binshare.hpp"
#include <cstring>
#include <vector>

namespace afsi {

namespace {

constexpr int CH_THREADS = 256;

struct Dim2 {
    int nx, ny;
};

template <class T>
struct Coef2 {
    T hx2, hy2;    // h^2 (operator divisions)
    T ihx2, ihy2;  // 1/h^2 (diagonal)
    T rx, ry;      // exact 1/h^2 when h^2 is a power of two, else 0
    __host__ __device__ bool exact() const { return rx != T(0) && ry != T(0); }
};

// 1/h when h is a positive power of two between 2^-100 and 2^100 (normal
// in FP32 and FP64 alike, and so is 1/h), else 0. For such h, a / h and
// a * (1/h) are the correctly rounded values of the same real number, so
// they are equal for every a: the operator's divisions by h^2 become
// multiplications without changing a bit.
__host__ __device__ inline double exact_recip(double h) {
    const uint64_t b = __builtin_bit_cast(uint64_t, h);
    const int64_t e = int64_t((b >> 52) & 0x7FF);
    if ((b & ((uint64_t(1) << 52) - 1)) != 0 || (b >> 63) != 0 || e < 1023 - 100 || e > 1023 + 100) return 0.0;
    return __builtin_bit_cast(double, uint64_t(2046 - e) << 52);
}

template <class T>
__host__ __device__ inline Coef2<T> coef2_of(double hx2, double hy2, double ihx2, double ihy2) {
    return Coef2<T>{T(hx2), T(hy2), T(ihx2), T(ihy2), T(exact_recip(hx2)), T(exact_recip(hy2))};
}

// a / h2, or a * (1/h2) when M (h2 a power of two: the same value).
template <bool M, class T>
__device__ inline T ch_quot(T a, T h2, T r) {
    if constexpr (M) return a * r;
    else return a / h2;
}

// -A p at cell (x, y), mac2d/_triton_mg.py:7-21 order: -(right diff)
// + (left diff) per axis, then the outlet term.
template <class T, bool P, bool M = false>
__device__ inline T ch_apply(const T* __restrict__ p, int x, int y, Dim2 d, const Coef2<T>& k,
                             T& centre) {
    const int i = x * d.ny + y;
    const T c = p[i];
    centre = c;
    T a = T(0);
    if (P || x + 1 < d.nx) a -= ch_quot<M>(p[x + 1 < d.nx ? i + d.ny : y] - c, k.hx2, k.rx);
    if (P || x > 0) a += ch_quot<M>(c - p[x > 0 ? i - d.ny : (d.nx - 1) * d.ny + y], k.hx2, k.rx);
    if (y + 1 < d.ny) a -= ch_quot<M>(p[i + 1] - c, k.hy2, k.ry);
    if (y > 0) a += ch_quot<M>(c - p[i - 1], k.hy2, k.ry);
    if (!P && x == d.nx - 1) a += ch_quot<M>(T(2) * c, k.hx2, k.rx);
    return a;
}

// Diagonal of A, mac2d/multigrid.py:27-33 order (2/hx^2 == 2*(1/hx^2)).
template <class T, bool P>
__device__ inline T ch_diag(int x, int y, Dim2 d, const Coef2<T>& k) {
    T g = T(0);
    if (P || x + 1 < d.nx) g += k.ihx2;
    if (P || x > 0) g += k.ihx2;
    if (y + 1 < d.ny) g += k.ihy2;
    if (y > 0) g += k.ihy2;
    if (!P && x == d.nx - 1) g += T(2) * k.ihx2;
    return g;
}

// Mode 0: out = rhs - A p. Mode 1: out = p + (omega*(rhs - A p))/diag.
template <class T, bool P, int Mode, bool M>
__global__ void __launch_bounds__(CH_THREADS) ch_stencil_kernel(const T* __restrict__ p,
                                                                const T* __restrict__ rhs,
                                                                T* __restrict__ out, Dim2 d,
                                                                Coef2<T> k, T omega) {
    const int n = d.nx * d.ny;
    for (int i = blockIdx.x * CH_THREADS + threadIdx.x; i < n; i += gridDim.x * CH_THREADS) {
        const int x = i / d.ny, y = i - x * d.ny;
        T c;
        const T r = rhs[i] - ch_apply<T, P, M>(p, x, y, d, k, c);
        out[i] = Mode == 0 ? r : c + (omega * r) / ch_diag<T, P>(x, y, d, k);
    }
}

// One colour of a red-black Gauss-Seidel sweep, in place: cells with
// (x + y) % 2 == color get p += (rhs - A p)/diag (their neighbours have the
// other colour, so no cell read here is written here).
template <class T, bool P, bool M>
__global__ void __launch_bounds__(CH_THREADS) ch_rbgs_kernel(T* __restrict__ p,
                                                             const T* __restrict__ rhs, Dim2 d,
                                                             Coef2<T> k, int color) {
    // ny even: row x holds ny/2 cells of this colour at y = 2j + ((x + color) & 1)
    const int hy = d.ny / 2, half = d.nx * hy;
    for (int e = blockIdx.x * CH_THREADS + threadIdx.x; e < half; e += gridDim.x * CH_THREADS) {
        const int x = e / hy, y = 2 * (e - x * hy) + ((x + color) & 1);
        const int i = x * d.ny + y;
        T c;
        const T r = rhs[i] - ch_apply<T, P, M>(p, x, y, d, k, c);
        p[i] = c + r / ch_diag<T, P>(x, y, d, k);
    }
}

// Fallback colour pass for odd ny: one thread per cell.
template <class T, bool P, bool M>
__global__ void __launch_bounds__(CH_THREADS) ch_rbgs_any_kernel(T* __restrict__ p,
                                                                 const T* __restrict__ rhs,
                                                                 Dim2 d, Coef2<T> k, int color) {
    const int n = d.nx * d.ny;
    for (int i = blockIdx.x * CH_THREADS + threadIdx.x; i < n; i += gridDim.x * CH_THREADS) {
        const int x = i / d.ny, y = i - x * d.ny;
        if (((x + y) & 1) != color) continue;
        T c;
        const T r = rhs[i] - ch_apply<T, P, M>(p, x, y, d, k, c);
        p[i] = c + r / ch_diag<T, P>(x, y, d, k);
    }
}

// coarse = (((r00 + r01) + r10) + r11) * 0.25 with r = rhs - A p on the fine
// level (mac2d/_triton_mg.py:40-51); d is the fine level.
template <class T, bool P, bool M>
__global__ void __launch_bounds__(CH_THREADS) ch_residual_restrict_kernel(
    const T* __restrict__ p, const T* __restrict__ rhs, T* __restrict__ coarse, Dim2 d,
    Coef2<T> k) {
    const int cy = d.ny / 2, n = (d.nx / 2) * cy;
    for (int i = blockIdx.x * CH_THREADS + threadIdx.x; i < n; i += gridDim.x * CH_THREADS) {
        const int X = i / cy, Y = i - X * cy;
        const int x = 2 * X, y = 2 * Y, j = x * d.ny + y;
        T c;
        const T r00 = rhs[j] - ch_apply<T, P, M>(p, x, y, d, k, c);
        const T r01 = rhs[j + 1] - ch_apply<T, P, M>(p, x, y + 1, d, k, c);
        const T r10 = rhs[j + d.ny] - ch_apply<T, P, M>(p, x + 1, y, d, k, c);
        const T r11 = rhs[j + d.ny + 1] - ch_apply<T, P, M>(p, x + 1, y + 1, d, k, c);
        coarse[i] = (((r00 + r01) + r10) + r11) * T(0.25);
    }
}

// Coarse value with ghosts: clamped (even) at the inlet and walls, negated
// (odd) beyond the outlet; periodic: wrapped in x.
template <class T, bool P>
__device__ inline T ch_coarse_value(const T* __restrict__ c, int x, int y, int NX, int NY) {
    const int yy = y < 0 ? 0 : (y >= NY ? NY - 1 : y);
    if (P) {
        const int xx = x < 0 ? x + NX : (x >= NX ? x - NX : x);
        return c[xx * NY + yy];
    }
    const T sign = x >= NX ? T(-1) : T(1);
    const int xx = x < 0 ? 0 : (x >= NX ? NX - 1 : x);
    return sign * c[xx * NY + yy];
}

// fine += bilinear prolongation of coarse (mac2d/_triton_mg.py:54-75);
// d is the fine level.
template <class T, bool P>
__global__ void __launch_bounds__(CH_THREADS) ch_prolong_add_kernel(const T* __restrict__ coarse,
                                                                    T* __restrict__ fine, Dim2 d) {
    const int n = d.nx * d.ny, NX = d.nx / 2, NY = d.ny / 2;
    for (int i = blockIdx.x * CH_THREADS + threadIdx.x; i < n; i += gridDim.x * CH_THREADS) {
        const int x = i / d.ny, y = i - x * d.ny;
        const int ax = x / 2 - 1 + (x & 1), ay = y / 2 - 1 + (y & 1);
        const T wx = (x & 1) ? T(0.25) : T(0.75), wy = (y & 1) ? T(0.25) : T(0.75);
        const T c00 = ch_coarse_value<T, P>(coarse, ax, ay, NX, NY);
        const T c01 = ch_coarse_value<T, P>(coarse, ax, ay + 1, NX, NY);
        const T c10 = ch_coarse_value<T, P>(coarse, ax + 1, ay, NX, NY);
        const T c11 = ch_coarse_value<T, P>(coarse, ax + 1, ay + 1, NX, NY);
        const T v = (c00 * (T(1) - wy) + c01 * wy) * (T(1) - wx) + (c10 * (T(1) - wy) + c11 * wy) * wx;
        fine[i] = fine[i] + v;
    }
}

// out = scale * ((u[x+1] - u[x])/hx + (v[y+1] - v[y])/hy)
// (mac2d/grid.py:46-48, then the caller's -rho/dt).
template <bool P>
__global__ void __launch_bounds__(CH_THREADS) ch_divergence_kernel(const double* __restrict__ u,
                                                                   const double* __restrict__ v,
                                                                   double* __restrict__ out,
                                                                   Dim2 d, double hx, double hy,
                                                                   double scale) {
    const int n = d.nx * d.ny;
    for (int i = blockIdx.x * CH_THREADS + threadIdx.x; i < n; i += gridDim.x * CH_THREADS) {
        const int x = i / d.ny, y = i - x * d.ny;
        const int xr = P ? (x + 1 < d.nx ? x + 1 : 0) : x + 1;
        const double du = (u[xr * d.ny + y] - u[i]) / hx;
        const double dv = (v[x * (d.ny + 1) + y + 1] - v[x * (d.ny + 1) + y]) / hy;
        out[i] = scale * (du + dv);
    }
}

// u -= coef*gx, v -= coef*gy with the channel gradient (mac2d/grid.py:51-58:
// gx[0] = 0, gx[nx] = -2 p[nx-1]/hx, gy = 0 on the walls) or the periodic
// one (gx[x] = (p[x] - p[x-1 mod nx])/hx). Faces whose gradient is zero
// keep their value exactly.
template <bool P>
__global__ void __launch_bounds__(CH_THREADS) ch_grad_correct_kernel(double* __restrict__ u,
                                                                     double* __restrict__ v,
                                                                     const double* __restrict__ p,
                                                                     Dim2 d, double hx, double hy,
                                                                     double coef) {
    const int nu = (P ? d.nx : d.nx + 1) * d.ny, nv = d.nx * (d.ny + 1);
    for (int i = blockIdx.x * CH_THREADS + threadIdx.x; i < nu + nv; i += gridDim.x * CH_THREADS) {
        if (i < nu) {
            const int x = i / d.ny, y = i - x * d.ny;
            double g;
            if (P) g = (p[x * d.ny + y] - p[(x > 0 ? x - 1 : d.nx - 1) * d.ny + y]) / hx;
            else if (x == 0) continue;
            else if (x == d.nx) g = -2 * p[(d.nx - 1) * d.ny + y] / hx;
            else g = (p[x * d.ny + y] - p[(x - 1) * d.ny + y]) / hx;
            u[i] = u[i] - coef * g;
        } else {
            const int j = i - nu, x = j / (d.ny + 1), y = j - x * (d.ny + 1);
            if (y == 0 || y == d.ny) continue;
            const double g = (p[x * d.ny + y] - p[x * d.ny + y - 1]) / hy;
            v[j] = v[j] - coef * g;
        }
    }
}

// Wall-reflected 16-link stencil of mac2d/transfer.py:61-81, one thread per
// (component, point): links k = 4 i + j at node (bx + i, by + j), weight
// phi(sx - node_x) * phi(sy - node_y); rows outside the walls reflect
// (u: j<0 -> -j-1, j>=ny -> 2ny-j-1; v: j<0 -> -j, j>ny -> 2ny-j) with the
// weight negated; periodic grids wrap the x node. Points outside the
// support region of mac2d/transfer.py:88-95 set flag bit 1.
__global__ void __launch_bounds__(CH_THREADS) ch_prepare_kernel(
    const double* __restrict__ xy, int64_t points, Dim2 d, double hx, double hy, double Lx,
    double Ly, int periodic, int64_t* __restrict__ ids_u, double* __restrict__ w_u,
    int64_t* __restrict__ ids_v, double* __restrict__ w_v, int* __restrict__ flags) {
    const int64_t total = 2 * points;
    for (int64_t t = int64_t(blockIdx.x) * CH_THREADS + threadIdx.x; t < total;
         t += int64_t(gridDim.x) * CH_THREADS) {
        const int c = int(t / points);
        const int64_t q = t - c * points;
        const double px = xy[2 * q], py = xy[2 * q + 1];
        bool ok = isfinite(px) && isfinite(py) && py >= -hy && py <= Ly + hy;
        if (!periodic) ok = ok && px >= 2 * hx && px <= Lx - 2 * hx;
        int64_t* ids = c == 0 ? ids_u : ids_v;
        double* w = c == 0 ? w_u : w_v;
        if (!ok) {
            raise_flag(flags, F_SUPPORT);
            for (int k = 0; k < 16; ++k) {
                ids[q * 16 + k] = 0;
                w[q * 16 + k] = 0.0;
            }
            continue;
        }
        const double ox = hx * (c == 0 ? 0.0 : 0.5), oy = hy * (c == 1 ? 0.0 : 0.5);
        const double sx = (px - ox) / hx, sy = (py - oy) / hy;
        const int64_t bx = int64_t(floor(sx - 1)), by = int64_t(floor(sy - 1));
        const int64_t fny = d.ny + (c == 1);
        double fx[4], fy[4];
        for (int a = 0; a < 4; ++a) {
            fx[a] = peskin_phi(sx - double(bx + a));
            fy[a] = peskin_phi(sy - double(by + a));
        }
        for (int i = 0; i < 4; ++i) {
            int64_t nxi = bx + i;
            if (periodic) nxi = ((nxi % d.nx) + d.nx) % d.nx;
            for (int j = 0; j < 4; ++j) {
                const int64_t row = by + j;
                bool outside;
                int64_t refl;
                if (c == 0) {
                    outside = row < 0 || row >= d.ny;
                    refl = row < 0 ? -row - 1 : (row >= d.ny ? 2 * d.ny - row - 1 : row);
                } else {
                    outside = row < 0 || row > d.ny;
                    refl = row < 0 ? -row : (row > d.ny ? 2 * d.ny - row : row);
                }
                const double wt = fx[i] * fy[j];
                ids[q * 16 + 4 * i + j] = nxi * fny + refl;
                w[q * 16 + 4 * i + j] = outside ? -wt : wt;
            }
        }
    }
}

bool dims2_ok(const int64_t n[2]) {
    return n && n[0] > 0 && n[1] > 0 && n[0] * n[1] < (int64_t(1) << 31);
}

template <class T>
Coef2<T> coef2(const double h2[2], const double inv_h2[2]) {
    return coef2_of<T>(h2[0], h2[1], inv_h2 ? inv_h2[0] : 0.0, inv_h2 ? inv_h2[1] : 0.0);
}

unsigned blocks_for(int64_t n) { return grid_for(n, CH_THREADS, 8192); }

template <class T, int Mode>
int ch_stencil(const T* p, const T* rhs, T* out, const int64_t n[2], const double h2[2],
               const double inv_h2[2], double omega, int periodic, hipStream_t s) {
    AFSI_REQUIRE(dims2_ok(n) && h2 && p && rhs && out && (Mode == 0 || inv_h2),
                 "invalid channel stencil arguments");
    const Dim2 d{int(n[0]), int(n[1])};
    const Coef2<T> k = coef2<T>(h2, inv_h2);
    const unsigned g = blocks_for(n[0] * n[1]);
    const bool m = k.exact();
    if (periodic && m) ch_stencil_kernel<T, true, Mode, true><<<g, CH_THREADS, 0, s>>>(p, rhs, out, d, k, T(omega));
    else if (periodic) ch_stencil_kernel<T, true, Mode, false><<<g, CH_THREADS, 0, s>>>(p, rhs, out, d, k, T(omega));
    else if (m) ch_stencil_kernel<T, false, Mode, true><<<g, CH_THREADS, 0, s>>>(p, rhs, out, d, k, T(omega));
    else ch_stencil_kernel<T, false, Mode, false><<<g, CH_THREADS, 0, s>>>(p, rhs, out, d, k, T(omega));
    AFSI_LAUNCH_CHECK();
    return OK;
}

template <class T>
int ch_rbgs(T* p, const T* rhs, const int64_t n[2], const double h2[2], const double inv_h2[2],
            int color, int periodic, hipStream_t s) {
    AFSI_REQUIRE(dims2_ok(n) && h2 && inv_h2 && p && rhs, "invalid channel smoothing arguments");
    const Dim2 d{int(n[0]), int(n[1])};
    const Coef2<T> k = coef2<T>(h2, inv_h2);
    const bool m = k.exact();
    if (d.ny % 2 == 0) {
        const unsigned g = blocks_for((n[0] * n[1] + 1) / 2);
        if (periodic && m) ch_rbgs_kernel<T, true, true><<<g, CH_THREADS, 0, s>>>(p, rhs, d, k, color);
        else if (periodic) ch_rbgs_kernel<T, true, false><<<g, CH_THREADS, 0, s>>>(p, rhs, d, k, color);
        else if (m) ch_rbgs_kernel<T, false, true><<<g, CH_THREADS, 0, s>>>(p, rhs, d, k, color);
        else ch_rbgs_kernel<T, false, false><<<g, CH_THREADS, 0, s>>>(p, rhs, d, k, color);
    } else {
        const unsigned g = blocks_for(n[0] * n[1]);
        if (periodic && m) ch_rbgs_any_kernel<T, true, true><<<g, CH_THREADS, 0, s>>>(p, rhs, d, k, color);
        else if (periodic) ch_rbgs_any_kernel<T, true, false><<<g, CH_THREADS, 0, s>>>(p, rhs, d, k, color);
        else if (m) ch_rbgs_any_kernel<T, false, true><<<g, CH_THREADS, 0, s>>>(p, rhs, d, k, color);
        else ch_rbgs_any_kernel<T, false, false><<<g, CH_THREADS, 0, s>>>(p, rhs, d, k, color);
    }
    AFSI_LAUNCH_CHECK();
    return OK;
}

template <class T>
int ch_residual_restrict(const T* p, const T* rhs, T* coarse, const int64_t n[2],
                         const double h2[2], int periodic, hipStream_t s) {
    AFSI_REQUIRE(dims2_ok(n) && h2 && p && rhs && coarse && n[0] % 2 == 0 && n[1] % 2 == 0,
                 "invalid channel restriction arguments");
    const Dim2 d{int(n[0]), int(n[1])};
    const Coef2<T> k = coef2<T>(h2, nullptr);
    const unsigned g = blocks_for(n[0] * n[1] / 4);
    const bool m = k.exact();
    if (periodic && m) ch_residual_restrict_kernel<T, true, true><<<g, CH_THREADS, 0, s>>>(p, rhs, coarse, d, k);
    else if (periodic) ch_residual_restrict_kernel<T, true, false><<<g, CH_THREADS, 0, s>>>(p, rhs, coarse, d, k);
    else if (m) ch_residual_restrict_kernel<T, false, true><<<g, CH_THREADS, 0, s>>>(p, rhs, coarse, d, k);
    else ch_residual_restrict_kernel<T, false, false><<<g, CH_THREADS, 0, s>>>(p, rhs, coarse, d, k);
    AFSI_LAUNCH_CHECK();
    return OK;
}

template <class T>
int ch_prolong(const T* coarse, T* fine, const int64_t n[2], int periodic, hipStream_t s) {
    AFSI_REQUIRE(dims2_ok(n) && coarse && fine && n[0] % 2 == 0 && n[1] % 2 == 0,
                 "invalid channel prolongation arguments");
    const Dim2 d{int(n[0]), int(n[1])};
    const unsigned g = blocks_for(n[0] * n[1]);
    if (periodic) ch_prolong_add_kernel<T, true><<<g, CH_THREADS, 0, s>>>(coarse, fine, d);
    else ch_prolong_add_kernel<T, false><<<g, CH_THREADS, 0, s>>>(coarse, fine, d);
    AFSI_LAUNCH_CHECK();
    return OK;
}

}  // namespace

// ---- whole V-cycles in one workgroup ----------------------------------------
// For grids up to CHB_MAX_CELLS cells the launches of a V-cycle (a few
// microseconds each) cost more than their work. One 1024-thread workgroup
// then runs `cycles` complete V-cycles: every phase of the multi-launch
// driver below becomes a block-strided loop over the same cells with the
// same per-cell expressions (ch_apply, ch_diag, ch_coarse_value; the coarse
// product with one wave per row and the same lane order and xor butterfly),
// separated by __syncthreads(), which orders the global-memory writes of one
// phase before the reads of the next within the workgroup. Results are the
// multi-launch driver's bit for bit, except the periodic mean removal, whose
// sum is taken in the workgroup's own fixed order.
constexpr int CHB_THREADS = 1024;
constexpr int64_t CHB_MAX_CELLS = 256 * 64;  // measured crossover: 2.05 vs 3.02 ms at 256x64, 6.4 vs 3.6 ms at 512x128
// A thread takes its cells of a phase in groups of CHB_ILP: every load of a
// group is issued before any of its stores, so the global-memory latencies
// of the group overlap (the compiler cannot move a load above a store to
// another array it may alias). Within a phase no cell reads what another
// cell of the phase writes (a colour reads only the other colour), so the
// values are those of the one-cell-at-a-time loop.
constexpr int CHB_ILP = 4;

template <class T>
__device__ void chb_remove_mean(T* x, int n, double* red) {
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += CHB_THREADS) v += double(x[i]);
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = CHB_THREADS / 2; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    const T mean = T(red[0] / double(n));
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += CHB_THREADS) x[i] = x[i] - mean;
    __syncthreads();
}

template <class T, bool P, bool M>
__device__ __forceinline__ void chb_smooth(const afsi_mg2d_plan& pl, int l, double* red) {
    const Dim2 d{int(pl.n[l][0]), int(pl.n[l][1])};
    const Coef2<T> k = coef2_of<T>(pl.h2[l][0], pl.h2[l][1], pl.inv_h2[l][0], pl.inv_h2[l][1]);
    T* p = static_cast<T*>(pl.p[l]);
    const T* rhs = static_cast<const T*>(pl.rhs[l]);
    const int n = d.nx * d.ny;
    if (pl.smoother == 1) {
        for (int it = 0; it < pl.nu; ++it)
            for (int color = 0; color < 2; ++color) {
                if (d.ny % 2 == 0) {
                    const int hy = d.ny / 2, half = d.nx * hy;
                    for (int e0 = threadIdx.x; e0 < half; e0 += CHB_THREADS * CHB_ILP) {
                        T val[CHB_ILP];
                        int at[CHB_ILP];
#pragma unroll
                        for (int g = 0; g < CHB_ILP; ++g) {
                            const int e = e0 + g * CHB_THREADS;
                            at[g] = -1;
                            if (e >= half) continue;
                            const int x = e / hy, y = 2 * (e - x * hy) + ((x + color) & 1);
                            const int i = x * d.ny + y;
                            T c;
                            const T r = rhs[i] - ch_apply<T, P, M>(p, x, y, d, k, c);
                            val[g] = c + r / ch_diag<T, P>(x, y, d, k);
                            at[g] = i;
                        }
#pragma unroll
                        for (int g = 0; g < CHB_ILP; ++g)
                            if (at[g] >= 0) p[at[g]] = val[g];
                    }
                } else {
                    for (int i0 = threadIdx.x; i0 < n; i0 += CHB_THREADS * CHB_ILP) {
                        T val[CHB_ILP];
                        int at[CHB_ILP];
#pragma unroll
                        for (int g = 0; g < CHB_ILP; ++g) {
                            const int i = i0 + g * CHB_THREADS;
                            at[g] = -1;
                            if (i >= n) continue;
                            const int x = i / d.ny, y = i - x * d.ny;
                            if (((x + y) & 1) != color) continue;
                            T c;
                            const T r = rhs[i] - ch_apply<T, P, M>(p, x, y, d, k, c);
                            val[g] = c + r / ch_diag<T, P>(x, y, d, k);
                            at[g] = i;
                        }
#pragma unroll
                        for (int g = 0; g < CHB_ILP; ++g)
                            if (at[g] >= 0) p[at[g]] = val[g];
                    }
                }
                __syncthreads();
            }
    } else {
        T* a = p;
        T* b = static_cast<T*>(pl.tmp[l]);
        const T omega = T(pl.omega);
        for (int it = 0; it < pl.nu; ++it) {
            for (int i0 = threadIdx.x; i0 < n; i0 += CHB_THREADS * CHB_ILP) {
                T val[CHB_ILP];
#pragma unroll
                for (int g = 0; g < CHB_ILP; ++g) {
                    const int i = i0 + g * CHB_THREADS;
                    if (i >= n) continue;
                    const int x = i / d.ny, y = i - x * d.ny;
                    T c;
                    const T r = rhs[i] - ch_apply<T, P, M>(a, x, y, d, k, c);
                    val[g] = c + (omega * r) / ch_diag<T, P>(x, y, d, k);
                }
#pragma unroll
                for (int g = 0; g < CHB_ILP; ++g)
                    if (i0 + g * CHB_THREADS < n) b[i0 + g * CHB_THREADS] = val[g];
            }
            __syncthreads();
            T* t = a; a = b; b = t;
        }
        if (a != p) {
            for (int i = threadIdx.x; i < n; i += CHB_THREADS) p[i] = a[i];
            __syncthreads();
        }
    }
    if (P) chb_remove_mean(p, n, red);
}

// Residual restriction of level l into rhs[l + 1] (the multi-launch
// ch_residual_restrict_kernel's expressions).
template <class T, bool P, bool M>
__device__ __forceinline__ void chb_restrict(const afsi_mg2d_plan& pl, int l) {
    const Dim2 d{int(pl.n[l][0]), int(pl.n[l][1])};
    const Coef2<T> k = coef2_of<T>(pl.h2[l][0], pl.h2[l][1], 0.0, 0.0);
    const T* p = static_cast<const T*>(pl.p[l]);
    const T* rhs = static_cast<const T*>(pl.rhs[l]);
    T* coarse = static_cast<T*>(pl.rhs[l + 1]);
    const int cy = d.ny / 2, nc = (d.nx / 2) * cy;
    for (int i0 = threadIdx.x; i0 < nc; i0 += CHB_THREADS * CHB_ILP) {
        T val[CHB_ILP];
#pragma unroll
        for (int g = 0; g < CHB_ILP; ++g) {
            const int i = i0 + g * CHB_THREADS;
            if (i >= nc) continue;
            const int X = i / cy, Y = i - X * cy;
            const int x = 2 * X, y = 2 * Y, j = x * d.ny + y;
            T c;
            const T r00 = rhs[j] - ch_apply<T, P, M>(p, x, y, d, k, c);
            const T r01 = rhs[j + 1] - ch_apply<T, P, M>(p, x, y + 1, d, k, c);
            const T r10 = rhs[j + d.ny] - ch_apply<T, P, M>(p, x + 1, y, d, k, c);
            const T r11 = rhs[j + d.ny + 1] - ch_apply<T, P, M>(p, x + 1, y + 1, d, k, c);
            val[g] = (((r00 + r01) + r10) + r11) * T(0.25);
        }
#pragma unroll
        for (int g = 0; g < CHB_ILP; ++g)
            if (i0 + g * CHB_THREADS < nc) coarse[i0 + g * CHB_THREADS] = val[g];
    }
}

// Whether level l's spacings are powers of two (exact multiplications).
__device__ inline bool chb_exact(const afsi_mg2d_plan& pl, int l) {
    return exact_recip(pl.h2[l][0]) != 0.0 && exact_recip(pl.h2[l][1]) != 0.0;
}

template <class T, bool P>
__device__ __forceinline__ void chb_smooth_level(const afsi_mg2d_plan& pl, int l, double* red) {
    if (chb_exact(pl, l)) chb_smooth<T, P, true>(pl, l, red);
    else chb_smooth<T, P, false>(pl, l, red);
}

// `cycles` V-cycles of levels l0..last (l0 > 0: the lower part of a cycle
// whose levels above l0 run as batched launches).
template <class T, bool P>
__device__ __forceinline__ void chb_cycles(const afsi_mg2d_plan& pl, int cycles, double* red, int l0 = 0) {
    const int last = pl.levels - 1;
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
    for (int cyc = 0; cyc < cycles; ++cyc) {
        for (int l = l0; l < last; ++l) {
            chb_smooth_level<T, P>(pl, l, red);
            if (chb_exact(pl, l)) chb_restrict<T, P, true>(pl, l);
            else chb_restrict<T, P, false>(pl, l);
            const Dim2 d{int(pl.n[l][0]), int(pl.n[l][1])};
            T* coarse = static_cast<T*>(pl.rhs[l + 1]);
            const int cy = d.ny / 2, nc = (d.nx / 2) * cy;
            __syncthreads();
            if (P) chb_remove_mean(coarse, nc, red);
            T* pc = static_cast<T*>(pl.p[l + 1]);
            for (int i = threadIdx.x; i < nc; i += CHB_THREADS) pc[i] = T(0);
            __syncthreads();
        }
        {   // coarsest: p = inverse * rhs, one wave per row
            const int m = int(pl.coarse_cells);
            const T* A = static_cast<const T*>(pl.coarse_inverse);
            const T* xv = static_cast<const T*>(pl.rhs[last]);
            T* y = static_cast<T*>(pl.p[last]);
            for (int r = wave; r < m; r += CHB_THREADS / WAVE) {
                T acc = T(0);
                for (int c = lane; c < m; c += WAVE) acc += A[int64_t(r) * m + c] * xv[c];
#pragma unroll
                for (int dd = 32; dd >= 1; dd >>= 1) acc += __shfl_xor(acc, dd, WAVE);
                if (lane == 0) y[r] = acc;
            }
            __syncthreads();
            if (P) chb_remove_mean(y, m, red);
        }
        for (int l = last - 1; l >= l0; --l) {
            const Dim2 d{int(pl.n[l][0]), int(pl.n[l][1])};
            const T* coarse = static_cast<const T*>(pl.p[l + 1]);
            T* fine = static_cast<T*>(pl.p[l]);
            const int n = d.nx * d.ny, NX = d.nx / 2, NY = d.ny / 2;
            for (int i0 = threadIdx.x; i0 < n; i0 += CHB_THREADS * CHB_ILP) {
                T val[CHB_ILP];
#pragma unroll
                for (int g = 0; g < CHB_ILP; ++g) {
                    const int i = i0 + g * CHB_THREADS;
                    if (i >= n) continue;
                    const int x = i / d.ny, y = i - x * d.ny;
                    const int ax = x / 2 - 1 + (x & 1), ay = y / 2 - 1 + (y & 1);
                    const T wx = (x & 1) ? T(0.25) : T(0.75), wy = (y & 1) ? T(0.25) : T(0.75);
                    const T c00 = ch_coarse_value<T, P>(coarse, ax, ay, NX, NY);
                    const T c01 = ch_coarse_value<T, P>(coarse, ax, ay + 1, NX, NY);
                    const T c10 = ch_coarse_value<T, P>(coarse, ax + 1, ay, NX, NY);
                    const T c11 = ch_coarse_value<T, P>(coarse, ax + 1, ay + 1, NX, NY);
                    const T v = (c00 * (T(1) - wy) + c01 * wy) * (T(1) - wx) + (c10 * (T(1) - wy) + c11 * wy) * wx;
                    val[g] = fine[i] + v;
                }
#pragma unroll
                for (int g = 0; g < CHB_ILP; ++g)
                    if (i0 + g * CHB_THREADS < n) fine[i0 + g * CHB_THREADS] = val[g];
            }
            __syncthreads();
            chb_smooth_level<T, P>(pl, l, red);
        }
    }
}

template <class T, bool P>
__global__ void __launch_bounds__(CHB_THREADS) ch_block_cycles_kernel(afsi_mg2d_plan pl, int cycles) {
    __shared__ double red[CHB_THREADS];
    chb_cycles<T, P>(pl, cycles, red);
}

// ---- native V-cycle driver ---------------------------------------------------

template <class T>
struct ChannelRunner {
    const afsi_mg2d_plan& pl;
    hipStream_t s;

    T* P(int l) const { return static_cast<T*>(pl.p[l]); }
    T* R(int l) const { return static_cast<T*>(pl.rhs[l]); }
    T* W(int l) const { return static_cast<T*>(pl.tmp[l]); }
    int64_t cells(int l) const { return pl.n[l][0] * pl.n[l][1]; }

    int remove_mean(T* x, int64_t count) const;

    int smooth(int l) const {
        if (pl.smoother == 1) {
            for (int k = 0; k < pl.nu; ++k)
                for (int color = 0; color < 2; ++color)
                    if (int r = ch_rbgs<T>(P(l), R(l), pl.n[l], pl.h2[l], pl.inv_h2[l], color,
                                           pl.periodic, s))
                        return r;
        } else {
            T* a = P(l);
            T* b = W(l);
            for (int k = 0; k < pl.nu; ++k) {
                if (int r = ch_stencil<T, 1>(a, R(l), b, pl.n[l], pl.h2[l], pl.inv_h2[l], pl.omega,
                                             pl.periodic, s))
                    return r;
                T* t = a; a = b; b = t;
            }
            if (a != P(l))
                AFSI_HIP(hipMemcpyAsync(P(l), a, size_t(cells(l)) * sizeof(T),
                                        hipMemcpyDeviceToDevice, s));
        }
        return pl.periodic ? remove_mean(P(l), cells(l)) : OK;
    }

    int cycle(int l) const {
        const int last = pl.levels - 1;
        if (l == last) {
            if (int r = dense_matvec(R(l), P(l))) return r;
            return pl.periodic ? remove_mean(P(l), cells(l)) : OK;
        }
        if (int r = smooth(l)) return r;
        if (int r = ch_residual_restrict<T>(P(l), R(l), R(l + 1), pl.n[l], pl.h2[l], pl.periodic, s))
            return r;
        if (pl.periodic)
            if (int r = remove_mean(R(l + 1), cells(l + 1))) return r;
        if (int r = fill(P(l + 1), cells(l + 1))) return r;
        if (int r = cycle(l + 1)) return r;
        if (int r = ch_prolong<T>(P(l + 1), P(l), pl.n[l], pl.periodic, s)) return r;
        return smooth(l);
    }

    int dense_matvec(const T* x, T* y) const;
    int fill(T* x, int64_t count) const;
};

template <>
int ChannelRunner<double>::remove_mean(double* x, int64_t count) const {
    if (int r = afsi_reduce_f64(x, count, pl.reduce_partials, pl.reduce_result, s)) return r;
    return afsi_sub_mean_f64(x, count, pl.reduce_result, s);
}
template <>
int ChannelRunner<float>::remove_mean(float* x, int64_t count) const {
    if (int r = afsi_reduce_f32(x, count, pl.reduce_partials, pl.reduce_result, s)) return r;
    return afsi_sub_mean_f32(x, count, pl.reduce_result, s);
}
template <>
int ChannelRunner<double>::dense_matvec(const double* x, double* y) const {
    return afsi_dense_matvec_f64(static_cast<const double*>(pl.coarse_inverse), x, y,
                                 pl.coarse_cells, s);
}
template <>
int ChannelRunner<float>::dense_matvec(const float* x, float* y) const {
    return afsi_dense_matvec_f32(static_cast<const float*>(pl.coarse_inverse), x, y,
                                 pl.coarse_cells, s);
}
template <>
int ChannelRunner<double>::fill(double* x, int64_t count) const {
    return afsi_fill_f64(x, count, 0.0, s);
}
template <>
int ChannelRunner<float>::fill(float* x, int64_t count) const {
    return afsi_fill_f32(x, count, 0.0, s);
}

template <class T>
static int channel_check(const afsi_mg2d_plan* pl) {
    AFSI_REQUIRE(pl && pl->levels >= 1 && pl->levels <= AFSI_MG_MAX_LEVELS, "invalid multigrid plan");
    AFSI_REQUIRE(pl->nu >= 1 && (pl->smoother == 0 || pl->smoother == 1), "invalid smoother");
    AFSI_REQUIRE(pl->reduce_partials && pl->reduce_result, "missing reduction scratch");
    for (int l = 0; l < pl->levels; ++l) {
        AFSI_REQUIRE(dims2_ok(pl->n[l]) && pl->p[l] && pl->rhs[l] && pl->tmp[l], "invalid level %d", l);
        if (l + 1 < pl->levels)
            for (int a = 0; a < 2; ++a)
                AFSI_REQUIRE(pl->n[l][a] == 2 * pl->n[l + 1][a], "levels %d/%d are not 2:1", l, l + 1);
    }
    const int last = pl->levels - 1;
    AFSI_REQUIRE(pl->coarse_inverse && pl->coarse_cells == pl->n[last][0] * pl->n[last][1] &&
                     pl->coarse_cells <= 512,
                 "coarse inverse does not match the coarsest level");
    AFSI_REQUIRE(4 * afsi_reduce_partials(pl->n[0][0] * pl->n[0][1]) <= pl->partials_capacity,
                 "reduction scratch too small");
    return OK;
}

template <class T>
static int channel_cycles(const afsi_mg2d_plan* pl, int32_t cycles, void* stream) {
    if (int r = channel_check<T>(pl)) return r;
    AFSI_REQUIRE(cycles >= 0, "negative cycle count");
    if (pl->n[0][0] * pl->n[0][1] <= CHB_MAX_CELLS) {
        if (cycles == 0) return OK;
        if (pl->periodic) ch_block_cycles_kernel<T, true><<<1, CHB_THREADS, 0, hipStream_t(stream)>>>(*pl, cycles);
        else ch_block_cycles_kernel<T, false><<<1, CHB_THREADS, 0, hipStream_t(stream)>>>(*pl, cycles);
        AFSI_LAUNCH_CHECK();
        return OK;
    }
    const ChannelRunner<T> run{*pl, hipStream_t(stream)};
    for (int c = 0; c < cycles; ++c)
        if (int r = run.cycle(0)) return r;
    return OK;
}

template <class T>
static int channel_residual(const afsi_mg2d_plan* pl, void* stream) {
    if (int r = channel_check<T>(pl)) return r;
    const hipStream_t s = hipStream_t(stream);
    T* tmp = static_cast<T*>(pl->tmp[0]);
    if (int r = ch_stencil<T, 0>(static_cast<const T*>(pl->p[0]), static_cast<const T*>(pl->rhs[0]),
                                 tmp, pl->n[0], pl->h2[0], nullptr, 0.0, pl->periodic, s))
        return r;
    const int64_t count = pl->n[0][0] * pl->n[0][1];
    if constexpr (sizeof(T) == 8)
        return afsi_reduce_f64(tmp, count, pl->reduce_partials, pl->reduce_result, s);
    else
        return afsi_reduce_f32(tmp, count, pl->reduce_partials, pl->reduce_result, s);
}

}  // namespace afsi

using namespace afsi;

extern "C" {

int afsi_mg_residual_2d_f64(const double* p, const double* rhs, double* out, const int64_t n[2],
                            const double h2[2], int periodic, void* stream) {
    return ch_stencil<double, 0>(p, rhs, out, n, h2, nullptr, 0.0, periodic, hipStream_t(stream));
}
int afsi_mg_residual_2d_f32(const float* p, const float* rhs, float* out, const int64_t n[2],
                            const double h2[2], int periodic, void* stream) {
    return ch_stencil<float, 0>(p, rhs, out, n, h2, nullptr, 0.0, periodic, hipStream_t(stream));
}
int afsi_mg_jacobi_2d_f64(const double* p, const double* rhs, double* out, const int64_t n[2],
                          const double h2[2], const double inv_h2[2], double omega, int periodic,
                          void* stream) {
    return ch_stencil<double, 1>(p, rhs, out, n, h2, inv_h2, omega, periodic, hipStream_t(stream));
}
int afsi_mg_jacobi_2d_f32(const float* p, const float* rhs, float* out, const int64_t n[2],
                          const double h2[2], const double inv_h2[2], double omega, int periodic,
                          void* stream) {
    return ch_stencil<float, 1>(p, rhs, out, n, h2, inv_h2, omega, periodic, hipStream_t(stream));
}
int afsi_mg_rbgs_2d_f64(double* p, const double* rhs, const int64_t n[2], const double h2[2],
                        const double inv_h2[2], int periodic, void* stream) {
    for (int color = 0; color < 2; ++color)
        if (int r = ch_rbgs<double>(p, rhs, n, h2, inv_h2, color, periodic, hipStream_t(stream)))
            return r;
    return OK;
}
int afsi_mg_rbgs_2d_f32(float* p, const float* rhs, const int64_t n[2], const double h2[2],
                        const double inv_h2[2], int periodic, void* stream) {
    for (int color = 0; color < 2; ++color)
        if (int r = ch_rbgs<float>(p, rhs, n, h2, inv_h2, color, periodic, hipStream_t(stream)))
            return r;
    return OK;
}
int afsi_mg_residual_restrict_2d_f64(const double* p, const double* rhs, double* coarse,
                                     const int64_t n[2], const double h2[2], int periodic,
                                     void* stream) {
    return ch_residual_restrict<double>(p, rhs, coarse, n, h2, periodic, hipStream_t(stream));
}
int afsi_mg_residual_restrict_2d_f32(const float* p, const float* rhs, float* coarse,
                                     const int64_t n[2], const double h2[2], int periodic,
                                     void* stream) {
    return ch_residual_restrict<float>(p, rhs, coarse, n, h2, periodic, hipStream_t(stream));
}
int afsi_mg_prolong_add_2d_f64(const double* coarse, double* fine, const int64_t n[2], int periodic,
                               void* stream) {
    return ch_prolong<double>(coarse, fine, n, periodic, hipStream_t(stream));
}
int afsi_mg_prolong_add_2d_f32(const float* coarse, float* fine, const int64_t n[2], int periodic,
                               void* stream) {
    return ch_prolong<float>(coarse, fine, n, periodic, hipStream_t(stream));
}

int afsi_divergence_2d_f64(const double* u, const double* v, double* out, const int64_t n[2],
                           const double h[2], double scale, int periodic, void* stream) {
    AFSI_REQUIRE(dims2_ok(n) && h && u && v && out, "invalid divergence arguments");
    const Dim2 d{int(n[0]), int(n[1])};
    const hipStream_t s = hipStream_t(stream);
    const unsigned g = blocks_for(n[0] * n[1]);
    if (periodic) ch_divergence_kernel<true><<<g, CH_THREADS, 0, s>>>(u, v, out, d, h[0], h[1], scale);
    else ch_divergence_kernel<false><<<g, CH_THREADS, 0, s>>>(u, v, out, d, h[0], h[1], scale);
    AFSI_LAUNCH_CHECK();
    return OK;
}

int afsi_grad_correct_2d_f64(double* u, double* v, const double* p, const int64_t n[2],
                             const double h[2], double coef, int periodic, void* stream) {
    AFSI_REQUIRE(dims2_ok(n) && h && u && v && p, "invalid gradient arguments");
    const Dim2 d{int(n[0]), int(n[1])};
    const hipStream_t s = hipStream_t(stream);
    const unsigned g = blocks_for((n[0] + 1) * n[1] + n[0] * (n[1] + 1));
    if (periodic) ch_grad_correct_kernel<true><<<g, CH_THREADS, 0, s>>>(u, v, p, d, h[0], h[1], coef);
    else ch_grad_correct_kernel<false><<<g, CH_THREADS, 0, s>>>(u, v, p, d, h[0], h[1], coef);
    AFSI_LAUNCH_CHECK();
    return OK;
}

int afsi_channel_prepare_f64(const double* x, int64_t points, const afsi_grid2* grid,
                             int64_t* ids_u, double* w_u, int64_t* ids_v, double* w_v, int* flags,
                             void* stream) {
    AFSI_REQUIRE(grid && dims2_ok(grid->n) && points >= 0 && flags, "invalid channel prepare arguments");
    AFSI_REQUIRE(grid->n[0] >= 4 && grid->n[1] >= 4, "channel grid needs >= 4 cells per axis");
    if (points == 0) return OK;
    AFSI_REQUIRE(x && ids_u && w_u && ids_v && w_v, "invalid channel prepare arguments");
    const Dim2 d{int(grid->n[0]), int(grid->n[1])};
    ch_prepare_kernel<<<blocks_for(2 * points), CH_THREADS, 0, hipStream_t(stream)>>>(
        x, points, d, grid->h[0], grid->h[1], grid->lengths[0], grid->lengths[1], grid->periodic,
        ids_u, w_u, ids_v, w_v, flags);
    AFSI_LAUNCH_CHECK();
    return OK;
}

int afsi_mg2d_cycles_f64(const afsi_mg2d_plan* plan, int32_t cycles, void* stream) {
    return channel_cycles<double>(plan, cycles, stream);
}
int afsi_mg2d_cycles_f32(const afsi_mg2d_plan* plan, int32_t cycles, void* stream) {
    return channel_cycles<float>(plan, cycles, stream);
}
int afsi_mg2d_residual_f64(const afsi_mg2d_plan* plan, void* stream) {
    return channel_residual<double>(plan, stream);
}
int afsi_mg2d_residual_f32(const afsi_mg2d_plan* plan, void* stream) {
    return channel_residual<float>(plan, stream);
}

}  // extern "C"

// ---- batched channel instances (build-own) ------------------------------------
// Many channels of uneven size advanced together: one launch per phase over
// all instances. Face arrays are concatenated (instance i's u faces start
// at u_off, its v faces at v_off); each instance's pressure hierarchy is its
// own plan. Stencils carry global face ids, so one indexed spread/gather
// per component serves the whole batch.

namespace afsi {
namespace {

struct ChbDesc {
    const double* x;   // (points, 2)
    int64_t points, point_begin;
    Dim2 d;
    double hx, hy, Lx, Ly;
    int32_t periodic, reserved;
    int64_t u_off, v_off, u_count, v_count, cell_count;
};

__device__ inline int chb_find(const ChbDesc* __restrict__ desc, int n, int64_t j, int which) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        const int64_t b = which == 0 ? desc[mid].point_begin : (which == 1 ? desc[mid].u_off : desc[mid].v_off);
        if (b <= j) lo = mid;
        else hi = mid;
    }
    return lo;
}

// The 16-link reflected stencil of every point (ch_prepare_kernel's
// expressions) with the instance's face offset added to the ids. A
// workgroup takes CH_THREADS consecutive points of one component; each
// thread forms its point's 16 ids and weights in registers, and the block's
// contiguous 16 * CH_THREADS ids, then weights, pass through LDS (rows
// padded to 17 entries against bank conflicts) so that consecutive lanes
// store consecutive entries (a thread storing its own 128-byte row leaves
// every store instruction 64 scattered words).
constexpr int CHS_PAD = 17;

// A point's support check and its stencil base for component c (the
// reference's reflected 16-link stencil, mac2d/transfer.py:83-101).
__device__ inline bool chb_point_base(const ChbDesc& D, int c, double px, double py, double& sx, double& sy,
                                      int64_t& bx, int64_t& by) {
    bool ok = isfinite(px) && isfinite(py) && py >= -D.hy && py <= D.Ly + D.hy;
    if (!D.periodic) ok = ok && px >= 2 * D.hx && px <= D.Lx - 2 * D.hx;
    if (!ok) return false;
    const double ox = D.hx * (c == 0 ? 0.0 : 0.5), oy = D.hy * (c == 1 ? 0.0 : 0.5);
    sx = (px - ox) / D.hx;
    sy = (py - oy) / D.hy;
    bx = int64_t(floor(sx - 1));
    by = int64_t(floor(sy - 1));
    return true;
}

// Row `row` of component c's face lattice reflected at the walls; outside:
// the link's weight is negated.
__device__ inline int64_t chb_reflect(int c, int64_t ny, int64_t row, bool& outside) {
    if (c == 0) {
        outside = row < 0 || row >= ny;
        return row < 0 ? -row - 1 : (row >= ny ? 2 * ny - row - 1 : row);
    }
    outside = row < 0 || row > ny;
    return row < 0 ? -row : (row > ny ? 2 * ny - row : row);
}

__global__ void __launch_bounds__(CH_THREADS) chb_stencil_kernel(const ChbDesc* __restrict__ desc, int n,
                                                                 int64_t total, int64_t* __restrict__ ids_u,
                                                                 double* __restrict__ w_u, int64_t* __restrict__ ids_v,
                                                                 double* __restrict__ w_v, int* __restrict__ flags) {
    __shared__ uint64_t stage[CH_THREADS * CHS_PAD];
    const int64_t per_comp = cdiv(total, CH_THREADS);
    for (int64_t blk = blockIdx.x; blk < 2 * per_comp; blk += gridDim.x) {
        const int c = int(blk / per_comp);
        const int64_t g0 = (blk - c * per_comp) * CH_THREADS;
        const int rows = int(total - g0 < CH_THREADS ? total - g0 : CH_THREADS);
        const int64_t g = g0 + threadIdx.x;
        int64_t ids[16];
        double w[16];
        if (int(threadIdx.x) < rows) {
            const ChbDesc& D = desc[chb_find(desc, n, g, 0)];
            const int64_t q = g - D.point_begin;
            const double px = D.x[2 * q], py = D.x[2 * q + 1];
            const Dim2 d = D.d;
            double sx, sy;
            int64_t bx, by;
            const bool ok = chb_point_base(D, c, px, py, sx, sy, bx, by);
            const int64_t off = c == 0 ? D.u_off : D.v_off;
            if (!ok) {
                raise_flag(flags, F_SUPPORT);
                for (int k = 0; k < 16; ++k) {
                    ids[k] = off;
                    w[k] = 0.0;
                }
            } else {
                const int64_t fny = d.ny + (c == 1);
                double fx[4], fy[4];
                for (int a = 0; a < 4; ++a) {
                    fx[a] = peskin_phi(sx - double(bx + a));
                    fy[a] = peskin_phi(sy - double(by + a));
                }
                for (int i = 0; i < 4; ++i) {
                    int64_t nxi = bx + i;
                    if (D.periodic) nxi = ((nxi % d.nx) + d.nx) % d.nx;
                    for (int j = 0; j < 4; ++j) {
                        bool outside;
                        const int64_t refl = chb_reflect(c, d.ny, by + j, outside);
                        const double wt = fx[i] * fy[j];
                        ids[4 * i + j] = off + nxi * fny + refl;
                        w[4 * i + j] = outside ? -wt : wt;
                    }
                }
            }
        }
        int64_t* out_ids = (c == 0 ? ids_u : ids_v) + g0 * 16;
        double* out_w = (c == 0 ? w_u : w_v) + g0 * 16;
        uint64_t* mine = stage + threadIdx.x * CHS_PAD;
        if (int(threadIdx.x) < rows)
            for (int k = 0; k < 16; ++k) mine[k] = uint64_t(ids[k]);
        __syncthreads();
        for (int e = threadIdx.x; e < rows * 16; e += CH_THREADS)
            out_ids[e] = int64_t(stage[(e >> 4) * CHS_PAD + (e & 15)]);
        __syncthreads();
        if (int(threadIdx.x) < rows)
            for (int k = 0; k < 16; ++k) mine[k] = __builtin_bit_cast(uint64_t, w[k]);
        __syncthreads();
        for (int e = threadIdx.x; e < rows * 16; e += CH_THREADS)
            out_w[e] = __builtin_bit_cast(double, stage[(e >> 4) * CHS_PAD + (e & 15)]);
        __syncthreads();
    }
}

// ---- tiled channel spread ---------------------------------------------------
// f_c[face] = sum over links of (force_c * w) for the whole batch, the links
// being the stencil kernel's (same ids and weights). The (point, component)
// pairs are stably sorted by the tile of their stencil base: tiles of
// CHS_T x CHS_T bases per instance and component, the x base wrapped on
// periodic channels, the y base offset by 3 (bases lie in [-3, ny]). One
// work item per (tile, chunk of <= CHS_CHUNK pairs) adds its pairs' 16 links
// into an FP64 LDS box of the tile's (CHS_T + 3)^2 unreflected face
// positions, then adds every nonzero entry to its reflected face in HBM with
// one atomic. As with the reference's index_add_, the order of the additions
// into a face is not fixed.
constexpr int CHS_T = 16;
constexpr int CHS_BOX = CHS_T + 3;
constexpr int CHS_CHUNK = 512;
constexpr int CHS_THREADS = 256;

struct ChsTile {
    int32_t begin, ntx, nty, reserved;
};

// x base of the key: wrapped on periodic channels (the box unwraps it).
__device__ inline int64_t chs_key_x(const ChbDesc& D, int64_t bx) {
    return D.periodic ? ((bx % D.d.nx) + D.d.nx) % D.d.nx : bx;
}

__global__ void chs_keys_kernel(const ChbDesc* __restrict__ desc, int n, int64_t total,
                                const ChsTile* __restrict__ tiles, int64_t space, uint32_t* __restrict__ keys,
                                uint32_t* __restrict__ vals) {
    for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < 2 * total;
         t += int64_t(gridDim.x) * blockDim.x) {
        const int c = int(t / total);
        const int64_t g = t - c * total;
        const int k = chb_find(desc, n, g, 0);
        const ChbDesc& D = desc[k];
        const int64_t q = g - D.point_begin;
        double sx, sy;
        int64_t bx, by;
        uint32_t key = uint32_t(space);  // unsupported points: sorted last, in no tile
        if (chb_point_base(D, c, D.x[2 * q], D.x[2 * q + 1], sx, sy, bx, by)) {
            const ChsTile T = tiles[2 * k + c];
            const int64_t kx = chs_key_x(D, bx), ky = by + 3;
            if (kx >= 0 && kx / CHS_T < T.ntx && ky >= 0 && ky / CHS_T < T.nty)
                key = uint32_t(T.begin + (kx / CHS_T) * T.nty + ky / CHS_T);
        }
        keys[t] = key;
        vals[t] = uint32_t(t);
    }
}

__global__ void __launch_bounds__(CHS_THREADS) chs_spread_kernel(
    const ChbDesc* __restrict__ desc, int n, const ChsTile* __restrict__ tiles, const int32_t* __restrict__ perm,
    const int32_t* __restrict__ tile_ptr, const int32_t* __restrict__ chunk_start, int64_t ntiles, int64_t total,
    const double* __restrict__ force, double* __restrict__ fu, double* __restrict__ fv, int* __restrict__ flags) {
    __shared__ double box[CHS_BOX * CHS_BOX];
    const int items = chunk_start[ntiles];
    for (int w = blockIdx.x; w < items; w += gridDim.x) {
        int lo = 0, hi = int(ntiles);
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (chunk_start[mid] <= w) lo = mid;
            else hi = mid;
        }
        const int t = lo;
        const int pbeg = tile_ptr[t] + (w - chunk_start[t]) * CHS_CHUNK;
        const int pend = min(tile_ptr[t + 1], pbeg + CHS_CHUNK);
        int a = 0, b = 2 * n;  // (instance, component) of tile t: last begin <= t
        while (b - a > 1) {
            const int mid = (a + b) >> 1;
            if (tiles[mid].begin <= t) a = mid;
            else b = mid;
        }
        const int inst = a / 2, c = a % 2;
        const ChbDesc& D = desc[inst];
        const ChsTile T = tiles[a];
        const int local = t - T.begin;
        const int64_t col0 = int64_t(local / T.nty) * CHS_T, row0 = int64_t(local % T.nty) * CHS_T - 3;
        for (int e = threadIdx.x; e < CHS_BOX * CHS_BOX; e += CHS_THREADS) box[e] = 0.0;
        __syncthreads();
        // lanes of one cell walk their 16 links from different rotations
        const int ri = (threadIdx.x >> 2) & 3, rj = threadIdx.x & 3;
        for (int j = pbeg + int(threadIdx.x); j < pend; j += CHS_THREADS) {
            const int64_t g = int64_t(perm[j]) - c * total;
            const int64_t q = g - D.point_begin;
            double sx, sy;
            int64_t bx, by;
            if (!chb_point_base(D, c, D.x[2 * q], D.x[2 * q + 1], sx, sy, bx, by)) continue;
            const int64_t lx = chs_key_x(D, bx) - col0, ly = by - row0;
            if (lx < 0 || lx >= CHS_T || ly < 0 || ly >= CHS_T) {
                raise_flag(flags, F_INDEX);
                continue;
            }
            double fx[4], fy[4];
            bool out[4];
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) {
                fx[s4] = peskin_phi(sx - double(bx + s4));
                fy[s4] = peskin_phi(sy - double(by + s4));
                chb_reflect(c, D.d.ny, by + s4, out[s4]);
            }
            double wxr[4], wyr[4];
            bool outr[4];
            int oxr[4], oyr[4];
#pragma unroll
            for (int q4 = 0; q4 < 4; ++q4) {
                const int iq = (q4 + ri) & 3, jq = (q4 + rj) & 3;
                wxr[q4] = iq == 0 ? fx[0] : (iq == 1 ? fx[1] : (iq == 2 ? fx[2] : fx[3]));
                wyr[q4] = jq == 0 ? fy[0] : (jq == 1 ? fy[1] : (jq == 2 ? fy[2] : fy[3]));
                outr[q4] = jq == 0 ? out[0] : (jq == 1 ? out[1] : (jq == 2 ? out[2] : out[3]));
                oxr[q4] = int(lx + iq) * CHS_BOX;
                oyr[q4] = int(ly + jq);
            }
            const double f = force[2 * g + c];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) {
                    const double wt = wxr[i] * wyr[jj];
                    atomicAdd(box + oxr[i] + oyr[jj], f * (outr[jj] ? -wt : wt));
                }
        }
        __syncthreads();
        double* out = c == 0 ? fu + D.u_off : fv + D.v_off;
        const int64_t fny = D.d.ny + c, count = c == 0 ? D.u_count : D.v_count;
        for (int e = threadIdx.x; e < CHS_BOX * CHS_BOX; e += CHS_THREADS) {
            const double v = box[e];
            if (v == 0.0) continue;
            int64_t col = col0 + e / CHS_BOX;
            if (D.periodic) col = col % D.d.nx;
            bool outside;
            const int64_t id = col * fny + chb_reflect(c, D.d.ny, row0 + e % CHS_BOX, outside);
            if (id < 0 || id >= count) {
                raise_flag(flags, F_INDEX);
                continue;
            }
            unsafeAtomicAdd(out + id, v);
        }
        __syncthreads();
    }
}

// rhs[0] of instance i = scale * D u (ch_divergence_kernel's expression).
__global__ void __launch_bounds__(CH_THREADS) chb_divergence_kernel(const ChbDesc* __restrict__ desc,
                                                                    const afsi_mg2d_plan* __restrict__ plans,
                                                                    const int64_t* __restrict__ cell_begin, int n,
                                                                    int64_t total, const double* __restrict__ u,
                                                                    const double* __restrict__ v, double scale) {
    for (int64_t j = int64_t(blockIdx.x) * CH_THREADS + threadIdx.x; j < total;
         j += int64_t(gridDim.x) * CH_THREADS) {
        int lo = 0, hi = n;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (cell_begin[mid] <= j) lo = mid;
            else hi = mid;
        }
        const ChbDesc& D = desc[lo];
        const Dim2 d = D.d;
        const int i = int(j - cell_begin[lo]);
        const int x = i / d.ny, y = i - x * d.ny;
        const double* ui = u + D.u_off;
        const double* vi = v + D.v_off;
        const int xr = D.periodic ? (x + 1 < d.nx ? x + 1 : 0) : x + 1;
        const double du = (ui[xr * d.ny + y] - ui[i]) / D.hx;
        const double dv = (vi[x * (d.ny + 1) + y + 1] - vi[x * (d.ny + 1) + y]) / D.hy;
        static_cast<double*>(plans[lo].rhs[0])[i] = scale * (du + dv);
    }
}

// u -= coef * G p with p = the instance's p[0] (ch_grad_correct_kernel's
// expressions); u faces of all instances first, then v faces.
__global__ void __launch_bounds__(CH_THREADS) chb_correct_kernel(const ChbDesc* __restrict__ desc,
                                                                 const afsi_mg2d_plan* __restrict__ plans, int n,
                                                                 int64_t nu, int64_t nv, double* __restrict__ u,
                                                                 double* __restrict__ v, double coef) {
    for (int64_t t = int64_t(blockIdx.x) * CH_THREADS + threadIdx.x; t < nu + nv;
         t += int64_t(gridDim.x) * CH_THREADS) {
        if (t < nu) {
            const int k = chb_find(desc, n, t, 1);
            const ChbDesc& D = desc[k];
            const Dim2 d = D.d;
            const double* p = static_cast<const double*>(plans[k].p[0]);
            const int i = int(t - D.u_off);
            const int x = i / d.ny, y = i - x * d.ny;
            double g;
            if (D.periodic) g = (p[x * d.ny + y] - p[(x > 0 ? x - 1 : d.nx - 1) * d.ny + y]) / D.hx;
            else if (x == 0) continue;
            else if (x == d.nx) g = -2 * p[(d.nx - 1) * d.ny + y] / D.hx;
            else g = (p[x * d.ny + y] - p[(x - 1) * d.ny + y]) / D.hx;
            u[t] = u[t] - coef * g;
        } else {
            const int64_t tv = t - nu;
            const int k = chb_find(desc, n, tv, 2);
            const ChbDesc& D = desc[k];
            const Dim2 d = D.d;
            const double* p = static_cast<const double*>(plans[k].p[0]);
            const int j = int(tv - D.v_off), x = j / (d.ny + 1), y = j - x * (d.ny + 1);
            if (y == 0 || y == d.ny) continue;
            const double g = (p[x * d.ny + y] - p[x * d.ny + y - 1]) / D.hy;
            v[tv] = v[tv] - coef * g;
        }
    }
}

// One workgroup per active instance: `cycles` V-cycles (ch_block_cycles).
// One instantiation per boundary type (each holds one inlined V-cycle, so
// the register budget of 1024 threads is not shared by two); workgroups of
// the other type return at once, and so do instances whose upper levels run
// as batched launches (split[i] > 0).
template <bool P>
__global__ void __launch_bounds__(CHB_THREADS) chb_cycles_kernel(const afsi_mg2d_plan* __restrict__ plans,
                                                                 const int32_t* __restrict__ split,
                                                                 const int32_t* __restrict__ active, int cycles) {
    __shared__ double red[CHB_THREADS];
    const afsi_mg2d_plan& pl = plans[blockIdx.x];
    if ((active && !active[blockIdx.x]) || (pl.periodic != 0) != P || split[blockIdx.x] > 0) return;
    chb_cycles<double, P>(pl, cycles, red);
}

// The lower part (levels split[i]..last) of one V-cycle of every active
// instance with split[i] > 0.
template <bool P>
__global__ void __launch_bounds__(CHB_THREADS) chb_tail_kernel(const afsi_mg2d_plan* __restrict__ plans,
                                                               const int32_t* __restrict__ split,
                                                               const int32_t* __restrict__ active) {
    __shared__ double red[CHB_THREADS];
    const afsi_mg2d_plan& pl = plans[blockIdx.x];
    const int l0 = split[blockIdx.x];
    if ((active && !active[blockIdx.x]) || (pl.periodic != 0) != P || l0 <= 0) return;
    chb_cycles<double, P>(pl, 1, red, l0);
}

// ---- upper levels of large instances as batched launches ---------------------
// A level with more than MLB_MIN_CELLS cells (RBGS instances) would keep one
// CU busy for hundreds of microseconds per cycle in the one-workgroup kernel:
// its FP64 stencil work is issued by that one CU. Such levels run instead as
// one launch per phase over every instance that has them, MLB_CHUNK cells of
// one instance per workgroup, with the per-cell expressions of the
// one-workgroup kernel (so smoothing, restriction and prolongation give the
// same values). Only the periodic mean removal sums in another fixed order:
// per-workgroup partial sums, then the instance's partials in workgroup
// order. Level l's work list holds the instances whose first split[i]
// levels include l; the levels below run in chb_tail_kernel.
constexpr int MLB_THREADS = 256;
constexpr int MLB_CHUNK = 2048;           // level-l cells per workgroup
constexpr int64_t MLB_MIN_CELLS = 4096;

struct MlbEntry {
    int32_t inst, block_begin;
};

__device__ inline MlbEntry mlb_find(const MlbEntry* __restrict__ list, int n) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (list[mid].block_begin <= int(blockIdx.x)) lo = mid;
        else hi = mid;
    }
    return list[lo];
}

template <bool P, bool M>
__device__ __forceinline__ void mlb_smooth_chunk(const afsi_mg2d_plan& pl, int l, int chunk, int color) {
    const Dim2 d{int(pl.n[l][0]), int(pl.n[l][1])};
    const Coef2<double> k = coef2_of<double>(pl.h2[l][0], pl.h2[l][1], pl.inv_h2[l][0], pl.inv_h2[l][1]);
    double* p = static_cast<double*>(pl.p[l]);
    const double* rhs = static_cast<const double*>(pl.rhs[l]);
    const int hy = d.ny / 2, half = d.nx * hy;
    constexpr int G = MLB_CHUNK / 2 / MLB_THREADS;
    const int e0 = chunk * (MLB_CHUNK / 2) + int(threadIdx.x);
    double val[G];
    int at[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int e = e0 + g * MLB_THREADS;
        at[g] = -1;
        if (e >= half) continue;
        const int x = e / hy, y = 2 * (e - x * hy) + ((x + color) & 1);
        const int i = x * d.ny + y;
        double c;
        const double r = rhs[i] - ch_apply<double, P, M>(p, x, y, d, k, c);
        val[g] = c + r / ch_diag<double, P>(x, y, d, k);
        at[g] = i;
    }
#pragma unroll
    for (int g = 0; g < G; ++g)
        if (at[g] >= 0) p[at[g]] = val[g];
}

// One colour of red-black iteration `it` on level l (instances with
// pl.nu <= it skip it).
__global__ void __launch_bounds__(MLB_THREADS) mlb_smooth_kernel(const afsi_mg2d_plan* __restrict__ plans,
                                                                 const MlbEntry* __restrict__ list, int n, int l,
                                                                 int it, int color,
                                                                 const int32_t* __restrict__ active) {
    const MlbEntry e = mlb_find(list, n);
    const afsi_mg2d_plan& pl = plans[e.inst];
    if ((active && !active[e.inst]) || it >= pl.nu) return;
    const int chunk = int(blockIdx.x) - e.block_begin;
    const bool m = chb_exact(pl, l);
    if (pl.periodic) {
        if (m) mlb_smooth_chunk<true, true>(pl, l, chunk, color);
        else mlb_smooth_chunk<true, false>(pl, l, chunk, color);
    } else {
        if (m) mlb_smooth_chunk<false, true>(pl, l, chunk, color);
        else mlb_smooth_chunk<false, false>(pl, l, chunk, color);
    }
}

// The range of a chunk of level l in the array being averaged: p[l]
// (coarse == 0, MLB_CHUNK cells) or rhs[l + 1] (coarse == 1, a quarter).
__device__ inline void mlb_mean_range(const afsi_mg2d_plan& pl, int l, int coarse, int chunk, double*& x,
                                      int& begin, int& end, int& count) {
    const int cells = int(pl.n[l][0] * pl.n[l][1]);
    const int per = coarse ? MLB_CHUNK / 4 : MLB_CHUNK;
    count = coarse ? cells / 4 : cells;
    x = static_cast<double*>(coarse ? pl.rhs[l + 1] : pl.p[l]);
    begin = chunk * per;
    end = begin + per < count ? begin + per : count;
}

// Partial sum of the chunk (periodic instances), in a fixed order.
__global__ void __launch_bounds__(MLB_THREADS) mlb_sum_kernel(const afsi_mg2d_plan* __restrict__ plans,
                                                              const MlbEntry* __restrict__ list, int n, int l,
                                                              int coarse, const int32_t* __restrict__ active,
                                                              double* __restrict__ partials) {
    __shared__ double red[MLB_THREADS];
    const MlbEntry e = mlb_find(list, n);
    const afsi_mg2d_plan& pl = plans[e.inst];
    if ((active && !active[e.inst]) || !pl.periodic) return;
    double* x;
    int begin, end, count;
    mlb_mean_range(pl, l, coarse, int(blockIdx.x) - e.block_begin, x, begin, end, count);
    double v = 0.0;
    for (int i = begin + int(threadIdx.x); i < end; i += MLB_THREADS) v += x[i];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = MLB_THREADS / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

// x -= mean, the mean from the instance's partials in workgroup order.
__global__ void __launch_bounds__(MLB_THREADS) mlb_sub_kernel(const afsi_mg2d_plan* __restrict__ plans,
                                                              const MlbEntry* __restrict__ list, int n, int l,
                                                              int coarse, const int32_t* __restrict__ active,
                                                              const double* __restrict__ partials) {
    __shared__ double total;
    const MlbEntry e = mlb_find(list, n);
    const afsi_mg2d_plan& pl = plans[e.inst];
    if ((active && !active[e.inst]) || !pl.periodic) return;
    double* x;
    int begin, end, count;
    mlb_mean_range(pl, l, coarse, int(blockIdx.x) - e.block_begin, x, begin, end, count);
    if (threadIdx.x == 0) {
        const int blocks = int((pl.n[l][0] * pl.n[l][1] + MLB_CHUNK - 1) / MLB_CHUNK);
        double t = 0.0;
        for (int b = 0; b < blocks; ++b) t += partials[e.block_begin + b];
        total = t;
    }
    __syncthreads();
    const double mean = total / double(count);
    for (int i = begin + int(threadIdx.x); i < end; i += MLB_THREADS) x[i] = x[i] - mean;
}

template <bool P, bool M>
__device__ __forceinline__ void mlb_restrict_chunk(const afsi_mg2d_plan& pl, int l, int chunk) {
    const Dim2 d{int(pl.n[l][0]), int(pl.n[l][1])};
    const Coef2<double> k = coef2_of<double>(pl.h2[l][0], pl.h2[l][1], 0.0, 0.0);
    const double* p = static_cast<const double*>(pl.p[l]);
    const double* rhs = static_cast<const double*>(pl.rhs[l]);
    double* coarse = static_cast<double*>(pl.rhs[l + 1]);
    double* pc = static_cast<double*>(pl.p[l + 1]);
    const int cy = d.ny / 2, nc = (d.nx / 2) * cy;
    constexpr int G = MLB_CHUNK / 4 / MLB_THREADS;
    const int i0 = chunk * (MLB_CHUNK / 4) + int(threadIdx.x);
    double val[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int i = i0 + g * MLB_THREADS;
        if (i >= nc) continue;
        const int X = i / cy, Y = i - X * cy;
        const int x = 2 * X, y = 2 * Y, j = x * d.ny + y;
        double c;
        const double r00 = rhs[j] - ch_apply<double, P, M>(p, x, y, d, k, c);
        const double r01 = rhs[j + 1] - ch_apply<double, P, M>(p, x, y + 1, d, k, c);
        const double r10 = rhs[j + d.ny] - ch_apply<double, P, M>(p, x + 1, y, d, k, c);
        const double r11 = rhs[j + d.ny + 1] - ch_apply<double, P, M>(p, x + 1, y + 1, d, k, c);
        val[g] = (((r00 + r01) + r10) + r11) * 0.25;
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int i = i0 + g * MLB_THREADS;
        if (i < nc) {
            coarse[i] = val[g];
            pc[i] = 0.0;
        }
    }
}

// rhs[l + 1] = restricted residual of level l; p[l + 1] = 0.
__global__ void __launch_bounds__(MLB_THREADS) mlb_restrict_kernel(const afsi_mg2d_plan* __restrict__ plans,
                                                                   const MlbEntry* __restrict__ list, int n, int l,
                                                                   const int32_t* __restrict__ active) {
    const MlbEntry e = mlb_find(list, n);
    const afsi_mg2d_plan& pl = plans[e.inst];
    if (active && !active[e.inst]) return;
    const int chunk = int(blockIdx.x) - e.block_begin;
    const bool m = chb_exact(pl, l);
    if (pl.periodic) {
        if (m) mlb_restrict_chunk<true, true>(pl, l, chunk);
        else mlb_restrict_chunk<true, false>(pl, l, chunk);
    } else {
        if (m) mlb_restrict_chunk<false, true>(pl, l, chunk);
        else mlb_restrict_chunk<false, false>(pl, l, chunk);
    }
}

template <bool P>
__device__ __forceinline__ void mlb_prolong_chunk(const afsi_mg2d_plan& pl, int l, int chunk) {
    const Dim2 d{int(pl.n[l][0]), int(pl.n[l][1])};
    const double* coarse = static_cast<const double*>(pl.p[l + 1]);
    double* fine = static_cast<double*>(pl.p[l]);
    const int n = d.nx * d.ny, NX = d.nx / 2, NY = d.ny / 2;
    constexpr int G = MLB_CHUNK / MLB_THREADS;
    const int i0 = chunk * MLB_CHUNK + int(threadIdx.x);
    double val[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int i = i0 + g * MLB_THREADS;
        if (i >= n) continue;
        const int x = i / d.ny, y = i - x * d.ny;
        const int ax = x / 2 - 1 + (x & 1), ay = y / 2 - 1 + (y & 1);
        const double wx = (x & 1) ? 0.25 : 0.75, wy = (y & 1) ? 0.25 : 0.75;
        const double c00 = ch_coarse_value<double, P>(coarse, ax, ay, NX, NY);
        const double c01 = ch_coarse_value<double, P>(coarse, ax, ay + 1, NX, NY);
        const double c10 = ch_coarse_value<double, P>(coarse, ax + 1, ay, NX, NY);
        const double c11 = ch_coarse_value<double, P>(coarse, ax + 1, ay + 1, NX, NY);
        const double v = (c00 * (1.0 - wy) + c01 * wy) * (1.0 - wx) + (c10 * (1.0 - wy) + c11 * wy) * wx;
        val[g] = fine[i] + v;
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int i = i0 + g * MLB_THREADS;
        if (i < n) fine[i] = val[g];
    }
}

// p[l] += prolongation of p[l + 1].
__global__ void __launch_bounds__(MLB_THREADS) mlb_prolong_kernel(const afsi_mg2d_plan* __restrict__ plans,
                                                                  const MlbEntry* __restrict__ list, int n, int l,
                                                                  const int32_t* __restrict__ active) {
    const MlbEntry e = mlb_find(list, n);
    const afsi_mg2d_plan& pl = plans[e.inst];
    if (active && !active[e.inst]) return;
    const int chunk = int(blockIdx.x) - e.block_begin;
    if (pl.periodic) mlb_prolong_chunk<true>(pl, l, chunk);
    else mlb_prolong_chunk<false>(pl, l, chunk);
}

// One workgroup per instance: [sum, sum of squares, sum |.|, non-finite
// count] of rhs[0] (which 0), p[0] (1) or rhs[0] - A p[0] (2), summed in a
// fixed order; which 3/4: subtract the mean out[4i]/n from rhs[0]/p[0].
// Value of cell i of level 0 for the statistics: rhs (which 0), p (1) or
// the residual rhs - A p (2).
__device__ inline double chb_stat_value(const afsi_mg2d_plan& pl, const Coef2<double>& k, bool m, int which,
                                        int i) {
    const Dim2 d{int(pl.n[0][0]), int(pl.n[0][1])};
    const double* rhs = static_cast<const double*>(pl.rhs[0]);
    const double* p = static_cast<const double*>(pl.p[0]);
    if (which == 0) return rhs[i];
    if (which == 1) return p[i];
    const int x = i / d.ny, y = i - x * d.ny;
    double c;
    if (pl.periodic)
        return rhs[i] - (m ? ch_apply<double, true, true>(p, x, y, d, k, c) : ch_apply<double, true, false>(p, x, y, d, k, c));
    return rhs[i] - (m ? ch_apply<double, false, true>(p, x, y, d, k, c) : ch_apply<double, false, false>(p, x, y, d, k, c));
}

__global__ void __launch_bounds__(CHB_THREADS) chb_stats_kernel(const afsi_mg2d_plan* __restrict__ plans,
                                                                const int32_t* __restrict__ split, int which,
                                                                double* __restrict__ out) {
    __shared__ double red[4][CHB_THREADS];
    if (split[blockIdx.x] > 0) return;  // level 0 is covered by mlb_stats_kernel
    const afsi_mg2d_plan& pl = plans[blockIdx.x];
    const Dim2 d{int(pl.n[0][0]), int(pl.n[0][1])};
    const int n = d.nx * d.ny;
    double* rhs = static_cast<double*>(pl.rhs[0]);
    double* p = static_cast<double*>(pl.p[0]);
    if (which >= 3) {  // mean removal: periodic channels only (the reference channel is non-singular)
        if (!pl.periodic) return;
        double* x = which == 3 ? rhs : p;
        const double mean = out[4 * blockIdx.x] / double(n);
        for (int i = threadIdx.x; i < n; i += CHB_THREADS) x[i] = x[i] - mean;
        return;
    }
    const Coef2<double> k = coef2_of<double>(pl.h2[0][0], pl.h2[0][1], 0.0, 0.0);
    const bool m = k.exact();
    double s = 0.0, q = 0.0, a = 0.0, bad = 0.0;
    for (int i = threadIdx.x; i < n; i += CHB_THREADS) {
        const double v = chb_stat_value(pl, k, m, which, i);
        if (isfinite(v)) {
            s += v;
            q += v * v;
            a += fabs(v);
        } else {
            bad += 1.0;
        }
    }
    red[0][threadIdx.x] = s;
    red[1][threadIdx.x] = q;
    red[2][threadIdx.x] = a;
    red[3][threadIdx.x] = bad;
    __syncthreads();
    for (int w = CHB_THREADS / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w)
            for (int m = 0; m < 4; ++m) red[m][threadIdx.x] += red[m][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 4) out[4 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// The same statistics for instances whose level 0 runs as batched launches
// (level-0 work list): partial [sum, squares, |.|, non-finite] per chunk in
// a fixed order, then mlb_stats_final_kernel adds an instance's partials in
// chunk order. which 3 / 4: subtract out[4i] / cells from the chunk.
__global__ void __launch_bounds__(MLB_THREADS) mlb_stats_kernel(const afsi_mg2d_plan* __restrict__ plans,
                                                                const MlbEntry* __restrict__ list, int n, int which,
                                                                double* __restrict__ partials,
                                                                double* __restrict__ out) {
    __shared__ double red[4][MLB_THREADS];
    const MlbEntry e = mlb_find(list, n);
    const afsi_mg2d_plan& pl = plans[e.inst];
    const int cells = int(pl.n[0][0] * pl.n[0][1]);
    const int begin = (int(blockIdx.x) - e.block_begin) * MLB_CHUNK;
    const int end = begin + MLB_CHUNK < cells ? begin + MLB_CHUNK : cells;
    if (which >= 3) {
        if (!pl.periodic) return;
        double* x = static_cast<double*>(which == 3 ? pl.rhs[0] : pl.p[0]);
        const double mean = out[4 * e.inst] / double(cells);
        for (int i = begin + int(threadIdx.x); i < end; i += MLB_THREADS) x[i] = x[i] - mean;
        return;
    }
    const Coef2<double> k = coef2_of<double>(pl.h2[0][0], pl.h2[0][1], 0.0, 0.0);
    const bool m = k.exact();
    double s = 0.0, q = 0.0, a = 0.0, bad = 0.0;
    for (int i = begin + int(threadIdx.x); i < end; i += MLB_THREADS) {
        const double v = chb_stat_value(pl, k, m, which, i);
        if (isfinite(v)) {
            s += v;
            q += v * v;
            a += fabs(v);
        } else {
            bad += 1.0;
        }
    }
    red[0][threadIdx.x] = s;
    red[1][threadIdx.x] = q;
    red[2][threadIdx.x] = a;
    red[3][threadIdx.x] = bad;
    __syncthreads();
    for (int w = MLB_THREADS / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w)
            for (int c = 0; c < 4; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 4) partials[4 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void mlb_stats_final_kernel(const afsi_mg2d_plan* __restrict__ plans, const MlbEntry* __restrict__ list,
                                       int n, const double* __restrict__ partials, double* __restrict__ out) {
    const int t = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= 4 * n) return;
    const MlbEntry e = list[t / 4];
    const int c = t % 4;
    const afsi_mg2d_plan& pl = plans[e.inst];
    const int blocks = int((pl.n[0][0] * pl.n[0][1] + MLB_CHUNK - 1) / MLB_CHUNK);
    double v = 0.0;
    for (int b = 0; b < blocks; ++b) v += partials[4 * (e.block_begin + b) + c];
    out[4 * e.inst + c] = v;
}

struct ChbLayout {
    int n = 0;
    int64_t points = 0, nu = 0, nv = 0, cells = 0;
    size_t desc = 0, plans = 0, cell_begin = 0, split = 0, lists = 0, partials = 0, tiles = 0, total = 0;
    int64_t ntiles = 0;  // spread tiles over all instances and both components
    // batched upper levels: levels 0..maxsplit-1 have work lists
    int maxsplit = 0;
    int count[AFSI_MG_MAX_LEVELS] = {}, blocks[AFSI_MG_MAX_LEVELS] = {}, max_nu[AFSI_MG_MAX_LEVELS] = {};
    bool mean[AFSI_MG_MAX_LEVELS] = {};
    bool whole_periodic = false, whole_closed = false, tail_periodic = false, tail_closed = false;
};

// Number of upper levels of an instance that run as batched launches.
int mlb_split(const afsi_mg2d_plan& pl) {
    if (pl.smoother != 1) return 0;
    int s = 0;
    while (s < pl.levels - 1 && pl.n[s][0] * pl.n[s][1] > MLB_MIN_CELLS && pl.n[s][1] % 2 == 0) ++s;
    return s;
}

size_t chb_align(size_t b) { return (b + 255) & ~size_t(255); }

int chb_layout(const afsi_chb_batch* b, ChbLayout& L, std::vector<ChbDesc>* descs,
               std::vector<int64_t>* cell_begin, std::vector<int32_t>* splits = nullptr,
               std::vector<MlbEntry>* lists = nullptr, std::vector<ChsTile>* tiles = nullptr) {
    AFSI_REQUIRE(b && b->items && b->count >= 1, "invalid channel batch");
    L.n = b->count;
    for (int i = 0; i < L.n; ++i) {
        const afsi_chb_item& it = b->items[i];
        const afsi_grid2& g = it.grid;
        AFSI_REQUIRE(dims2_ok(g.n) && g.n[0] >= 4 && g.n[1] >= 4, "channel batch item %d: invalid grid", i);
        AFSI_REQUIRE(it.points >= 0 && (it.points == 0 || it.x), "channel batch item %d: invalid points", i);
        const afsi_mg2d_plan& pl = it.plan;
        if (int r = channel_check<double>(&pl)) return r;
        AFSI_REQUIRE(pl.n[0][0] == g.n[0] && pl.n[0][1] == g.n[1] && pl.periodic == g.periodic,
                     "channel batch item %d: plan does not match its grid", i);
        ChbDesc d{};
        d.x = it.x;
        d.points = it.points;
        d.point_begin = L.points;
        d.d = Dim2{int(g.n[0]), int(g.n[1])};
        d.hx = g.h[0];
        d.hy = g.h[1];
        d.Lx = g.lengths[0];
        d.Ly = g.lengths[1];
        d.periodic = g.periodic;
        d.u_off = L.nu;
        d.v_off = L.nv;
        d.u_count = (g.periodic ? g.n[0] : g.n[0] + 1) * g.n[1];
        d.v_count = g.n[0] * (g.n[1] + 1);
        d.cell_count = g.n[0] * g.n[1];
        if (descs) descs->push_back(d);
        if (cell_begin) cell_begin->push_back(L.cells);
        for (int c = 0; c < 2; ++c) {
            // key x: wrapped base (periodic) or base <= n (closed); key y = base + 3 in [0, ny + 4)
            const int64_t kx = g.periodic ? g.n[0] : g.n[0] + 1, ky = g.n[1] + 4;
            const ChsTile t{int32_t(L.ntiles), int32_t(cdiv(kx, CHS_T)), int32_t(cdiv(ky, CHS_T)), 0};
            if (tiles) tiles->push_back(t);
            L.ntiles += int64_t(t.ntx) * t.nty;
        }
        const int sp = mlb_split(pl);
        if (splits) splits->push_back(sp);
        if (sp > L.maxsplit) L.maxsplit = sp;
        if (sp == 0) (pl.periodic ? L.whole_periodic : L.whole_closed) = true;
        else (pl.periodic ? L.tail_periodic : L.tail_closed) = true;
        L.points += it.points;
        L.nu += d.u_count;
        L.nv += d.v_count;
        L.cells += d.cell_count;
    }
    AFSI_REQUIRE(L.points < (int64_t(1) << 29) && L.nu < (int64_t(1) << 31) && L.nv < (int64_t(1) << 31) &&
                     L.ntiles < (int64_t(1) << 30),
                 "channel batch too large");
    if (lists) lists->assign(size_t(AFSI_MG_MAX_LEVELS) * L.n, MlbEntry{0, 0});
    int max_blocks = 1;
    for (int l = 0; l < L.maxsplit; ++l) {
        int64_t blocks = 0;
        for (int i = 0; i < L.n; ++i) {
            const afsi_mg2d_plan& pl = b->items[i].plan;
            if (mlb_split(pl) <= l) continue;
            if (lists) (*lists)[size_t(l) * L.n + L.count[l]] = MlbEntry{i, int32_t(blocks)};
            ++L.count[l];
            blocks += cdiv(pl.n[l][0] * pl.n[l][1], MLB_CHUNK);
            if (pl.nu > L.max_nu[l]) L.max_nu[l] = pl.nu;
            if (pl.periodic) L.mean[l] = true;
        }
        AFSI_REQUIRE(blocks < (int64_t(1) << 30), "channel batch level too large");
        L.blocks[l] = int(blocks);
        if (L.blocks[l] > max_blocks) max_blocks = L.blocks[l];
    }
    size_t off = 0;
    L.desc = off; off = chb_align(off + size_t(L.n) * sizeof(ChbDesc));
    L.plans = off; off = chb_align(off + size_t(L.n) * sizeof(afsi_mg2d_plan));
    L.cell_begin = off; off = chb_align(off + size_t(L.n) * sizeof(int64_t));
    L.split = off; off = chb_align(off + size_t(L.n) * sizeof(int32_t));
    L.lists = off; off = chb_align(off + size_t(AFSI_MG_MAX_LEVELS) * L.n * sizeof(MlbEntry));
    L.partials = off; off = chb_align(off + 4 * size_t(max_blocks) * sizeof(double));
    L.tiles = off; off = chb_align(off + 2 * size_t(L.n) * sizeof(ChsTile));
    L.total = off;
    return OK;
}

int chb_ready(const afsi_chb_batch* b, ChbLayout& L) {
    if (int r = chb_layout(b, L, nullptr, nullptr)) return r;
    AFSI_REQUIRE(b->workspace && b->workspace_bytes >= L.total, "channel batch workspace too small");
    return OK;
}

template <class D>
D* chb_at(const afsi_chb_batch* b, size_t off) {
    return reinterpret_cast<D*>(static_cast<char*>(b->workspace) + off);
}

}  // namespace
}  // namespace afsi

extern "C" {

int afsi_chb_sizes(const afsi_chb_batch* batch, int64_t* points, int64_t* u_faces, int64_t* v_faces,
                   int64_t* cells, size_t* workspace_bytes) {
    ChbLayout L;
    if (int r = chb_layout(batch, L, nullptr, nullptr)) return r;
    if (points) *points = L.points;
    if (u_faces) *u_faces = L.nu;
    if (v_faces) *v_faces = L.nv;
    if (cells) *cells = L.cells;
    if (workspace_bytes) *workspace_bytes = L.total;
    return OK;
}

int afsi_chb_prepare(const afsi_chb_batch* batch, void* stream) {
    ChbLayout L;
    std::vector<ChbDesc> descs;
    std::vector<int64_t> begins;
    std::vector<int32_t> splits;
    std::vector<MlbEntry> lists;
    std::vector<ChsTile> tiles;
    if (int r = chb_layout(batch, L, &descs, &begins, &splits, &lists, &tiles)) return r;
    AFSI_REQUIRE(batch->workspace && batch->workspace_bytes >= L.total, "channel batch workspace too small");
    std::vector<char> host(L.total, 0);
    std::memcpy(host.data() + L.desc, descs.data(), descs.size() * sizeof(ChbDesc));
    for (int i = 0; i < L.n; ++i)
        std::memcpy(host.data() + L.plans + size_t(i) * sizeof(afsi_mg2d_plan), &batch->items[i].plan,
                    sizeof(afsi_mg2d_plan));
    std::memcpy(host.data() + L.cell_begin, begins.data(), begins.size() * sizeof(int64_t));
    std::memcpy(host.data() + L.split, splits.data(), splits.size() * sizeof(int32_t));
    std::memcpy(host.data() + L.lists, lists.data(), lists.size() * sizeof(MlbEntry));
    std::memcpy(host.data() + L.tiles, tiles.data(), tiles.size() * sizeof(ChsTile));
    const hipStream_t s = hipStream_t(stream);
    AFSI_HIP(hipMemcpyAsync(batch->workspace, host.data(), host.size(), hipMemcpyHostToDevice, s));
    AFSI_HIP(hipStreamSynchronize(s));
    return OK;
}

int afsi_chb_stencil_f64(const afsi_chb_batch* batch, int64_t* ids_u, double* w_u, int64_t* ids_v, double* w_v,
                         int* flags, void* stream) {
    ChbLayout L;
    if (int r = chb_ready(batch, L)) return r;
    if (L.points == 0) return OK;
    AFSI_REQUIRE(ids_u && w_u && ids_v && w_v && flags, "null channel stencil buffer");
    const int64_t stencil_blocks = 2 * cdiv(L.points, CH_THREADS);
    chb_stencil_kernel<<<unsigned(stencil_blocks < 8192 ? stencil_blocks : 8192), CH_THREADS, 0,
                         hipStream_t(stream)>>>(
        chb_at<ChbDesc>(batch, L.desc), L.n, L.points, ids_u, w_u, ids_v, w_v, flags);
    AFSI_LAUNCH_CHECK();
    return OK;
}

// Spread workspace: the sort of 2 * points (key, value) pairs, the sorted
// keys and values, then tile_ptr and chunk_start over ntiles + 1 entries.
static size_t chs_align(size_t b) { return (b + 255) & ~size_t(255); }

size_t afsi_chb_spread_workspace_bytes(const afsi_chb_batch* batch) {
    ChbLayout L;
    if (chb_layout(batch, L, nullptr, nullptr)) return 0;
    const int64_t m = 2 * L.points > 0 ? 2 * L.points : 1;
    return chs_align(bin_workspace_bytes(m)) + 2 * chs_align(size_t(m) * 4) + 2 * chs_align(size_t(L.ntiles + 1) * 4);
}

int afsi_chb_spread_f64(const afsi_chb_batch* batch, const double* force, double* fu, double* fv, void* workspace,
                        size_t workspace_bytes, int* flags, void* stream) {
    ChbLayout L;
    if (int r = chb_ready(batch, L)) return r;
    AFSI_REQUIRE(fu && fv && flags, "null channel spread buffer");
    const hipStream_t s = hipStream_t(stream);
    if (int r = afsi_fill_f64(fu, L.nu, 0.0, stream)) return r;
    if (int r = afsi_fill_f64(fv, L.nv, 0.0, stream)) return r;
    if (L.points == 0) return OK;
    AFSI_REQUIRE(force && workspace, "null channel spread buffer");
    const size_t need = afsi_chb_spread_workspace_bytes(batch);
    AFSI_REQUIRE(need > 0 && workspace_bytes >= need, "channel spread workspace too small");
    const int64_t m = 2 * L.points;
    char* ws = static_cast<char*>(workspace);
    size_t off = chs_align(bin_workspace_bytes(m));
    uint32_t* keys = reinterpret_cast<uint32_t*>(ws + off);
    off += chs_align(size_t(m) * 4);
    int32_t* perm = reinterpret_cast<int32_t*>(ws + off);
    off += chs_align(size_t(m) * 4);
    int32_t* tile_ptr = reinterpret_cast<int32_t*>(ws + off);
    off += chs_align(size_t(L.ntiles + 1) * 4);
    int32_t* chunk_start = reinterpret_cast<int32_t*>(ws + off);
    uint32_t *k0, *v0;
    sort_key_slots(ws, m, &k0, &v0);
    const ChbDesc* desc = chb_at<ChbDesc>(batch, L.desc);
    const ChsTile* tiles = chb_at<ChsTile>(batch, L.tiles);
    chs_keys_kernel<<<grid_for(m, 256, 8192), 256, 0, s>>>(desc, L.n, L.points, tiles, L.ntiles, k0, v0);
    AFSI_LAUNCH_CHECK();
    if (int r = sort_pairs(m, L.ntiles, ws, keys, perm, tile_ptr, s)) return r;
    if (int r = launch_chunk_scan(tile_ptr, L.ntiles, 1, CHS_CHUNK, chunk_start, s)) return r;
    const int64_t upper = L.ntiles + m / CHS_CHUNK + 1;
    chs_spread_kernel<<<unsigned(upper < 8192 ? upper : 8192), CHS_THREADS, 0, s>>>(
        desc, L.n, tiles, perm, tile_ptr, chunk_start, L.ntiles, L.points, force, fu, fv, flags);
    AFSI_LAUNCH_CHECK();
    return OK;
}

int afsi_chb_divergence_f64(const afsi_chb_batch* batch, const double* u, const double* v, double scale,
                            void* stream) {
    ChbLayout L;
    if (int r = chb_ready(batch, L)) return r;
    AFSI_REQUIRE(u && v, "null channel velocity");
    chb_divergence_kernel<<<blocks_for(L.cells), CH_THREADS, 0, hipStream_t(stream)>>>(
        chb_at<ChbDesc>(batch, L.desc), chb_at<afsi_mg2d_plan>(batch, L.plans), chb_at<int64_t>(batch, L.cell_begin),
        L.n, L.cells, u, v, scale);
    AFSI_LAUNCH_CHECK();
    return OK;
}

int afsi_chb_correct_f64(const afsi_chb_batch* batch, double* u, double* v, double coef, void* stream) {
    ChbLayout L;
    if (int r = chb_ready(batch, L)) return r;
    AFSI_REQUIRE(u && v, "null channel velocity");
    chb_correct_kernel<<<blocks_for(L.nu + L.nv), CH_THREADS, 0, hipStream_t(stream)>>>(
        chb_at<ChbDesc>(batch, L.desc), chb_at<afsi_mg2d_plan>(batch, L.plans), L.n, L.nu, L.nv, u, v, coef);
    AFSI_LAUNCH_CHECK();
    return OK;
}

int afsi_chb_cycles_f64(const afsi_chb_batch* batch, const int32_t* active, int32_t cycles, void* stream) {
    ChbLayout L;
    if (int r = chb_ready(batch, L)) return r;
    AFSI_REQUIRE(cycles >= 0, "negative cycle count");
    if (!cycles) return OK;
    const hipStream_t s = hipStream_t(stream);
    const afsi_mg2d_plan* plans = chb_at<afsi_mg2d_plan>(batch, L.plans);
    const int32_t* split = chb_at<int32_t>(batch, L.split);
    const unsigned n = unsigned(L.n);
    // instances without batched levels: whole cycles, one workgroup each
    if (L.whole_periodic) chb_cycles_kernel<true><<<n, CHB_THREADS, 0, s>>>(plans, split, active, cycles);
    if (L.whole_closed) chb_cycles_kernel<false><<<n, CHB_THREADS, 0, s>>>(plans, split, active, cycles);
    AFSI_LAUNCH_CHECK();
    if (L.maxsplit == 0) return OK;
    // the others: upper levels as batched launches around a one-workgroup tail
    double* partials = chb_at<double>(batch, L.partials);
    const auto list = [&](int l) { return chb_at<MlbEntry>(batch, L.lists) + size_t(l) * L.n; };
    const auto smooth = [&](int l) {
        for (int it = 0; it < L.max_nu[l]; ++it)
            for (int color = 0; color < 2; ++color)
                mlb_smooth_kernel<<<unsigned(L.blocks[l]), MLB_THREADS, 0, s>>>(plans, list(l), L.count[l], l, it,
                                                                              color, active);
    };
    const auto remove_mean = [&](int l, int coarse) {
        if (!L.mean[l]) return;
        mlb_sum_kernel<<<unsigned(L.blocks[l]), MLB_THREADS, 0, s>>>(plans, list(l), L.count[l], l, coarse, active,
                                                                      partials);
        mlb_sub_kernel<<<unsigned(L.blocks[l]), MLB_THREADS, 0, s>>>(plans, list(l), L.count[l], l, coarse, active,
                                                                      partials);
    };
    for (int c = 0; c < cycles; ++c) {
        for (int l = 0; l < L.maxsplit; ++l) {
            smooth(l);
            remove_mean(l, 0);
            mlb_restrict_kernel<<<unsigned(L.blocks[l]), MLB_THREADS, 0, s>>>(plans, list(l), L.count[l], l, active);
            remove_mean(l, 1);
        }
        if (L.tail_periodic) chb_tail_kernel<true><<<n, CHB_THREADS, 0, s>>>(plans, split, active);
        if (L.tail_closed) chb_tail_kernel<false><<<n, CHB_THREADS, 0, s>>>(plans, split, active);
        for (int l = L.maxsplit - 1; l >= 0; --l) {
            mlb_prolong_kernel<<<unsigned(L.blocks[l]), MLB_THREADS, 0, s>>>(plans, list(l), L.count[l], l, active);
            smooth(l);
            remove_mean(l, 0);
        }
        AFSI_LAUNCH_CHECK();
    }
    return OK;
}

int afsi_chb_stats_f64(const afsi_chb_batch* batch, int which, double* out, void* stream) {
    ChbLayout L;
    if (int r = chb_ready(batch, L)) return r;
    AFSI_REQUIRE(which >= 0 && which <= 4 && out, "invalid channel batch statistics request");
    const hipStream_t s = hipStream_t(stream);
    const afsi_mg2d_plan* plans = chb_at<afsi_mg2d_plan>(batch, L.plans);
    chb_stats_kernel<<<unsigned(L.n), CHB_THREADS, 0, s>>>(plans, chb_at<int32_t>(batch, L.split), which, out);
    AFSI_LAUNCH_CHECK();
    if (L.maxsplit > 0) {
        const MlbEntry* list = chb_at<MlbEntry>(batch, L.lists);
        double* partials = chb_at<double>(batch, L.partials);
        mlb_stats_kernel<<<unsigned(L.blocks[0]), MLB_THREADS, 0, s>>>(plans, list, L.count[0], which, partials, out);
        if (which < 3)
            mlb_stats_final_kernel<<<unsigned(cdiv(4 * L.count[0], 256)), 256, 0, s>>>(plans, list, L.count[0],
                                                                                       partials, out);
        AFSI_LAUNCH_CHECK();
    }
    return OK;
}

}  // extern "C"
