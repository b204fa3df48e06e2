// 3D pressure multigrid kernels, reductions and MAC projection pieces.
//
// Cell arrays are C-order (x,y,z), z fastest, one thread per output cell so
// a wave reads 64 consecutive z-cells (coalesced); the six neighbours come
// through L1/L2. The operator A=-D G is evaluated in the reference's order
// (mac/_triton_mg.py:26-32): a = -(p_x+ - p)/h2x + (p - p_x-)/h2x - ... with
// absent Neumann neighbours skipped, so with -ffp-contract=off every
// elementwise result is bit-identical to the NumPy oracle. The diagonal is
// rebuilt on the fly in the order of mac/multigrid.py:38-41 instead of being
// read from HBM (one stream fewer per sweep).
#include "mgcommon.hpp"
#include <vector>

namespace afsi {

template <class T>
int launch_fill(T* x, int64_t count, T value, hipStream_t s);

// mode 0: out = rhs - A p ; mode 1: out = p + omega*(rhs - A p)/diag
template <class T, bool P, int Mode>
__global__ void __launch_bounds__(256) stencil_kernel(const T* __restrict__ p,
                                                      const T* __restrict__ rhs,
                                                      T* __restrict__ out, Dims d, Coef<T> k,
                                                      T omega) {
    const int64_t total = d.n0 * d.n1 * d.n2;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += int64_t(gridDim.x) * blockDim.x) {
        const int64_t z = i % d.n2, y = (i / d.n2) % d.n1, x = i / (d.n1 * d.n2);
        T c;
        const T a = apply_A<T, P>(p, i, x, y, z, d, k, c);
        const T r = rhs[i] - a;
        if (Mode == 0) out[i] = r;
        else out[i] = c + (omega * r) / diag_of<T, P>(x, y, z, d, k);
    }
}

// In-place Gauss-Seidel update of cells with (x+y+z)%2 == color. Requires
// an even n2 (true on every smoothed level of the hierarchy).
template <class T, bool P>
__global__ void __launch_bounds__(256) rbgs_kernel(T* __restrict__ p, const T* __restrict__ rhs,
                                                   Dims d, Coef<T> k, int color, T relax) {
    const int64_t half = d.n2 / 2;
    const int64_t total = d.n0 * d.n1 * half;
    for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total;
         t += int64_t(gridDim.x) * blockDim.x) {
        const int64_t row = t / half, kk = t % half;
        const int64_t x = row / d.n1, y = row % d.n1;
        const int64_t z = 2 * kk + ((x + y + color) & 1);
        const int64_t i = (x * d.n1 + y) * d.n2 + z;
        T c;
        const T a = apply_A<T, P>(p, i, x, y, z, d, k, c);
        const T dp = (rhs[i] - a) / diag_of<T, P>(x, y, z, d, k);
        p[i] = relax == T(1) ? c + dp : c + relax * dp;  // over-relaxed when relax != 1
    }
}

template <class T>
__global__ void __launch_bounds__(256) restrict_kernel(const T* __restrict__ fine,
                                                       T* __restrict__ coarse, Dims c) {
    const int64_t total = c.n0 * c.n1 * c.n2;
    const int64_t fy = 2 * c.n2, fx = 4 * c.n1 * c.n2;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += int64_t(gridDim.x) * blockDim.x) {
        const int64_t z = i % c.n2, y = (i / c.n2) % c.n1, x = i / (c.n1 * c.n2);
        const int64_t j = (2 * x * 2 * c.n1 + 2 * y) * (2 * c.n2) + 2 * z;
        T v = fine[j];
        v += fine[j + 1];
        v += fine[j + fy];
        v += fine[j + fy + 1];
        v += fine[j + fx];
        v += fine[j + fx + 1];
        v += fine[j + fx + fy];
        v += fine[j + fx + fy + 1];
        coarse[i] = v * T(0.125);
    }
}

template <class T, bool P>
__global__ void __launch_bounds__(256) residual_restrict_kernel(const T* __restrict__ p,
                                                                const T* __restrict__ rhs,
                                                                T* __restrict__ coarse, Dims d,
                                                                Coef<T> k) {
    const int64_t c1 = d.n1 / 2, c2 = d.n2 / 2;
    const int64_t total = (d.n0 / 2) * c1 * c2;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += int64_t(gridDim.x) * blockDim.x) {
        const int64_t z = i % c2, y = (i / c2) % c1, x = i / (c1 * c2);
        T v = T(0);
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int cc = 0; cc < 2; ++cc) {
                    const int64_t fx = 2 * x + a, fy = 2 * y + b, fz = 2 * z + cc;
                    const int64_t j = (fx * d.n1 + fy) * d.n2 + fz;
                    T pc;
                    const T ap = apply_A<T, P>(p, j, fx, fy, fz, d, k, pc);
                    v += rhs[j] - ap;
                }
        coarse[i] = v * T(0.125);
    }
}

template <class T, bool P>
__global__ void __launch_bounds__(256) prolong_add_kernel(const T* __restrict__ coarse,
                                                          T* __restrict__ fine, Dims d) {
    const int64_t C0 = d.n0 / 2, C1 = d.n1 / 2, C2 = d.n2 / 2;
    const int64_t total = d.n0 * d.n1 * d.n2;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += int64_t(gridDim.x) * blockDim.x) {
        const int64_t z = i % d.n2, y = (i / d.n2) % d.n1, x = i / (d.n1 * d.n2);
        const int64_t ax = x / 2 - 1 + x % 2, ay = y / 2 - 1 + y % 2, az = z / 2 - 1 + z % 2;
        int64_t x0, x1, y0, y1, z0, z1;
        if (P) {
            x0 = wrap(ax, C0); x1 = wrap(ax + 1, C0);
            y0 = wrap(ay, C1); y1 = wrap(ay + 1, C1);
            z0 = wrap(az, C2); z1 = wrap(az + 1, C2);
        } else {
            x0 = ax > 0 ? ax : 0; x1 = ax + 1 < C0 - 1 ? ax + 1 : C0 - 1;
            y0 = ay > 0 ? ay : 0; y1 = ay + 1 < C1 - 1 ? ay + 1 : C1 - 1;
            z0 = az > 0 ? az : 0; z1 = az + 1 < C2 - 1 ? az + 1 : C2 - 1;
        }
        const T wx = (x % 2 == 0) ? T(0.75) : T(0.25);
        const T wy = (y % 2 == 0) ? T(0.75) : T(0.25);
        const T wz = (z % 2 == 0) ? T(0.75) : T(0.25);
        auto at = [&](int64_t a, int64_t b, int64_t c) { return coarse[(a * C1 + b) * C2 + c]; };
        const T c00 = at(x0, y0, z0) * (T(1) - wz) + at(x0, y0, z1) * wz;
        const T c01 = at(x0, y1, z0) * (T(1) - wz) + at(x0, y1, z1) * wz;
        const T c10 = at(x1, y0, z0) * (T(1) - wz) + at(x1, y0, z1) * wz;
        const T c11 = at(x1, y1, z0) * (T(1) - wz) + at(x1, y1, z1) * wz;
        const T corr = (c00 * (T(1) - wy) + c01 * wy) * (T(1) - wx) + (c10 * (T(1) - wy) + c11 * wy) * wx;
        fine[i] = fine[i] + corr;
    }
}

template <class T>
__global__ void __launch_bounds__(256) matvec_kernel(const T* __restrict__ A, const T* __restrict__ xv,
                                                     T* __restrict__ y, int64_t n) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int64_t r = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / WAVE;
    if (r >= n) return;
    T s = T(0);
    for (int64_t c = lane; c < n; c += WAVE) s += A[r * n + c] * xv[c];
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d, WAVE);
    if (lane == 0) y[r] = s;
}

// ---- deterministic reductions ----------------------------------------------

constexpr int RED_THREADS = 256;

static int64_t reduce_blocks(int64_t count) {
    int64_t b = cdiv(count, 8 * RED_THREADS);
    if (b < 1) b = 1;
    if (b > 1024) b = 1024;
    return b;
}

__device__ inline void block_reduce4(double v[4], double (*sh)[RED_THREADS]) {
    for (int q = 0; q < 4; ++q) sh[q][threadIdx.x] = v[q];
    __syncthreads();
    for (int s = RED_THREADS / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int q = 0; q < 4; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + s];
        __syncthreads();
    }
    for (int q = 0; q < 4; ++q) v[q] = sh[q][0];
}

template <class T>
__global__ void __launch_bounds__(RED_THREADS) reduce_partial_kernel(const T* __restrict__ x,
                                                                     int64_t count,
                                                                     double* __restrict__ partials) {
    __shared__ double sh[4][RED_THREADS];
    const int64_t per = cdiv(count, gridDim.x);
    const int64_t beg = int64_t(blockIdx.x) * per;
    const int64_t end = beg + per < count ? beg + per : count;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t i = beg + threadIdx.x; i < end; i += RED_THREADS) {
        const double a = double(x[i]);
        if (isfinite(a)) {
            v[0] += a;
            v[1] += a * a;
            v[2] += fabs(a);
        } else {
            v[3] += 1.0;
        }
    }
    block_reduce4(v, sh);
    if (threadIdx.x == 0)
        for (int q = 0; q < 4; ++q) partials[4 * blockIdx.x + q] = v[q];
}

__global__ void __launch_bounds__(RED_THREADS) reduce_final_kernel(const double* __restrict__ partials,
                                                                   int64_t nparts,
                                                                   double* __restrict__ result) {
    __shared__ double sh[4][RED_THREADS];
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t b = threadIdx.x; b < nparts; b += RED_THREADS)
        for (int q = 0; q < 4; ++q) v[q] += partials[4 * b + q];
    block_reduce4(v, sh);
    if (threadIdx.x == 0)
        for (int q = 0; q < 4; ++q) result[q] = v[q];
}

template <class T>
__global__ void sub_mean_kernel(T* __restrict__ x, int64_t count, const double* __restrict__ sum) {
    const T mean = T(sum[0] / double(count));
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
         i += int64_t(gridDim.x) * blockDim.x)
        x[i] = x[i] - mean;
}

// ---- projection pieces --------------------------------------------------------

// Face value of the divergence's input: u, or u + a f (the IB predictor
// folded into the projection, formed as the axpy kernel forms it).
template <bool F>
__device__ __forceinline__ double face_value(const double* __restrict__ u, const double* __restrict__ f,
                                             double a, int64_t i) {
    return F ? u[i] + a * f[i] : u[i];
}

// scale * D(u + a f) at cell i (F) or scale * D u (!F).
template <bool P, bool F>
__device__ __forceinline__ double divergence_at(const double* __restrict__ u, const double* __restrict__ v,
                                                const double* __restrict__ w, const Dims& d, double hx,
                                                double hy, double hz, double scale, const double* __restrict__ fu,
                                                const double* __restrict__ fv, const double* __restrict__ fw,
                                                double a, int64_t i) {
    const int64_t e = P ? 0 : 1;
    // 32-bit index arithmetic (the entry points keep every array below 2^31)
    const uint32_t i32 = uint32_t(i), m1 = uint32_t(d.n1), m2 = uint32_t(d.n2);
    const int64_t z = i32 % m2, y = (i32 / m2) % m1, x = i32 / (m1 * m2);
    const int64_t iu = (x * d.n1 + y) * d.n2 + z;
    const int64_t iu1 = P ? ((x + 1 < d.n0 ? x + 1 : 0) * d.n1 + y) * d.n2 + z : iu + d.n1 * d.n2;
    const int64_t iv = (x * (d.n1 + e) + y) * d.n2 + z;
    const int64_t iv1 = P ? (x * d.n1 + (y + 1 < d.n1 ? y + 1 : 0)) * d.n2 + z : iv + d.n2;
    const int64_t iw = (x * d.n1 + y) * (d.n2 + e) + z;
    const int64_t iw1 = P ? (x * d.n1 + y) * d.n2 + (z + 1 < d.n2 ? z + 1 : 0) : iw + 1;
    double s = (face_value<F>(u, fu, a, iu1) - face_value<F>(u, fu, a, iu)) / hx;
    s = s + (face_value<F>(v, fv, a, iv1) - face_value<F>(v, fv, a, iv)) / hy;
    s = s + (face_value<F>(w, fw, a, iw1) - face_value<F>(w, fw, a, iw)) / hz;
    return scale * s;
}

template <bool P, bool F = false>
__global__ void __launch_bounds__(256) divergence_kernel(const double* __restrict__ u,
                                                         const double* __restrict__ v,
                                                         const double* __restrict__ w,
                                                         double* __restrict__ out, Dims d,
                                                         double hx, double hy, double hz,
                                                         double scale,
                                                         const double* __restrict__ fu = nullptr,
                                                         const double* __restrict__ fv = nullptr,
                                                         const double* __restrict__ fw = nullptr,
                                                         double a = 0.0) {
    const int64_t total = d.n0 * d.n1 * d.n2;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += int64_t(gridDim.x) * blockDim.x)
        out[i] = divergence_at<P, F>(u, v, w, d, hx, hy, hz, scale, fu, fv, fw, a, i);
}

// Faces of component C per axis.
template <int C, bool P>
__device__ __forceinline__ void face_extents(const Dims& d, int64_t& f0, int64_t& f1, int64_t& f2) {
    f0 = d.n0 + (C == 0 && !P);
    f1 = d.n1 + (C == 1 && !P);
    f2 = d.n2 + (C == 2 && !P);
}

// Face i of component C: u = u* - coef * (p[q]-p[q-1])/h with u* = u + a f
// (F) or u (!F), on interior faces (all faces if periodic); boundary faces
// of a closed box keep u*.
template <int C, bool P, bool F>
__device__ __forceinline__ void grad_correct_at(double* __restrict__ u, const double* __restrict__ p,
                                                const Dims& d, double h, double coef,
                                                const double* __restrict__ f, double a, int64_t i) {
    int64_t f0, f1, f2;
    face_extents<C, P>(d, f0, f1, f2);
    // 32-bit index arithmetic (the entry points keep every array below 2^31)
    const uint32_t i32 = uint32_t(i), m1 = uint32_t(f1), m2 = uint32_t(f2);
    const int64_t z = i32 % m2, y = (i32 / m2) % m1, x = i32 / (m1 * m2);
    const int64_t q = C == 0 ? x : (C == 1 ? y : z);
    const int64_t nq = C == 0 ? d.n0 : (C == 1 ? d.n1 : d.n2);
    const double star = face_value<F>(u, f, a, i);
    if (!P && (q == 0 || q == nq)) {
        if (F) u[i] = star;
        return;
    }
    const int64_t qm = q > 0 ? q - 1 : nq - 1;  // face q lies between cells q-1 and q
    auto cell = [&](int64_t aa, int64_t b, int64_t c) { return (aa * d.n1 + b) * d.n2 + c; };
    int64_t hi, lo;
    if (C == 0) { hi = cell(q, y, z); lo = cell(qm, y, z); }
    else if (C == 1) { hi = cell(x, q, z); lo = cell(x, qm, z); }
    else { hi = cell(x, y, q); lo = cell(x, y, qm); }
    const double g = (p[hi] - p[lo]) / h;
    u[i] = star - coef * g;
}

template <int C, bool P, bool F = false>
__global__ void __launch_bounds__(256) grad_correct_kernel(double* __restrict__ u,
                                                           const double* __restrict__ p, Dims d,
                                                           double h, double coef,
                                                           const double* __restrict__ f = nullptr,
                                                           double a = 0.0) {
    int64_t f0, f1, f2;
    face_extents<C, P>(d, f0, f1, f2);
    const int64_t total = f0 * f1 * f2;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += int64_t(gridDim.x) * blockDim.x)
        grad_correct_at<C, P, F>(u, p, d, h, coef, f, a, i);
}

__global__ void axpy_kernel(double* __restrict__ y, const double* __restrict__ x, int64_t count,
                            double a) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
         i += int64_t(gridDim.x) * blockDim.x)
        y[i] = y[i] + a * x[i];
}

// ---- elementwise passes of many 3D IB instances, one launch each ----------
// Workgroup b of a pass handles elements [(b - begin) * IBB_CHUNK, + CHUNK)
// of the instance whose range holds b (binary search over the begins); each
// element is formed exactly as in the single-instance kernels above.

constexpr int IBB_CHUNK = 2048;

struct IbbDesc {
    double* u[3];
    const double* f[3];
    double* rhs;
    const double* p[2];
    double* x;
    const double* U;
    int64_t cells, faces[3], coords;
    Dims d;
    double h[3];
    int begin[3];  // first workgroup of the rhs, correct and advance passes
};

__device__ inline int ibb_find(const IbbDesc* __restrict__ desc, int n, int pass, int b) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (desc[mid].begin[pass] <= b) lo = mid;
        else hi = mid;
    }
    return __builtin_amdgcn_readfirstlane(lo);
}

template <bool P>
__global__ void __launch_bounds__(256) ibb_rhs_kernel(const IbbDesc* __restrict__ desc, int n, double a,
                                                      double scale) {
    const IbbDesc& s = desc[ibb_find(desc, n, 0, blockIdx.x)];
    const int64_t lo = int64_t(int(blockIdx.x) - s.begin[0]) * IBB_CHUNK;
    const int64_t hi = lo + IBB_CHUNK < s.cells ? lo + IBB_CHUNK : s.cells;
    for (int64_t i = lo + threadIdx.x; i < hi; i += 256)
        s.rhs[i] = divergence_at<P, true>(s.u[0], s.u[1], s.u[2], s.d, s.h[0], s.h[1], s.h[2], scale, s.f[0],
                                          s.f[1], s.f[2], a, i);
}

template <bool P>
__global__ void __launch_bounds__(256) ibb_correct_kernel(const IbbDesc* __restrict__ desc, int n, int which,
                                                          double a, double coef) {
    const IbbDesc& s = desc[ibb_find(desc, n, 1, blockIdx.x)];
    const int64_t total = s.faces[0] + s.faces[1] + s.faces[2];
    const int64_t lo = int64_t(int(blockIdx.x) - s.begin[1]) * IBB_CHUNK;
    const int64_t hi = lo + IBB_CHUNK < total ? lo + IBB_CHUNK : total;
    const double* p = s.p[which];
    for (int64_t j = lo + threadIdx.x; j < hi; j += 256) {
        if (j < s.faces[0]) grad_correct_at<0, P, true>(s.u[0], p, s.d, s.h[0], coef, s.f[0], a, j);
        else if (j < s.faces[0] + s.faces[1])
            grad_correct_at<1, P, true>(s.u[1], p, s.d, s.h[1], coef, s.f[1], a, j - s.faces[0]);
        else grad_correct_at<2, P, true>(s.u[2], p, s.d, s.h[2], coef, s.f[2], a, j - s.faces[0] - s.faces[1]);
    }
}

__global__ void __launch_bounds__(256) ibb_advance_kernel(const IbbDesc* __restrict__ desc, int n, double dt) {
    const IbbDesc& s = desc[ibb_find(desc, n, 2, blockIdx.x)];
    const int64_t lo = int64_t(int(blockIdx.x) - s.begin[2]) * IBB_CHUNK;
    const int64_t hi = lo + IBB_CHUNK < s.coords ? lo + IBB_CHUNK : s.coords;
    for (int64_t i = lo + threadIdx.x; i < hi; i += 256) s.x[i] = s.x[i] + dt * s.U[i];
}

// ---- host wrappers ---------------------------------------------------------

// Positive extents whose face arrays (n + 1 per axis at most) stay below
// 2^31 entries, so the elementwise passes index them with 32-bit arithmetic.
static bool dims_ok(const int64_t n[3]) {
    return n && n[0] > 0 && n[1] > 0 && n[2] > 0 && n[0] < (int64_t(1) << 31) && n[1] < (int64_t(1) << 31) &&
           n[2] < (int64_t(1) << 31) && (n[0] + 1) * (n[1] + 1) * (n[2] + 1) < (int64_t(1) << 31);
}

template <class T>
static Coef<T> make_coef(const double h2[3], const double* w) {
    Coef<T> k;
    for (int a = 0; a < 3; ++a) {
        k.h2[a] = T(h2[a]);
        k.w[a] = w ? T(w[a]) : T(0);
    }
    return k;
}

template <class T, int Mode>
static int stencil(const T* p, const T* rhs, T* out, const int64_t n[3], const double h2[3],
                   const double* w, double omega, int periodic, void* stream) {
    AFSI_REQUIRE(dims_ok(n) && h2 && p && rhs && out, "invalid multigrid arguments");
    const Dims d{n[0], n[1], n[2]};
    const Coef<T> k = make_coef<T>(h2, w);
    const int64_t total = d.n0 * d.n1 * d.n2;
    const unsigned g = grid_for(total, 256, 16384);
    const hipStream_t s = hipStream_t(stream);
    if (periodic) stencil_kernel<T, true, Mode><<<g, 256, 0, s>>>(p, rhs, out, d, k, T(omega));
    else stencil_kernel<T, false, Mode><<<g, 256, 0, s>>>(p, rhs, out, d, k, T(omega));
    AFSI_LAUNCH_CHECK();
    return OK;
}

template <class T>
static int rbgs(T* p, const T* rhs, const int64_t n[3], const double h2[3], const double w[3],
                int color, int periodic, void* stream, double relax = 1.0) {
    AFSI_REQUIRE(dims_ok(n) && h2 && w && p && rhs, "invalid multigrid arguments");
    AFSI_REQUIRE(n[2] % 2 == 0, "red-black smoothing needs an even z extent");
    AFSI_REQUIRE(!periodic || (n[0] % 2 == 0 && n[1] % 2 == 0), "periodic red-black needs even extents");
    const Dims d{n[0], n[1], n[2]};
    const Coef<T> k = make_coef<T>(h2, w);
    const unsigned g = grid_for(d.n0 * d.n1 * (d.n2 / 2), 256, 16384);
    const hipStream_t s = hipStream_t(stream);
    const T rw = T(relax > 0.0 ? relax : 1.0);
    if (periodic) rbgs_kernel<T, true><<<g, 256, 0, s>>>(p, rhs, d, k, color & 1, rw);
    else rbgs_kernel<T, false><<<g, 256, 0, s>>>(p, rhs, d, k, color & 1, rw);
    AFSI_LAUNCH_CHECK();
    return OK;
}

template <class T>
static int restrict3(const T* fine, T* coarse, const int64_t nc[3], void* stream) {
    AFSI_REQUIRE(dims_ok(nc) && fine && coarse, "invalid restriction arguments");
    const Dims c{nc[0], nc[1], nc[2]};
    restrict_kernel<T><<<grid_for(c.n0 * c.n1 * c.n2, 256, 16384), 256, 0, hipStream_t(stream)>>>(
        fine, coarse, c);
    AFSI_LAUNCH_CHECK();
    return OK;
}

template <class T>
static int residual_restrict3(const T* p, const T* rhs, T* coarse, const int64_t n[3],
                              const double h2[3], int periodic, void* stream) {
    AFSI_REQUIRE(dims_ok(n) && h2 && p && rhs && coarse, "invalid restriction arguments");
    AFSI_REQUIRE(n[0] % 2 == 0 && n[1] % 2 == 0 && n[2] % 2 == 0, "restriction needs even extents");
    const Dims d{n[0], n[1], n[2]};
    const Coef<T> k = make_coef<T>(h2, nullptr);
    const unsigned g = grid_for((d.n0 / 2) * (d.n1 / 2) * (d.n2 / 2), 256, 16384);
    const hipStream_t s = hipStream_t(stream);
    if (periodic) residual_restrict_kernel<T, true><<<g, 256, 0, s>>>(p, rhs, coarse, d, k);
    else residual_restrict_kernel<T, false><<<g, 256, 0, s>>>(p, rhs, coarse, d, k);
    AFSI_LAUNCH_CHECK();
    return OK;
}

template <class T>
static int prolong3(const T* coarse, T* fine, const int64_t n[3], int periodic, void* stream) {
    AFSI_REQUIRE(dims_ok(n) && coarse && fine, "invalid prolongation arguments");
    AFSI_REQUIRE(n[0] % 2 == 0 && n[1] % 2 == 0 && n[2] % 2 == 0, "prolongation needs even extents");
    const Dims d{n[0], n[1], n[2]};
    const unsigned g = grid_for(d.n0 * d.n1 * d.n2, 256, 16384);
    const hipStream_t s = hipStream_t(stream);
    if (periodic) prolong_add_kernel<T, true><<<g, 256, 0, s>>>(coarse, fine, d);
    else prolong_add_kernel<T, false><<<g, 256, 0, s>>>(coarse, fine, d);
    AFSI_LAUNCH_CHECK();
    return OK;
}

template <class T>
static int matvec(const T* A, const T* x, T* y, int64_t n, void* stream) {
    AFSI_REQUIRE(n > 0 && A && x && y, "invalid matvec arguments");
    matvec_kernel<T><<<grid_for(n * WAVE, 256), 256, 0, hipStream_t(stream)>>>(A, x, y, n);
    AFSI_LAUNCH_CHECK();
    return OK;
}

template <class T>
static int reduce(const T* x, int64_t count, double* partials, double* result, void* stream) {
    AFSI_REQUIRE(count >= 0 && partials && result && (x || count == 0), "invalid reduction");
    const hipStream_t s = hipStream_t(stream);
    const int64_t nb = reduce_blocks(count);
    reduce_partial_kernel<T><<<unsigned(nb), RED_THREADS, 0, s>>>(x, count, partials);
    AFSI_LAUNCH_CHECK();
    reduce_final_kernel<<<1, RED_THREADS, 0, s>>>(partials, nb, result);
    AFSI_LAUNCH_CHECK();
    return OK;
}

int launch_reduce_final(const double* partials, int64_t nparts, double* result, hipStream_t s) {
    reduce_final_kernel<<<1, RED_THREADS, 0, s>>>(partials, nparts, result);
    AFSI_LAUNCH_CHECK();
    return OK;
}

template <class T>
static int sub_mean(T* x, int64_t count, const double* sum, void* stream) {
    AFSI_REQUIRE(count > 0 && x && sum, "invalid mean removal");
    sub_mean_kernel<T><<<grid_for(count, 256, 16384), 256, 0, hipStream_t(stream)>>>(x, count, sum);
    AFSI_LAUNCH_CHECK();
    return OK;
}

}  // namespace afsi

using namespace afsi;

extern "C" {

int afsi_mg_residual_3d_f64(const double* p, const double* rhs, double* out, const int64_t n[3],
                            const double h2[3], int periodic, void* stream) {
    return stencil<double, 0>(p, rhs, out, n, h2, nullptr, 0.0, periodic, stream);
}
int afsi_mg_residual_3d_f32(const float* p, const float* rhs, float* out, const int64_t n[3],
                            const double h2[3], int periodic, void* stream) {
    return stencil<float, 0>(p, rhs, out, n, h2, nullptr, 0.0, periodic, stream);
}
int afsi_mg_jacobi_3d_f64(const double* p, const double* rhs, double* out, const int64_t n[3],
                          const double h2[3], const double inv_h2[3], double omega, int periodic,
                          void* stream) {
    AFSI_REQUIRE(inv_h2, "null inverse spacing");
    return stencil<double, 1>(p, rhs, out, n, h2, inv_h2, omega, periodic, stream);
}
int afsi_mg_jacobi_3d_f32(const float* p, const float* rhs, float* out, const int64_t n[3],
                          const double h2[3], const double inv_h2[3], double omega, int periodic,
                          void* stream) {
    AFSI_REQUIRE(inv_h2, "null inverse spacing");
    return stencil<float, 1>(p, rhs, out, n, h2, inv_h2, omega, periodic, stream);
}
int afsi_mg_rbgs_3d_f64(double* p, const double* rhs, const int64_t n[3], const double h2[3],
                        const double inv_h2[3], int color, int periodic, void* stream) {
    return rbgs<double>(p, rhs, n, h2, inv_h2, color, periodic, stream);
}
int afsi_mg_rbgs_3d_f32(float* p, const float* rhs, const int64_t n[3], const double h2[3],
                        const double inv_h2[3], int color, int periodic, void* stream) {
    return rbgs<float>(p, rhs, n, h2, inv_h2, color, periodic, stream);
}
int afsi_mg_restrict_3d_f64(const double* fine, double* coarse, const int64_t nc[3], void* stream) {
    return restrict3<double>(fine, coarse, nc, stream);
}
int afsi_mg_restrict_3d_f32(const float* fine, float* coarse, const int64_t nc[3], void* stream) {
    return restrict3<float>(fine, coarse, nc, stream);
}
int afsi_mg_residual_restrict_3d_f64(const double* p, const double* rhs, double* coarse,
                                     const int64_t n[3], const double h2[3], int periodic,
                                     void* stream) {
    return residual_restrict3<double>(p, rhs, coarse, n, h2, periodic, stream);
}
int afsi_mg_residual_restrict_3d_f32(const float* p, const float* rhs, float* coarse,
                                     const int64_t n[3], const double h2[3], int periodic,
                                     void* stream) {
    return residual_restrict3<float>(p, rhs, coarse, n, h2, periodic, stream);
}
int afsi_mg_prolong_add_3d_f64(const double* coarse, double* fine, const int64_t n[3],
                               int periodic, void* stream) {
    return prolong3<double>(coarse, fine, n, periodic, stream);
}
int afsi_mg_prolong_add_3d_f32(const float* coarse, float* fine, const int64_t n[3], int periodic,
                               void* stream) {
    return prolong3<float>(coarse, fine, n, periodic, stream);
}
int afsi_dense_matvec_f64(const double* A, const double* x, double* y, int64_t n, void* stream) {
    return matvec<double>(A, x, y, n, stream);
}
int afsi_dense_matvec_f32(const float* A, const float* x, float* y, int64_t n, void* stream) {
    return matvec<float>(A, x, y, n, stream);
}
int64_t afsi_reduce_partials(int64_t count) { return reduce_blocks(count); }
int afsi_reduce_f64(const double* x, int64_t count, double* partials, double* result,
                    void* stream) {
    return reduce<double>(x, count, partials, result, stream);
}
int afsi_reduce_f32(const float* x, int64_t count, double* partials, double* result,
                    void* stream) {
    return reduce<float>(x, count, partials, result, stream);
}
int afsi_sub_mean_f64(double* x, int64_t count, const double* sum, void* stream) {
    return sub_mean<double>(x, count, sum, stream);
}
int afsi_sub_mean_f32(float* x, int64_t count, const double* sum, void* stream) {
    return sub_mean<float>(x, count, sum, stream);
}

int afsi_divergence_3d_f64(const double* u, const double* v, const double* w, double* out,
                           const int64_t n[3], const double h[3], double scale, int periodic,
                           void* stream) {
    AFSI_REQUIRE(dims_ok(n) && h && u && v && w && out, "invalid divergence arguments");
    const Dims d{n[0], n[1], n[2]};
    const unsigned g = grid_for(d.n0 * d.n1 * d.n2, 256, 16384);
    const hipStream_t s = hipStream_t(stream);
    if (periodic) divergence_kernel<true><<<g, 256, 0, s>>>(u, v, w, out, d, h[0], h[1], h[2], scale);
    else divergence_kernel<false><<<g, 256, 0, s>>>(u, v, w, out, d, h[0], h[1], h[2], scale);
    AFSI_LAUNCH_CHECK();
    return OK;
}

int afsi_grad_correct_3d_f64(double* u, double* v, double* w, const double* p, const int64_t n[3],
                             const double h[3], double coef, int periodic, void* stream) {
    AFSI_REQUIRE(dims_ok(n) && h && u && v && w && p, "invalid gradient arguments");
    const Dims d{n[0], n[1], n[2]};
    const hipStream_t s = hipStream_t(stream);
    const int64_t cells = d.n0 * d.n1 * d.n2;
    const unsigned g = grid_for(cells + d.n1 * d.n2 + d.n0 * d.n2 + d.n0 * d.n1, 256, 16384);
    if (periodic) {
        grad_correct_kernel<0, true><<<g, 256, 0, s>>>(u, p, d, h[0], coef);
        grad_correct_kernel<1, true><<<g, 256, 0, s>>>(v, p, d, h[1], coef);
        grad_correct_kernel<2, true><<<g, 256, 0, s>>>(w, p, d, h[2], coef);
    } else {
        grad_correct_kernel<0, false><<<g, 256, 0, s>>>(u, p, d, h[0], coef);
        grad_correct_kernel<1, false><<<g, 256, 0, s>>>(v, p, d, h[1], coef);
        grad_correct_kernel<2, false><<<g, 256, 0, s>>>(w, p, d, h[2], coef);
    }
    AFSI_LAUNCH_CHECK();
    return OK;
}

int afsi_project_rhs_3d_f64(const double* u, const double* v, const double* w, const double* fu,
                            const double* fv, const double* fw, double a, double* out,
                            const int64_t n[3], const double h[3], double scale, int periodic,
                            void* stream) {
    AFSI_REQUIRE(dims_ok(n) && h && u && v && w && fu && fv && fw && out,
                 "invalid projection arguments");
    const Dims d{n[0], n[1], n[2]};
    const unsigned g = grid_for(d.n0 * d.n1 * d.n2, 256, 16384);
    const hipStream_t s = hipStream_t(stream);
    if (periodic)
        divergence_kernel<true, true><<<g, 256, 0, s>>>(u, v, w, out, d, h[0], h[1], h[2], scale, fu, fv, fw, a);
    else
        divergence_kernel<false, true><<<g, 256, 0, s>>>(u, v, w, out, d, h[0], h[1], h[2], scale, fu, fv, fw, a);
    AFSI_LAUNCH_CHECK();
    return OK;
}

int afsi_project_correct_3d_f64(double* u, double* v, double* w, const double* fu,
                                const double* fv, const double* fw, double a, const double* p,
                                const int64_t n[3], const double h[3], double coef, int periodic,
                                void* stream) {
    AFSI_REQUIRE(dims_ok(n) && h && u && v && w && fu && fv && fw && p, "invalid projection arguments");
    const Dims d{n[0], n[1], n[2]};
    const hipStream_t s = hipStream_t(stream);
    const int64_t cells = d.n0 * d.n1 * d.n2;
    const unsigned g = grid_for(cells + d.n1 * d.n2 + d.n0 * d.n2 + d.n0 * d.n1, 256, 16384);
    if (periodic) {
        grad_correct_kernel<0, true, true><<<g, 256, 0, s>>>(u, p, d, h[0], coef, fu, a);
        grad_correct_kernel<1, true, true><<<g, 256, 0, s>>>(v, p, d, h[1], coef, fv, a);
        grad_correct_kernel<2, true, true><<<g, 256, 0, s>>>(w, p, d, h[2], coef, fw, a);
    } else {
        grad_correct_kernel<0, false, true><<<g, 256, 0, s>>>(u, p, d, h[0], coef, fu, a);
        grad_correct_kernel<1, false, true><<<g, 256, 0, s>>>(v, p, d, h[1], coef, fv, a);
        grad_correct_kernel<2, false, true><<<g, 256, 0, s>>>(w, p, d, h[2], coef, fw, a);
    }
    AFSI_LAUNCH_CHECK();
    return OK;
}

int afsi_axpy_f64(double* y, const double* x, int64_t count, double a, void* stream) {
    AFSI_REQUIRE(count >= 0 && (count == 0 || (x && y)), "invalid axpy");
    if (count == 0) return OK;
    axpy_kernel<<<grid_for(count, 256, 16384), 256, 0, hipStream_t(stream)>>>(y, x, count, a);
    AFSI_LAUNCH_CHECK();
    return OK;
}

}  // extern "C"

// ---- native V-cycle driver ---------------------------------------------------

namespace afsi {

template <class T>
struct PlanRunner {
    const afsi_mg3d_plan& pl;
    hipStream_t s;

    T* P(int l) const { return static_cast<T*>(pl.p[l]); }
    T* R(int l) const { return static_cast<T*>(pl.rhs[l]); }
    T* W(int l) const { return static_cast<T*>(pl.tmp[l]); }
    int64_t cells(int l) const { return pl.n[l][0] * pl.n[l][1] * pl.n[l][2]; }

    int remove_mean(T* x, int64_t count) const {
        if (int r = reduce<T>(x, count, pl.reduce_partials, pl.reduce_result, s)) return r;
        return sub_mean<T>(x, count, pl.reduce_result, s);
    }

    int smooth(int l) const {
        const int64_t* n = pl.n[l];
        if (pl.smoother == 1) {
            for (int k = 0; k < pl.nu; ++k)
                for (int color = 0; color < 2; ++color)
                    if (int r = rbgs<T>(P(l), R(l), n, pl.h2[l], pl.inv_h2[l], color, pl.periodic, s, pl.relax))
                        return r;
        } else {
            T* a = P(l);
            T* b = W(l);
            for (int k = 0; k < pl.nu; ++k) {
                if (int r = stencil<T, 1>(a, R(l), b, n, pl.h2[l], pl.inv_h2[l], pl.omega,
                                          pl.periodic, s))
                    return r;
                T* t = a; a = b; b = t;
            }
            if (a != P(l)) AFSI_HIP(hipMemcpyAsync(P(l), a, size_t(cells(l)) * sizeof(T),
                                                   hipMemcpyDeviceToDevice, s));
        }
        return remove_mean(P(l), cells(l));
    }

    int cycle(int l) const {
        const int last = pl.levels - 1;
        if (l == last) {
            if (int r = matvec<T>(static_cast<const T*>(pl.coarse_inverse), R(l), P(l),
                                  pl.coarse_cells, s))
                return r;
            return remove_mean(P(l), cells(l));
        }
        if (int r = smooth(l)) return r;
        if (int r = residual_restrict3<T>(P(l), R(l), R(l + 1), pl.n[l], pl.h2[l], pl.periodic, s))
            return r;
        if (int r = remove_mean(R(l + 1), cells(l + 1))) return r;
        if (int r = launch_fill<T>(P(l + 1), cells(l + 1), T(0), s)) return r;
        if (int r = cycle(l + 1)) return r;
        if (int r = prolong3<T>(P(l + 1), P(l), pl.n[l], pl.periodic, s)) return r;
        return smooth(l);
    }
};

template <class T>
static int plan_check(const afsi_mg3d_plan* pl) {
    AFSI_REQUIRE(pl && pl->levels >= 1 && pl->levels <= AFSI_MG_MAX_LEVELS, "invalid multigrid plan");
    AFSI_REQUIRE(pl->nu >= 1 && (pl->smoother == 0 || pl->smoother == 1), "invalid smoother");
    AFSI_REQUIRE(pl->reduce_partials && pl->reduce_result, "missing reduction scratch");
    for (int l = 0; l < pl->levels; ++l) {
        AFSI_REQUIRE(dims_ok(pl->n[l]) && pl->p[l] && pl->rhs[l], "invalid level %d", l);
        if (l + 1 < pl->levels) {
            AFSI_REQUIRE(pl->tmp[l] || pl->smoother == 1, "missing Jacobi buffer at level %d", l);
            for (int a = 0; a < 3; ++a)
                AFSI_REQUIRE(pl->n[l][a] == 2 * pl->n[l + 1][a], "levels %d/%d are not 2:1", l, l + 1);
        }
    }
    const int last = pl->levels - 1;
    AFSI_REQUIRE(pl->coarse_inverse &&
                     pl->coarse_cells == pl->n[last][0] * pl->n[last][1] * pl->n[last][2],
                 "coarse inverse does not match the coarsest level");
    return OK;
}

template <class T>
static int plan_cycles(const afsi_mg3d_plan* pl, int32_t cycles, void* stream) {
    if (int r = plan_check<T>(pl)) return r;
    AFSI_REQUIRE(cycles >= 0, "negative cycle count");
    const PlanRunner<T> run{*pl, hipStream_t(stream)};
    for (int c = 0; c < cycles; ++c)
        if (int r = run.cycle(0)) return r;
    return OK;
}

// Descriptors of a batch of IB instances (host side) and the workgroups of
// each pass.
struct IbbLayout {
    std::vector<IbbDesc> desc;
    int blocks[3] = {0, 0, 0};
    int periodic = 0;
};

static int ibb_layout(const afsi_ib3b_batch* bt, IbbLayout& L) {
    AFSI_REQUIRE(bt && bt->items && bt->count >= 1, "invalid IB batch");
    L.desc.resize(size_t(bt->count));
    L.periodic = bt->items[0].grid.periodic;
    for (int i = 0; i < bt->count; ++i) {
        const afsi_ib3b_item& it = bt->items[i];
        const afsi_grid3& g = it.grid;
        AFSI_REQUIRE(g.periodic == L.periodic, "instances of one IB batch share the boundary type");
        AFSI_REQUIRE(dims_ok(g.n) && it.points >= 0, "invalid grid or point count (instance %d)", i);
        for (int c = 0; c < 3; ++c)
            AFSI_REQUIRE(it.velocity[c] && it.force[c], "missing face array (instance %d)", i);
        AFSI_REQUIRE(it.rhs && it.pressure[0] && it.pressure[1], "missing pressure buffers (instance %d)", i);
        AFSI_REQUIRE(it.points == 0 || (it.x && it.point_velocity), "missing point arrays (instance %d)", i);
        IbbDesc& s = L.desc[size_t(i)];
        for (int c = 0; c < 3; ++c) {
            s.u[c] = it.velocity[c];
            s.f[c] = it.force[c];
            s.h[c] = g.h[c];
        }
        s.rhs = it.rhs;
        s.p[0] = it.pressure[0];
        s.p[1] = it.pressure[1];
        s.x = it.x;
        s.U = it.point_velocity;
        s.d = Dims{g.n[0], g.n[1], g.n[2]};
        s.cells = g.n[0] * g.n[1] * g.n[2];
        const int e = g.periodic ? 0 : 1;
        s.faces[0] = (g.n[0] + e) * g.n[1] * g.n[2];
        s.faces[1] = g.n[0] * (g.n[1] + e) * g.n[2];
        s.faces[2] = g.n[0] * g.n[1] * (g.n[2] + e);
        s.coords = 3 * it.points;
        const int64_t work[3] = {s.cells, s.faces[0] + s.faces[1] + s.faces[2], s.coords};
        for (int q = 0; q < 3; ++q) {
            s.begin[q] = L.blocks[q];
            const int64_t nb = cdiv(work[q], IBB_CHUNK);
            AFSI_REQUIRE(int64_t(L.blocks[q]) + nb < (int64_t(1) << 31), "IB batch too large");
            L.blocks[q] += int(nb);
        }
    }
    return OK;
}

static int ibb_ready(const afsi_ib3b_batch* bt, IbbLayout& L) {
    if (int r = ibb_layout(bt, L)) return r;
    AFSI_REQUIRE(bt->workspace && bt->workspace_bytes >= L.desc.size() * sizeof(IbbDesc),
                 "IB batch workspace too small");
    return OK;
}

template <class T>
static int plan_residual(const afsi_mg3d_plan* pl, void* stream) {
    if (int r = plan_check<T>(pl)) return r;
    AFSI_REQUIRE(pl->tmp[0], "missing level-0 scratch");
    const hipStream_t s = hipStream_t(stream);
    const T* p = static_cast<const T*>(pl->p[0]);
    if (int r = stencil<T, 0>(p, static_cast<const T*>(pl->rhs[0]), static_cast<T*>(pl->tmp[0]),
                              pl->n[0], pl->h2[0], nullptr, 0.0, pl->periodic, s))
        return r;
    const int64_t count = pl->n[0][0] * pl->n[0][1] * pl->n[0][2];
    return reduce<T>(static_cast<const T*>(pl->tmp[0]), count, pl->reduce_partials,
                     pl->reduce_result, s);
}

}  // namespace afsi

extern "C" {
int afsi_mg3d_cycles_unfused_f64(const afsi_mg3d_plan* plan, int32_t cycles, void* stream) {
    return afsi::plan_cycles<double>(plan, cycles, stream);
}
int afsi_mg3d_cycles_unfused_f32(const afsi_mg3d_plan* plan, int32_t cycles, void* stream) {
    return afsi::plan_cycles<float>(plan, cycles, stream);
}
int afsi_mg3d_residual_unfused_f64(const afsi_mg3d_plan* plan, void* stream) {
    return afsi::plan_residual<double>(plan, stream);
}
int afsi_mg3d_residual_unfused_f32(const afsi_mg3d_plan* plan, void* stream) {
    return afsi::plan_residual<float>(plan, stream);
}

size_t afsi_ib3b_workspace_bytes(const afsi_ib3b_batch* batch) {
    afsi::IbbLayout L;
    return afsi::ibb_layout(batch, L) == afsi::OK ? L.desc.size() * sizeof(afsi::IbbDesc) : 0;
}
int afsi_ib3b_prepare(const afsi_ib3b_batch* batch, void* stream) {
    using namespace afsi;
    IbbLayout L;
    if (int r = ibb_ready(batch, L)) return r;
    const hipStream_t s = hipStream_t(stream);
    AFSI_HIP(hipMemcpyAsync(batch->workspace, L.desc.data(), L.desc.size() * sizeof(IbbDesc),
                            hipMemcpyHostToDevice, s));
    AFSI_HIP(hipStreamSynchronize(s));
    return OK;
}
int afsi_ib3b_rhs(const afsi_ib3b_batch* batch, double a, double scale, void* stream) {
    using namespace afsi;
    IbbLayout L;
    if (int r = ibb_ready(batch, L)) return r;
    const IbbDesc* d = static_cast<const IbbDesc*>(batch->workspace);
    const hipStream_t s = hipStream_t(stream);
    if (L.periodic) ibb_rhs_kernel<true><<<unsigned(L.blocks[0]), 256, 0, s>>>(d, batch->count, a, scale);
    else ibb_rhs_kernel<false><<<unsigned(L.blocks[0]), 256, 0, s>>>(d, batch->count, a, scale);
    AFSI_LAUNCH_CHECK();
    return OK;
}
int afsi_ib3b_correct(const afsi_ib3b_batch* batch, int32_t which, double a, double coef, void* stream) {
    using namespace afsi;
    IbbLayout L;
    if (int r = ibb_ready(batch, L)) return r;
    AFSI_REQUIRE(which == 0 || which == 1, "invalid pressure index");
    const IbbDesc* d = static_cast<const IbbDesc*>(batch->workspace);
    const hipStream_t s = hipStream_t(stream);
    if (L.periodic)
        ibb_correct_kernel<true><<<unsigned(L.blocks[1]), 256, 0, s>>>(d, batch->count, which, a, coef);
    else
        ibb_correct_kernel<false><<<unsigned(L.blocks[1]), 256, 0, s>>>(d, batch->count, which, a, coef);
    AFSI_LAUNCH_CHECK();
    return OK;
}
int afsi_ib3b_advance(const afsi_ib3b_batch* batch, double dt, void* stream) {
    using namespace afsi;
    IbbLayout L;
    if (int r = ibb_ready(batch, L)) return r;
    if (L.blocks[2] == 0) return OK;
    ibb_advance_kernel<<<unsigned(L.blocks[2]), 256, 0, hipStream_t(stream)>>>(
        static_cast<const IbbDesc*>(batch->workspace), batch->count, dt);
    AFSI_LAUNCH_CHECK();
    return OK;
}
}  // extern "C"
