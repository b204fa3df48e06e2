// Stencil preparation, compact separable and generic indexed IB transfer.
//
// Layouts follow the reference op boundary (mac/csrc/transfer.cpp:49-52):
// compact stencils base (P,3) int64 + phi (P,3,4) per component; face arrays
// flat C-order (z fastest). One 64-lane wavefront per point: lane k owns link
// (k/16, (k/4)%4, k%4), so a point's 64 links map onto exactly one wave64
// with no idle lanes and the gather reduction is a single wave reduction.
#include "common.hpp"
#include <cstring>
#include <string>

namespace afsi {

static thread_local char g_error[512] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
}

// ---------------------------------------------------------------------------
// prepare: one thread per (component, point, axis)

__global__ void prepare_compact_kernel(const double* __restrict__ x, int64_t points,
                                       afsi_grid3 g, int64_t* __restrict__ base,
                                       double* __restrict__ phi, int* flags) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= 3 * points * 3) return;
    const int a = int(t % 3);
    const int64_t p = (t / 3) % points;
    const int c = int(t / (3 * points));
    const double xa = x[3 * p + a];
    const double o = (a == c) ? g.origin[a] : g.center_origin[a];
    const double scaled = (xa - o) / g.h[a];
    if (c == 0 && !g.periodic) {
        // mac/compact_transfer.py:94-96,115-117 on the box-origin scale.
        const double s0 = (xa - g.origin[a]) / g.h[a];
        if (!(s0 >= 2.0) || !(s0 < double(g.n[a]) - 2.0)) raise_flag(flags, F_SUPPORT);
    }
    const double fl = floor(scaled - 1.0);
    int64_t b = isfinite(fl) ? int64_t(fl) : 0;
    base[t] = b;
    double* ph = phi + 4 * t;
#pragma unroll
    for (int k = 0; k < 4; ++k) ph[k] = peskin_phi(scaled - double(b + k));
}

// ---------------------------------------------------------------------------
// compact gather / spread, one wave per point, lane = link

template <class T>
__device__ inline T wave_sum(T v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, WAVE);
    return v;
}

template <class T, bool Spread>
__global__ void __launch_bounds__(256) compact_kernel(const T* __restrict__ f, int64_t fstride,
                                                      const int64_t* __restrict__ base,
                                                      const T* __restrict__ phi, T* out,
                                                      int64_t ostride, int64_t points,
                                                      int64_t nx, int64_t ny, int64_t nz,
                                                      T volume, int* flags) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int64_t p = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / WAVE;
    if (p >= points) return;
    const int i = lane >> 4, j = (lane >> 2) & 3, l = lane & 3;
    const int64_t xi = base[3 * p] + i, yj = base[3 * p + 1] + j, zl = base[3 * p + 2] + l;
    const bool ok = xi >= 0 && xi < nx && yj >= 0 && yj < ny && zl >= 0 && zl < nz;
    const T wgt = (phi[12 * p + i] * phi[12 * p + 4 + j]) * phi[12 * p + 8 + l];
    const int64_t id = (xi * ny + yj) * nz + zl;
    if (!ok) raise_flag(flags, F_INDEX);
    if constexpr (Spread) {
        if (ok) unsafeAtomicAdd(out + id, (wgt * f[p * fstride]) / volume);
    } else {
        T s = ok ? f[id] * wgt : T(0);
        s = wave_sum(s);
        if (lane == 0) out[p * ostride] = s;
    }
}

template <class T>
__global__ void fill_kernel(T* x, int64_t count, T value) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
         i += int64_t(gridDim.x) * blockDim.x)
        x[i] = value;
}

// ---------------------------------------------------------------------------
// indexed gather / spread, one wave per (point, component), lanes over K

template <class T, bool Spread>
__global__ void __launch_bounds__(256) indexed_kernel(const T* __restrict__ f,
                                                      const int64_t* __restrict__ ids,
                                                      const T* __restrict__ w, T* out,
                                                      int64_t points, int64_t neighbors,
                                                      int64_t comps, int64_t nodes, T volume,
                                                      int* flags) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int64_t task = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / WAVE;
    if (task >= points * comps) return;
    const int64_t p = task / comps, c = task % comps;
    T s = T(0);
    for (int64_t k = lane; k < neighbors; k += WAVE) {
        const int64_t j = ids[p * neighbors + k];
        if (j < 0 || j >= nodes) { raise_flag(flags, F_INDEX); continue; }
        const T wt = w[p * neighbors + k];
        if constexpr (Spread) unsafeAtomicAdd(out + j * comps + c, (f[p * comps + c] * wt) / volume);
        else s += f[j * comps + c] * wt;
    }
    if constexpr (!Spread) {
        s = wave_sum(s);
        if (lane == 0) out[p * comps + c] = s;
    }
}

template <class T>
int launch_fill(T* x, int64_t count, T value, hipStream_t s) {
    if (count <= 0) return OK;
    fill_kernel<T><<<grid_for(count, 256, 4096), 256, 0, s>>>(x, count, value);
    AFSI_LAUNCH_CHECK();
    return OK;
}
template int launch_fill<double>(double*, int64_t, double, hipStream_t);
template int launch_fill<float>(float*, int64_t, float, hipStream_t);

template <class T>
static int compact_gather(const T* field, const int64_t* base, const T* phi, int64_t points,
                          int64_t nx, int64_t ny, int64_t nz, T* out, int64_t stride,
                          int* flags, void* stream) {
    AFSI_REQUIRE(points >= 0 && nx > 0 && ny > 0 && nz > 0 && stride > 0,
                 "invalid compact IB field/base");
    if (points == 0) return OK;
    AFSI_REQUIRE(field && base && phi && out, "null pointer");
    const hipStream_t s = hipStream_t(stream);
    compact_kernel<T, false><<<grid_for(points * WAVE, 256), 256, 0, s>>>(
        field, 0, base, phi, out, stride, points, nx, ny, nz, T(1), flags);
    AFSI_LAUNCH_CHECK();
    return OK;
}

template <class T>
static int compact_spread(const T* force, int64_t fstride, const int64_t* base, const T* phi,
                          int64_t points, int64_t nx, int64_t ny, int64_t nz, double volume,
                          T* out, int* flags, void* stream) {
    AFSI_REQUIRE(points >= 0 && nx > 0 && ny > 0 && nz > 0 && fstride > 0 &&
                     std::isfinite(volume) && volume > 0,
                 "invalid compact grid/volume");
    AFSI_REQUIRE(out, "null pointer");
    const hipStream_t s = hipStream_t(stream);
    if (int r = launch_fill<T>(out, nx * ny * nz, T(0), s)) return r;
    if (points == 0) return OK;
    AFSI_REQUIRE(force && base && phi, "null pointer");
    compact_kernel<T, true><<<grid_for(points * WAVE, 256), 256, 0, s>>>(
        force, fstride, base, phi, out, 0, points, nx, ny, nz, T(volume), flags);
    AFSI_LAUNCH_CHECK();
    return OK;
}

template <class T>
static int indexed_gather(const T* field, int64_t nodes, int64_t comps, const int64_t* ids,
                          const T* w, int64_t points, int64_t neighbors, T* out, int* flags,
                          void* stream) {
    AFSI_REQUIRE(nodes > 0, "empty IB grid");
    AFSI_REQUIRE(comps > 0 && neighbors > 0 && points >= 0, "IB indexed stencil shape mismatch");
    if (points == 0) return OK;
    AFSI_REQUIRE(field && ids && w && out, "null pointer");
    const hipStream_t s = hipStream_t(stream);
    indexed_kernel<T, false><<<grid_for(points * comps * WAVE, 256), 256, 0, s>>>(
        field, ids, w, out, points, neighbors, comps, nodes, T(1), flags);
    AFSI_LAUNCH_CHECK();
    return OK;
}

template <class T>
static int indexed_spread(const T* field, int64_t points, int64_t comps, const int64_t* ids,
                          const T* w, int64_t neighbors, int64_t size, double volume, T* out,
                          int* flags, void* stream) {
    AFSI_REQUIRE(size > 0 && std::isfinite(volume) && volume > 0, "invalid IB output size/volume");
    AFSI_REQUIRE(comps > 0 && neighbors > 0 && points >= 0, "IB indexed stencil shape mismatch");
    AFSI_REQUIRE(out, "null pointer");
    const hipStream_t s = hipStream_t(stream);
    if (int r = launch_fill<T>(out, size * comps, T(0), s)) return r;
    if (points == 0) return OK;
    AFSI_REQUIRE(field && ids && w, "null pointer");
    indexed_kernel<T, true><<<grid_for(points * comps * WAVE, 256), 256, 0, s>>>(
        field, ids, w, out, points, neighbors, comps, size, T(volume), flags);
    AFSI_LAUNCH_CHECK();
    return OK;
}

}  // namespace afsi

using namespace afsi;

extern "C" {

const char* afsi_last_error(void) { return afsi::g_error; }
int afsi_abi_version(void) { return 1; }

int afsi_prepare_compact_f64(const double* x, int64_t points, const afsi_grid3* grid,
                             int64_t* base, double* phi, int* flags, void* stream) {
    AFSI_REQUIRE(grid && points >= 0, "invalid prepare arguments");
    AFSI_REQUIRE(grid->n[0] > 0 && grid->n[1] > 0 && grid->n[2] > 0, "invalid grid");
    if (points == 0) return OK;
    AFSI_REQUIRE(x && base && phi, "null pointer");
    const int64_t work = 9 * points;
    prepare_compact_kernel<<<grid_for(work, 256), 256, 0, hipStream_t(stream)>>>(
        x, points, *grid, base, phi, flags);
    AFSI_LAUNCH_CHECK();
    return OK;
}

int afsi_compact_gather_f64(const double* field, const int64_t* base, const double* phi,
                            int64_t points, int64_t nx, int64_t ny, int64_t nz, double* out,
                            int64_t out_stride, int* flags, void* stream) {
    return compact_gather<double>(field, base, phi, points, nx, ny, nz, out, out_stride, flags,
                                  stream);
}
int afsi_compact_gather_f32(const float* field, const int64_t* base, const float* phi,
                            int64_t points, int64_t nx, int64_t ny, int64_t nz, float* out,
                            int64_t out_stride, int* flags, void* stream) {
    return compact_gather<float>(field, base, phi, points, nx, ny, nz, out, out_stride, flags,
                                 stream);
}
int afsi_compact_spread_f64(const double* force, int64_t force_stride, const int64_t* base,
                            const double* phi, int64_t points, int64_t nx, int64_t ny,
                            int64_t nz, double volume, double* out, int* flags, void* stream) {
    return compact_spread<double>(force, force_stride, base, phi, points, nx, ny, nz, volume,
                                  out, flags, stream);
}
int afsi_compact_spread_f32(const float* force, int64_t force_stride, const int64_t* base,
                            const float* phi, int64_t points, int64_t nx, int64_t ny, int64_t nz,
                            double volume, float* out, int* flags, void* stream) {
    return compact_spread<float>(force, force_stride, base, phi, points, nx, ny, nz, volume,
                                 out, flags, stream);
}

int afsi_indexed_gather_f64(const double* field, int64_t nodes, int64_t comps,
                            const int64_t* ids, const double* weights, int64_t points,
                            int64_t neighbors, double* out, int* flags, void* stream) {
    return indexed_gather<double>(field, nodes, comps, ids, weights, points, neighbors, out,
                                  flags, stream);
}
int afsi_indexed_gather_f32(const float* field, int64_t nodes, int64_t comps, const int64_t* ids,
                            const float* weights, int64_t points, int64_t neighbors, float* out,
                            int* flags, void* stream) {
    return indexed_gather<float>(field, nodes, comps, ids, weights, points, neighbors, out,
                                 flags, stream);
}
int afsi_indexed_spread_f64(const double* field, int64_t points, int64_t comps,
                            const int64_t* ids, const double* weights, int64_t neighbors,
                            int64_t size, double volume, double* out, int* flags, void* stream) {
    return indexed_spread<double>(field, points, comps, ids, weights, neighbors, size, volume,
                                  out, flags, stream);
}
int afsi_indexed_spread_f32(const float* field, int64_t points, int64_t comps, const int64_t* ids,
                            const float* weights, int64_t neighbors, int64_t size, double volume,
                            float* out, int* flags, void* stream) {
    return indexed_spread<float>(field, points, comps, ids, weights, neighbors, size, volume, out,
                                 flags, stream);
}

int afsi_fill_f64(double* x, int64_t count, double value, void* stream) {
    AFSI_REQUIRE(count >= 0 && (x || count == 0), "invalid fill");
    return launch_fill<double>(x, count, value, hipStream_t(stream));
}
int afsi_fill_f32(float* x, int64_t count, double value, void* stream) {
    AFSI_REQUIRE(count >= 0 && (x || count == 0), "invalid fill");
    return launch_fill<float>(x, count, float(value), hipStream_t(stream));
}

}  // extern "C"
