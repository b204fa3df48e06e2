// Shared definitions for the afsi_hip C-ABI library (gfx950 / CDNA4).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstdarg>
#include <cmath>
#include "afsi_hip.h"

namespace afsi {

// Status codes returned by every entry point.
enum : int { OK = 0, E_ARG = 1, E_HIP = 2, E_LIMIT = 3 };

// Per-thread last error message (read through afsi_last_error()).
void set_error(const char* fmt, ...);

#define AFSI_REQUIRE(cond, ...)                                                  \
    do {                                                                         \
        if (!(cond)) { ::afsi::set_error(__VA_ARGS__); return ::afsi::E_ARG; }   \
    } while (0)

#define AFSI_HIP(expr)                                                           \
    do {                                                                         \
        hipError_t e_ = (expr);                                                  \
        if (e_ != hipSuccess) {                                                  \
            ::afsi::set_error("%s: %s", #expr, hipGetErrorString(e_));           \
            return ::afsi::E_HIP;                                                \
        }                                                                        \
    } while (0)

// Check the launch that was just issued.
#define AFSI_LAUNCH_CHECK() AFSI_HIP(hipGetLastError())

// Raise kernel fn's dynamic-LDS limit to what the device allows (once per
// kernel and device); E_LIMIT when `need` bytes do not fit.
int dynamic_lds(const void* fn, size_t need);

constexpr int WAVE = 64;

// Device error bits written by kernels into a caller-provided int word.
enum : int { F_SUPPORT = 1, F_INDEX = 2, F_NONFINITE = 4, F_CAPACITY = 8 };

__device__ inline void raise_flag(int* flags, int bit) {
    if (flags) atomicOr(flags, bit);
}

__host__ __device__ inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

inline unsigned grid_for(int64_t work, int block, int64_t cap = 1 << 20) {
    int64_t g = cdiv(work, block);
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return unsigned(g);
}

// Four Peskin weights for offsets 0..3 from f = s - base - 1 in [0,1):
// |r| = 1+f, f, 1-f, 2-f all share the radical sqrt(1+4f-4f^2).
template <class T>
__device__ inline void peskin_weights(T f, T w[4]) {
    const T R = sqrt(T(1) + T(4) * f - T(4) * (f * f));
    const T t = T(3) - T(2) * f;
    const T u = T(1) + T(2) * f;
    w[0] = (t - R) * T(0.125);
    w[1] = (t + R) * T(0.125);
    w[2] = (u + R) * T(0.125);
    w[3] = (u - R) * T(0.125);
}

// phi(r) with the two clamped branches of the reference kernel (ib.py:57-67).
__device__ inline double peskin_phi(double d) {
    const double r = fabs(d);
    const double a = fmin(r, 1.0);
    const double b = fmin(fmax(r, 1.0), 2.0);
    const double inner = (3.0 - 2.0 * a + sqrt(1.0 + 4.0 * a - 4.0 * (a * a))) / 8.0;
    const double outer = (5.0 - 2.0 * b - sqrt(-7.0 + 12.0 * b - 4.0 * (b * b))) / 8.0;
    return r < 1.0 ? inner : (r < 2.0 ? outer : 0.0);
}

__device__ inline int64_t wrap(int64_t i, int64_t n) {
    int64_t r = i % n;
    return r < 0 ? r + n : r;
}

}  // namespace afsi
