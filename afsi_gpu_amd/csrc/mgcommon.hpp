// Shared device helpers of the 3D multigrid kernels.
#pragma once
#include "common.hpp"

namespace afsi {

struct Dims {
    int64_t n0, n1, n2;
};

template <class T>
struct Coef {
    T h2[3];
    T w[3];  // 1/h^2
};

template <class T, bool P>
__device__ inline T apply_A(const T* __restrict__ p, int64_t i, int64_t x, int64_t y, int64_t z,
                            const Dims& d, const Coef<T>& k, T& pc) {
    const int64_t sx = d.n1 * d.n2, sy = d.n2;
    const T c = p[i];
    pc = c;
    T a = T(0);
    if (P) {
        const T xp = p[i + (x + 1 < d.n0 ? sx : -(d.n0 - 1) * sx)];
        const T xm = p[i + (x > 0 ? -sx : (d.n0 - 1) * sx)];
        a -= (xp - c) / k.h2[0];
        a += (c - xm) / k.h2[0];
        const T yp = p[i + (y + 1 < d.n1 ? sy : -(d.n1 - 1) * sy)];
        const T ym = p[i + (y > 0 ? -sy : (d.n1 - 1) * sy)];
        a -= (yp - c) / k.h2[1];
        a += (c - ym) / k.h2[1];
        const T zp = p[i + (z + 1 < d.n2 ? 1 : -(d.n2 - 1))];
        const T zm = p[i + (z > 0 ? -1 : (d.n2 - 1))];
        a -= (zp - c) / k.h2[2];
        a += (c - zm) / k.h2[2];
    } else {
        if (x + 1 < d.n0) a -= (p[i + sx] - c) / k.h2[0];
        if (x > 0) a += (c - p[i - sx]) / k.h2[0];
        if (y + 1 < d.n1) a -= (p[i + sy] - c) / k.h2[1];
        if (y > 0) a += (c - p[i - sy]) / k.h2[1];
        if (z + 1 < d.n2) a -= (p[i + 1] - c) / k.h2[2];
        if (z > 0) a += (c - p[i - 1]) / k.h2[2];
    }
    return a;
}

template <class T, bool P>
__device__ inline T diag_of(int64_t x, int64_t y, int64_t z, const Dims& d, const Coef<T>& k) {
    T s = T(0);
    if (P || x + 1 < d.n0) s += k.w[0];
    if (P || x > 0) s += k.w[0];
    if (P || y + 1 < d.n1) s += k.w[1];
    if (P || y > 0) s += k.w[1];
    if (P || z + 1 < d.n2) s += k.w[2];
    if (P || z > 0) s += k.w[2];
    return s;
}

}  // namespace afsi
