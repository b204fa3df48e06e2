// Explicit MAC predictor of the closed box (gfx950): MACFlow.step's
//   u* = zero_normal(u + dt (-N(u) + (mu/rho) L u + f/rho)),  u = zero_normal(u0)
// (mac/flow.py:59-64), with the centered conservative momentum flux N of
// mac/grid.py:108-134 and the no-slip velocity Laplacian L of
// mac/grid.py:96-105 (odd ghosts across the walls), and the transport
// screen's component speeds (mac/flow.py:142-146).
//
// One thread per face of one component; every term is formed in the
// reference's operation order (axes accumulated 0, 1, 2; no contraction), so
// the result is bit-identical to the FP64 oracle. Wall-normal faces are read
// as zero (zero_normal of the input) and written as zero.
#include "common.hpp"

namespace afsi {
namespace {

struct Box {
    int n[3];        // cells
    double h[3];
};

// Face array of component a: extent n[a] + 1 along a, n elsewhere.
__device__ inline int ext(const Box& b, int comp, int axis) { return b.n[axis] + (axis == comp ? 1 : 0); }

__device__ inline int64_t at(const Box& b, int comp, int i0, int i1, int i2) {
    return (int64_t(i0) * ext(b, comp, 1) + i1) * ext(b, comp, 2) + i2;
}

// Value of zero_normal(velocity)[comp] at face (i0, i1, i2).
__device__ inline double vel(const double* __restrict__ u, const Box& b, int comp, int i0, int i1, int i2) {
    const int i[3] = {i0, i1, i2};
    if (i[comp] == 0 || i[comp] == b.n[comp]) return 0.0;
    return u[at(b, comp, i0, i1, i2)];
}

template <int C>
__global__ void predict_kernel(const double* __restrict__ u0, const double* __restrict__ u1,
                               const double* __restrict__ u2, const double* __restrict__ f,
                               double* __restrict__ out, Box b, double dt, double rho, double mu) {
    const double* us[3] = {u0, u1, u2};
    const int e0 = ext(b, C, 0), e1 = ext(b, C, 1), e2 = ext(b, C, 2);
    const int64_t total = int64_t(e0) * e1 * e2;
    const double nu = mu / rho;
    for (int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < total;
         p += int64_t(gridDim.x) * blockDim.x) {
        const int i2 = int(p % e2);
        const int64_t r = p / e2;
        const int i1 = int(r % e1), i0 = int(r / e1);
        const int idx[3] = {i0, i1, i2};
        if (idx[C] == 0 || idx[C] == b.n[C]) {
            out[p] = 0.0;
            continue;
        }
        auto U = [&](int a0, int a1, int a2) { return vel(us[C], b, C, a0, a1, a2); };
        const double u = U(i0, i1, i2);
        double conv = 0.0, lap = 0.0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double h = b.h[a];
            // convection term of axis a
            if (a == C) {
                int m[3] = {i0, i1, i2}, pl[3] = {i0, i1, i2};
                m[C] -= 1;
                pl[C] += 1;
                const double sp = .5 * (U(pl[0], pl[1], pl[2]) + u);
                const double sm = .5 * (u + U(m[0], m[1], m[2]));
                conv += (sp * sp - sm * sm) / h;
            } else {
                // E(j) = advector(j) * carried(j) at the edges j = idx[a], idx[a] + 1
                const int ja = idx[a], na = b.n[a];
                double E[2];
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    const int j = ja + s;  // face index of component a along a
                    int q[3] = {i0, i1, i2};
                    q[a] = j;
                    int qm[3] = {q[0], q[1], q[2]};
                    qm[C] -= 1;
                    const double adv = .5 * (vel(us[a], b, a, q[0], q[1], q[2]) + vel(us[a], b, a, qm[0], qm[1], qm[2]));
                    // carried: padded u along a with odd ghosts, P(m) = -u[0] (m = 0),
                    // u[m - 1] (1 <= m <= na), -u[na - 1] (m = na + 1); carried(j) = .5 (P(j+1) + P(j))
                    auto P = [&](int mm) {
                        int w[3] = {i0, i1, i2};
                        if (mm == 0) { w[a] = 0; return -U(w[0], w[1], w[2]); }
                        if (mm == na + 1) { w[a] = na - 1; return -U(w[0], w[1], w[2]); }
                        w[a] = mm - 1;
                        return U(w[0], w[1], w[2]);
                    };
                    const double carried = .5 * (P(j + 1) + P(j));
                    E[s] = adv * carried;
                }
                conv += (E[1] - E[0]) / h;
            }
            // Laplacian term of axis a: ((left - 2u) + right) / h^2, odd ghosts
            const int last = ext(b, C, a) - 1;
            int l[3] = {i0, i1, i2}, rr[3] = {i0, i1, i2};
            l[a] -= 1;
            rr[a] += 1;
            const double left = idx[a] == 0 ? -u : U(l[0], l[1], l[2]);
            const double right = idx[a] == last ? -u : U(rr[0], rr[1], rr[2]);
            lap += (left - 2 * u + right) / (h * h);
        }
        out[p] = u + dt * ((-conv + nu * lap) + f[p] / rho);
    }
}

constexpr int SPEED_BLOCKS = 256;

// Per block: max |u_c| for c = 0..2 and the count of non-finite entries of
// u and f; one block then reduces the blocks in order.
__global__ void __launch_bounds__(256) speeds_partial_kernel(const double* __restrict__ u0,
                                                            const double* __restrict__ u1,
                                                            const double* __restrict__ u2,
                                                            const double* __restrict__ f0,
                                                            const double* __restrict__ f1,
                                                            const double* __restrict__ f2, Box b,
                                                            double* __restrict__ partials) {
    __shared__ double sh[4][256];
    const double* us[3] = {u0, u1, u2};
    const double* fs[3] = {f0, f1, f2};
    double m[3] = {0.0, 0.0, 0.0}, bad = 0.0;
    for (int c = 0; c < 3; ++c) {
        const int64_t total = int64_t(ext(b, c, 0)) * ext(b, c, 1) * ext(b, c, 2);
        for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < total; i += int64_t(SPEED_BLOCKS) * 256) {
            const double v = us[c][i], w = fs[c][i];
            if (!isfinite(v)) bad += 1.0;
            else m[c] = fmax(m[c], fabs(v));
            if (!isfinite(w)) bad += 1.0;
        }
    }
    for (int c = 0; c < 3; ++c) sh[c][threadIdx.x] = m[c];
    sh[3][threadIdx.x] = bad;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) {
            for (int c = 0; c < 3; ++c) sh[c][threadIdx.x] = fmax(sh[c][threadIdx.x], sh[c][threadIdx.x + s]);
            sh[3][threadIdx.x] += sh[3][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x < 4) partials[4 * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ void speeds_final_kernel(const double* __restrict__ partials, double* __restrict__ out) {
    if (threadIdx.x != 0) return;
    double m[3] = {0.0, 0.0, 0.0}, bad = 0.0;
    for (int k = 0; k < SPEED_BLOCKS; ++k) {
        for (int c = 0; c < 3; ++c) m[c] = fmax(m[c], partials[4 * k + c]);
        bad += partials[4 * k + 3];
    }
    out[0] = m[0];
    out[1] = m[1];
    out[2] = m[2];
    out[3] = bad;
}

int make_box(const int64_t n[3], const double h[3], Box& b) {
    AFSI_REQUIRE(n && h, "null grid");
    for (int a = 0; a < 3; ++a) {
        AFSI_REQUIRE(n[a] >= 2 && n[a] < (int64_t(1) << 20), "invalid grid extent");
        AFSI_REQUIRE(h[a] > 0 && std::isfinite(h[a]), "invalid grid spacing");
        b.n[a] = int(n[a]);
        b.h[a] = h[a];
    }
    AFSI_REQUIRE((n[0] + 1) * (n[1] + 1) * (n[2] + 1) < (int64_t(1) << 40), "grid too large");
    return OK;
}

}  // namespace
}  // namespace afsi

using namespace afsi;

extern "C" {

int afsi_mac_predict_f64(const double* u0, const double* u1, const double* u2, const double* f0,
                         const double* f1, const double* f2, double* s0, double* s1, double* s2,
                         const int64_t n[3], const double h[3], double dt, double rho, double mu, void* stream) {
    Box b;
    if (int r = make_box(n, h, b)) return r;
    AFSI_REQUIRE(u0 && u1 && u2 && f0 && f1 && f2 && s0 && s1 && s2, "null pointer");
    AFSI_REQUIRE(s0 != u0 && s1 != u1 && s2 != u2 && s0 != u1 && s0 != u2 && s1 != u0 && s1 != u2 &&
                 s2 != u0 && s2 != u1, "the predictor writes out of place");
    AFSI_REQUIRE(std::isfinite(dt) && std::isfinite(rho) && std::isfinite(mu) && dt > 0 && rho > 0 && mu > 0,
                 "this explicit-viscous MAC scheme requires positive dt/rho/mu");
    const hipStream_t s = hipStream_t(stream);
    const int64_t t0 = int64_t(b.n[0] + 1) * b.n[1] * b.n[2], t1 = int64_t(b.n[0]) * (b.n[1] + 1) * b.n[2],
                  t2 = int64_t(b.n[0]) * b.n[1] * (b.n[2] + 1);
    predict_kernel<0><<<grid_for(t0, 256, 16384), 256, 0, s>>>(u0, u1, u2, f0, s0, b, dt, rho, mu);
    AFSI_LAUNCH_CHECK();
    predict_kernel<1><<<grid_for(t1, 256, 16384), 256, 0, s>>>(u0, u1, u2, f1, s1, b, dt, rho, mu);
    AFSI_LAUNCH_CHECK();
    predict_kernel<2><<<grid_for(t2, 256, 16384), 256, 0, s>>>(u0, u1, u2, f2, s2, b, dt, rho, mu);
    AFSI_LAUNCH_CHECK();
    return OK;
}

size_t afsi_mac_speeds_workspace_bytes(void) { return size_t(SPEED_BLOCKS) * 4 * sizeof(double); }

int afsi_mac_speeds_f64(const double* u0, const double* u1, const double* u2, const double* f0,
                        const double* f1, const double* f2, const int64_t n[3], double* out,
                        void* workspace, size_t workspace_bytes, void* stream) {
    Box b;
    const double h[3] = {1.0, 1.0, 1.0};
    if (int r = make_box(n, h, b)) return r;
    AFSI_REQUIRE(u0 && u1 && u2 && f0 && f1 && f2 && out && workspace, "null pointer");
    AFSI_REQUIRE(workspace_bytes >= afsi_mac_speeds_workspace_bytes(), "speeds workspace too small");
    const hipStream_t s = hipStream_t(stream);
    double* partials = static_cast<double*>(workspace);
    speeds_partial_kernel<<<SPEED_BLOCKS, 256, 0, s>>>(u0, u1, u2, f0, f1, f2, b, partials);
    AFSI_LAUNCH_CHECK();
    speeds_final_kernel<<<1, 64, 0, s>>>(partials, out);
    AFSI_LAUNCH_CHECK();
    return OK;
}

}  // extern "C"
