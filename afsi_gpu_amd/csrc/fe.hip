// FE interaction-point transfer and consistent-mass Jacobi PCG (gfx950).
//
// The reference's quadrature IB equations (mac/transfer.py:1-5):
//   M F = b,  f = V^-1 K^T W B F,  M U = B^T W K u,
// with B the FE evaluation at interaction points (N_qa), W the physical
// quadrature weights and K the Peskin-4 MAC stencil (the tiled transfer
// kernels). This file holds B, W B, B^T W, the CSR mass action and the PCG of
// fluid/solvers.py:39-122 (recurrence, check/restart schedule and
// true-residual acceptance), driven by a host loop that reads the device
// state back only at the schedule's check iterations.
//
// B^T W is a gather per node over a node-incidence CSR (deterministic order),
// not an atomic scatter. PCG reductions are two-stage with a fixed number of
// partial blocks, so every solve is bit-reproducible.
#include "common.hpp"

namespace afsi {
namespace {

constexpr int FE_MAX_NODES = 32;  // nodes per cell
constexpr int FE_MAX_QUAD = 64;   // quadrature points per cell
constexpr int RED_BLOCKS = 256;   // partial sums of the vector reductions
constexpr int RED_THREADS = 256;
constexpr int SPMV_LANES = 8;     // lanes per CSR row
constexpr int SPMV_BLOCKS = 2048; // partial sums of the product's d . Ad

// B (scale = false) or W B (scale = true): one thread per interaction point.
template <bool SCALE>
__global__ void fe_evaluate_kernel(afsi_fe_geometry g, const double* __restrict__ nodal,
                                   double* __restrict__ out) {
    const int64_t total = g.elements * g.quad;
    for (int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < total;
         p += int64_t(gridDim.x) * blockDim.x) {
        const int64_t e = p / g.quad;
        const int q = int(p - e * g.quad);
        const int64_t* cell = g.cells + e * g.nodes_per_cell;
        const double* N = g.values + int64_t(q) * g.nodes_per_cell;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        for (int a = 0; a < g.nodes_per_cell; ++a) {
            const double* v = nodal + 3 * cell[a];
            s0 += N[a] * v[0];
            s1 += N[a] * v[1];
            s2 += N[a] * v[2];
        }
        if (SCALE) {
            const double w = g.weights[p];
            s0 *= w; s1 *= w; s2 *= w;
        }
        out[3 * p] = s0;
        out[3 * p + 1] = s1;
        out[3 * p + 2] = s2;
    }
}

// B^T W: node n sums, over its incidences (e, a) in node_items order,
// sum_q N_qa W_eq U_(e,q). One wave per node: lane l takes quadrature point
// q = l % quad of the incidences slot, slot + 64/quad, ... (slot = l / quad),
// accumulating in that fixed order; the 64 lane sums are then combined by a
// fixed xor butterfly, so the result is reproducible.
__global__ void __launch_bounds__(256) fe_assemble_kernel(afsi_fe_geometry g, const double* __restrict__ point_values,
                                                          double* __restrict__ out) {
    const int lane = threadIdx.x % WAVE;
    const int per = WAVE / g.quad;  // incidences per pass (quad <= 64)
    const int q = lane % g.quad, slot = lane / g.quad;
    const bool active = slot < per;
    const int64_t waves = int64_t(gridDim.x) * (blockDim.x / WAVE);
    for (int64_t n = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / WAVE; n < g.node_count; n += waves) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        if (active) {
            for (int32_t k = g.node_ptr[n] + slot; k < g.node_ptr[n + 1]; k += per) {
                const int64_t item = g.node_items[k];
                const int64_t e = item / g.nodes_per_cell;
                const int a = int(item - e * g.nodes_per_cell);
                const double c = g.values[int64_t(q) * g.nodes_per_cell + a] * g.weights[e * g.quad + q];
                const double* U = point_values + 3 * (e * g.quad + q);
                s0 += c * U[0];
                s1 += c * U[1];
                s2 += c * U[2];
            }
        }
#pragma unroll
        for (int m = WAVE / 2; m > 0; m >>= 1) {
            s0 += __shfl_xor(s0, m);
            s1 += __shfl_xor(s1, m);
            s2 += __shfl_xor(s2, m);
        }
        if (lane == 0) {
            out[3 * n] = s0;
            out[3 * n + 1] = s1;
            out[3 * n + 2] = s2;
        }
    }
}

// Row r of A x (three columns) on SPMV_LANES consecutive lanes: lane j
// takes the entries row_ptr[r] + j, + j + SPMV_LANES, ...; the lane sums
// are combined by a fixed xor butterfly, so every lane of the group holds
// the same, reproducible total.
__device__ inline void csr_row3(const afsi_csr& a, const double* __restrict__ x, int64_t r, int sub,
                                double& s0, double& s1, double& s2) {
    s0 = s1 = s2 = 0.0;
    const int32_t end = a.row_ptr[r + 1];
    for (int32_t k = a.row_ptr[r] + sub; k < end; k += SPMV_LANES) {
        const double v = a.vals[k];
        const double* xc = x + 3 * int64_t(a.cols[k]);
        s0 += v * xc[0];
        s1 += v * xc[1];
        s2 += v * xc[2];
    }
#pragma unroll
    for (int m = SPMV_LANES / 2; m > 0; m >>= 1) {
        s0 += __shfl_xor(s0, m);
        s1 += __shfl_xor(s1, m);
        s2 += __shfl_xor(s2, m);
    }
}

// y = A x, three columns.
__global__ void __launch_bounds__(256) csr_spmv3_kernel(afsi_csr a, const double* __restrict__ x,
                                                        double* __restrict__ y) {
    const int sub = threadIdx.x % SPMV_LANES;
    const int64_t groups = int64_t(gridDim.x) * blockDim.x / SPMV_LANES;
    for (int64_t r = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / SPMV_LANES; r < a.rows; r += groups) {
        double s0, s1, s2;
        csr_row3(a, x, r, sub, s0, s1, s2);
        if (sub == 0) {
            y[3 * r] = s0;
            y[3 * r + 1] = s1;
            y[3 * r + 2] = s2;
        }
    }
}

// ---------------------------------------------------------------------------
// PCG. Device state (doubles): the scalars of the recurrence and its flags.
enum : int { S_RZ = 0, S_CURV, S_ALPHA, S_RN, S_TOL, S_ACTIVE, S_BROKEN, S_WORKING, S_TRUE, S_RHS, S_RZOLD, S_COUNT };

__device__ inline double block_sum(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = RED_THREADS / 2; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    return sh[0];
}

// Partial sums over blocks of sum_i a_i b_i (mode 0), sum_i a_i a_i / d_i
// (mode 1: r . z with z = r / diag), sum_i a_i^2 (mode 2).
__global__ void __launch_bounds__(RED_THREADS) pcg_dot_kernel(int mode, const double* __restrict__ a,
                                                              const double* __restrict__ b,
                                                              const double* __restrict__ diag, int64_t n,
                                                              double* __restrict__ partials) {
    __shared__ double sh[RED_THREADS];
    double s = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * RED_THREADS + threadIdx.x; i < n; i += int64_t(RED_BLOCKS) * RED_THREADS)
        s += mode == 0 ? a[i] * b[i] : (mode == 1 ? a[i] * (a[i] / diag[i]) : a[i] * a[i]);
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// Ad = A d (three columns) and SPMV_BLOCKS partial sums of d . Ad, row
// groups strided over the fixed grid.
__global__ void __launch_bounds__(RED_THREADS) pcg_spmv_dot_kernel(afsi_csr a, const double* __restrict__ d,
                                                                   double* __restrict__ Ad,
                                                                   double* __restrict__ partials) {
    __shared__ double sh[RED_THREADS];
    const int sub = threadIdx.x % SPMV_LANES;
    const int64_t groups = int64_t(SPMV_BLOCKS) * RED_THREADS / SPMV_LANES;
    double s = 0.0;
    for (int64_t r = (int64_t(blockIdx.x) * RED_THREADS + threadIdx.x) / SPMV_LANES; r < a.rows; r += groups) {
        double s0, s1, s2;
        csr_row3(a, d, r, sub, s0, s1, s2);
        if (sub == 0) {
            Ad[3 * r] = s0;
            Ad[3 * r + 1] = s1;
            Ad[3 * r + 2] = s2;
            s += (d[3 * r] * s0 + d[3 * r + 1] * s1) + d[3 * r + 2] * s2;
        }
    }
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// Sum of `count` partials in the fixed order of pcg_scalar_kernel (thread k
// adds entries k, k + RED_THREADS, ..., then the block tree): every block
// that calls it gets the same bits.
__device__ inline double sum_partials(const double* __restrict__ p, int count, double* sh) {
    double own = 0.0;
    for (int k = threadIdx.x; k < count; k += RED_THREADS) own += p[k];
    return block_sum(own, sh);
}

// One PCG iteration's update with its scalar step folded in: every block
// forms curvature = d . Ad from the product's SPMV_BLOCKS partials and
// alpha exactly as the curvature step does; block 0 records the scalars
// (S_BROKEN is re-derived identically by any block that reads the updated
// value, and S_RZ is kept in S_RZOLD for the direction step). Then
// x += alpha d, r -= alpha Ad with partial sums of r^2 (out[0, B)) and of
// r . (r / diag) (out[B, 2B)).
__global__ void __launch_bounds__(RED_THREADS) pcg_update_kernel(double* __restrict__ x, double* __restrict__ r,
                                                                 const double* __restrict__ d,
                                                                 const double* __restrict__ Ad,
                                                                 const double* __restrict__ diag, int64_t n,
                                                                 double* __restrict__ st,
                                                                 const double* __restrict__ curv_partials,
                                                                 double* __restrict__ out) {
    __shared__ double sh[RED_THREADS];
    const double curvature = sum_partials(curv_partials, SPMV_BLOCKS, sh);
    const double rz = st[S_RZ];
    const bool active = st[S_ACTIVE] != 0.0;
    const bool valid = isfinite(curvature) && curvature > 0.0 && isfinite(rz) && rz > 0.0;
    const bool broken = st[S_BROKEN] != 0.0 || (active && !valid);
    const bool working = active && valid && !broken;
    const double alpha = (working ? rz : 0.0) / (working ? curvature : 1.0);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st[S_CURV] = curvature;
        st[S_ALPHA] = alpha;
        st[S_BROKEN] = broken ? 1.0 : 0.0;
        st[S_WORKING] = working ? 1.0 : 0.0;
        st[S_RZOLD] = rz;
    }
    __syncthreads();
    double s = 0.0, q = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * RED_THREADS + threadIdx.x; i < n; i += int64_t(RED_BLOCKS) * RED_THREADS) {
        x[i] = x[i] + alpha * d[i];
        const double ri = r[i] - alpha * Ad[i];
        r[i] = ri;
        s += ri * ri;
        q += ri * (ri / diag[i]);
    }
    const double t = block_sum(s, sh);
    __syncthreads();
    const double u = block_sum(q, sh);
    if (threadIdx.x == 0) {
        out[blockIdx.x] = t;
        out[RED_BLOCKS + blockIdx.x] = u;
    }
}

// The norm/active/beta step folded into the direction: every block forms
// rn, r.z, active and beta = r.z / rz_old as the scalar step does (rz_old
// from S_RZOLD, which this launch does not write); block 0 records S_RN,
// S_BROKEN, S_ACTIVE, beta and the new rz. d = active ? z + beta d : 0.
__global__ void __launch_bounds__(RED_THREADS) pcg_next_direction_kernel(double* __restrict__ d,
                                                                         const double* __restrict__ r,
                                                                         const double* __restrict__ diag, int64_t n,
                                                                         double* __restrict__ st,
                                                                         const double* __restrict__ sums) {
    __shared__ double sh[RED_THREADS];
    const double t = sum_partials(sums, RED_BLOCKS, sh);
    __syncthreads();
    const double t2 = sum_partials(sums + RED_BLOCKS, RED_BLOCKS, sh);
    const double rn = sqrt(t);
    const bool broken = st[S_BROKEN] != 0.0 || !isfinite(rn);
    const bool active = st[S_WORKING] != 0.0 && rn > st[S_TOL];
    const double beta = t2 / (active ? st[S_RZOLD] : 1.0);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st[S_RN] = rn;
        st[S_BROKEN] = broken ? 1.0 : 0.0;
        st[S_ACTIVE] = active ? 1.0 : 0.0;
        st[S_CURV] = beta;
        st[S_RZ] = t2;
    }
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        const double z = r[i] / diag[i];
        d[i] = active ? z + beta * d[i] : 0.0;
    }
}

// r = rhs - Ax; partial sums of r^2.
__global__ void __launch_bounds__(RED_THREADS) pcg_residual_kernel(double* __restrict__ r,
                                                                   const double* __restrict__ rhs,
                                                                   const double* __restrict__ Ax, int64_t n,
                                                                   double* __restrict__ partials) {
    __shared__ double sh[RED_THREADS];
    double s = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * RED_THREADS + threadIdx.x; i < n; i += int64_t(RED_BLOCKS) * RED_THREADS) {
        const double ri = rhs[i] - Ax[i];
        r[i] = ri;
        s += ri * ri;
    }
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// The scalar steps outside the iteration, one block summing the RED_BLOCKS
// partials in order: what 2 restart r.z (rz, active = 1, beta = 0); 3 true
// residual norm; 4 initial r.z (rz, active = 1, not broken); 6 rhs norm ->
// tolerance max(atol, rtol ||rhs||).
__global__ void __launch_bounds__(RED_THREADS) pcg_scalar_kernel(int what, const double* __restrict__ partials,
                                                                 double* __restrict__ st, double atol = 0.0,
                                                                 double rtol = 0.0) {
    __shared__ double sh[RED_THREADS];
    const double t = sum_partials(partials, RED_BLOCKS, sh);
    if (threadIdx.x != 0) return;
    if (what == 6) {
        const double rhs_norm = sqrt(t);
        st[S_RHS] = rhs_norm;
        st[S_TOL] = fmax(atol, rtol * rhs_norm);
    } else if (what == 2) {
        st[S_ACTIVE] = 1.0;
        st[S_CURV] = 0.0;
        st[S_RZ] = t;
    } else if (what == 3) {
        st[S_TRUE] = sqrt(t);
    } else {
        st[S_RZ] = t;
        st[S_ACTIVE] = 1.0;
        st[S_BROKEN] = 0.0;
    }
}

// d = z (restart) or active ? z + beta d : 0, with z = r / diag.
__global__ void pcg_direction_kernel(double* __restrict__ d, const double* __restrict__ r,
                                     const double* __restrict__ diag, int64_t n, int restart,
                                     const double* __restrict__ st) {
    const double beta = st[S_CURV];
    const bool active = st[S_ACTIVE] != 0.0;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        const double z = r[i] / diag[i];
        d[i] = restart ? z : (active ? z + beta * d[i] : 0.0);
    }
}

struct PcgSpace {
    double *r, *d, *Ad, *partials, *state;
};

// partials: [0, SPMV_BLOCKS) the product's d . Ad (and the vector
// reductions' RED_BLOCKS sums), then 2 * RED_BLOCKS for the update's sums.
constexpr int PARTIALS = SPMV_BLOCKS + 2 * RED_BLOCKS;

size_t pcg_bytes(int64_t rows) {
    const size_t v = (size_t(rows) * 3 * sizeof(double) + 255) & ~size_t(255);
    return 3 * v + PARTIALS * sizeof(double) + 256;
}

PcgSpace pcg_space(void* ws, int64_t rows) {
    char* p = static_cast<char*>(ws);
    const size_t v = (size_t(rows) * 3 * sizeof(double) + 255) & ~size_t(255);
    PcgSpace s;
    s.r = reinterpret_cast<double*>(p);
    s.d = reinterpret_cast<double*>(p + v);
    s.Ad = reinterpret_cast<double*>(p + 2 * v);
    s.partials = reinterpret_cast<double*>(p + 3 * v);
    s.state = reinterpret_cast<double*>(p + 3 * v + PARTIALS * sizeof(double));
    return s;
}

// Pinned host slots for look-ahead state readbacks, one set per host thread
// and device (created on first use, kept for the life of the process).
struct Readback {
    int device = -1;
    double* host = nullptr;  // 2 slots of S_COUNT doubles
    hipEvent_t done[2] = {nullptr, nullptr};
};
thread_local Readback readback;

int ensure_readback() {
    int device = 0;
    AFSI_HIP(hipGetDevice(&device));
    if (readback.host && readback.device == device) return OK;
    if (!readback.host) AFSI_HIP(hipHostMalloc(reinterpret_cast<void**>(&readback.host), 2 * S_COUNT * sizeof(double),
                                               hipHostMallocDefault));
    for (int k = 0; k < 2; ++k) {
        if (readback.done[k]) AFSI_HIP(hipEventDestroy(readback.done[k]));
        AFSI_HIP(hipEventCreateWithFlags(&readback.done[k], hipEventDisableTiming));
    }
    readback.device = device;
    return OK;
}

int check_geometry(const afsi_fe_geometry* g) {
    AFSI_REQUIRE(g && g->cells && g->values && g->weights, "null FE geometry");
    AFSI_REQUIRE(g->elements >= 0 && g->node_count > 0, "invalid FE geometry sizes");
    AFSI_REQUIRE(g->nodes_per_cell >= 1 && g->nodes_per_cell <= FE_MAX_NODES, "nodes per cell must be in [1, %d]",
                 FE_MAX_NODES);
    AFSI_REQUIRE(g->quad >= 1 && g->quad <= FE_MAX_QUAD, "quadrature points per cell must be in [1, %d]", FE_MAX_QUAD);
    AFSI_REQUIRE(g->elements * g->quad < (int64_t(1) << 40), "too many interaction points");
    return OK;
}

}  // namespace
}  // namespace afsi

using namespace afsi;

extern "C" {

int afsi_fe_evaluate_f64(const afsi_fe_geometry* g, const double* nodal, double* out, void* stream) {
    if (int r = check_geometry(g)) return r;
    AFSI_REQUIRE(nodal && out, "null pointer");
    if (g->elements == 0) return OK;
    fe_evaluate_kernel<false><<<grid_for(g->elements * g->quad, 256, 8192), 256, 0, hipStream_t(stream)>>>(*g, nodal, out);
    AFSI_LAUNCH_CHECK();
    return OK;
}

int afsi_fe_weighted_f64(const afsi_fe_geometry* g, const double* nodal, double* out, void* stream) {
    if (int r = check_geometry(g)) return r;
    AFSI_REQUIRE(nodal && out, "null pointer");
    if (g->elements == 0) return OK;
    fe_evaluate_kernel<true><<<grid_for(g->elements * g->quad, 256, 8192), 256, 0, hipStream_t(stream)>>>(*g, nodal, out);
    AFSI_LAUNCH_CHECK();
    return OK;
}

int afsi_fe_assemble_f64(const afsi_fe_geometry* g, const double* point_values, double* out, void* stream) {
    if (int r = check_geometry(g)) return r;
    AFSI_REQUIRE(g->node_ptr && g->node_items && point_values && out, "null pointer");
    fe_assemble_kernel<<<grid_for(g->node_count * WAVE, 256, 16384), 256, 0, hipStream_t(stream)>>>(*g, point_values,
                                                                                                  out);
    AFSI_LAUNCH_CHECK();
    return OK;
}

int afsi_csr_spmv3_f64(const afsi_csr* a, const double* x, double* y, void* stream) {
    AFSI_REQUIRE(a && a->row_ptr && a->cols && a->vals && x && y && a->rows > 0, "invalid CSR product arguments");
    csr_spmv3_kernel<<<grid_for(a->rows * SPMV_LANES, 256, 8192), 256, 0, hipStream_t(stream)>>>(*a, x, y);
    AFSI_LAUNCH_CHECK();
    return OK;
}

size_t afsi_mass_pcg_workspace_bytes(int64_t rows) { return rows > 0 ? pcg_bytes(rows) : 0; }

int afsi_mass_pcg_f64(const afsi_csr* a, const double* diagonal, const double* rhs, const double* initial,
                      double* x, const afsi_pcg_options* opt, void* workspace, size_t workspace_bytes,
                      afsi_pcg_info* info, void* stream) {
    AFSI_REQUIRE(a && a->row_ptr && a->cols && a->vals && a->rows > 0, "invalid mass matrix");
    AFSI_REQUIRE(diagonal && rhs && x && opt && info && workspace, "null pointer");
    AFSI_REQUIRE(workspace_bytes >= pcg_bytes(a->rows), "PCG workspace too small");
    AFSI_REQUIRE(opt->max_iterations >= 1 && opt->recompute_every >= 1 && opt->check_every >= 1,
                 "PCG iteration counts must be positive");
    AFSI_REQUIRE(std::isfinite(opt->rtol) && std::isfinite(opt->atol) && opt->rtol >= 0 && opt->atol >= 0 &&
                 opt->rtol + opt->atol > 0, "invalid PCG tolerances");
    const hipStream_t s = hipStream_t(stream);
    const int64_t n = a->rows * 3;
    const PcgSpace w = pcg_space(workspace, a->rows);
    const unsigned gv = grid_for(a->rows * SPMV_LANES, 256, 8192), ge = grid_for(n, 256, 8192);
    double host[S_COUNT];
    auto read_state = [&]() -> int {
        AFSI_HIP(hipMemcpyAsync(host, w.state, sizeof(host), hipMemcpyDeviceToHost, s));
        AFSI_HIP(hipStreamSynchronize(s));
        return OK;
    };
    auto scalar = [&](int what) -> int {
        pcg_scalar_kernel<<<1, RED_THREADS, 0, s>>>(what, w.partials, w.state);
        AFSI_LAUNCH_CHECK();
        return OK;
    };
    // true residual r = rhs - A x and its norm (state S_TRUE)
    auto true_residual = [&]() -> int {
        csr_spmv3_kernel<<<gv, 256, 0, s>>>(*a, x, w.Ad);
        AFSI_LAUNCH_CHECK();
        pcg_residual_kernel<<<RED_BLOCKS, RED_THREADS, 0, s>>>(w.r, rhs, w.Ad, n, w.partials);
        AFSI_LAUNCH_CHECK();
        return scalar(3);
    };
    // rhs norm and tolerance (device), initial iterate, true residual: one
    // readback for both norms.
    pcg_dot_kernel<<<RED_BLOCKS, RED_THREADS, 0, s>>>(2, rhs, rhs, diagonal, n, w.partials);
    AFSI_LAUNCH_CHECK();
    pcg_scalar_kernel<<<1, RED_THREADS, 0, s>>>(6, w.partials, w.state, opt->atol, opt->rtol);
    AFSI_LAUNCH_CHECK();
    if (initial) {
        if (initial != x) AFSI_HIP(hipMemcpyAsync(x, initial, size_t(n) * sizeof(double), hipMemcpyDeviceToDevice, s));
    } else {
        AFSI_HIP(hipMemsetAsync(x, 0, size_t(n) * sizeof(double), s));
    }
    if (int r = true_residual()) return r;
    if (int r = read_state()) return r;
    const double rhs_norm = host[S_RHS];
    const double tol = host[S_TOL];
    info->rhs_norm = rhs_norm;
    info->tolerance = tol;
    info->iterations = 0;
    double true_norm = host[S_TRUE];
    if (true_norm <= tol) {
        info->status = 0;
        info->residual_norm = true_norm;
        return OK;
    }
    // d = z = r / diag, rz = r . z, active, not broken
    pcg_dot_kernel<<<RED_BLOCKS, RED_THREADS, 0, s>>>(1, w.r, w.r, diagonal, n, w.partials);
    AFSI_LAUNCH_CHECK();
    if (int r = scalar(4)) return r;
    pcg_direction_kernel<<<ge, 256, 0, s>>>(w.d, w.r, diagonal, n, 1, w.state);
    AFSI_LAUNCH_CHECK();
    // Per iteration: A d with d . Ad; alpha folded into the update with |r|^2
    // and r . z; the norm/active/beta step folded into the direction (three
    // launches). A restart replaces that direction by z afterwards.
    //
    // Check iterations read the device state one check period late: the copy
    // goes to a pinned slot and the host looks at it when it reaches the next
    // check, with the iterations in between already queued. A recurrence that
    // has converged (rn <= tol) or broken down is frozen on the device
    // (active = 0 gives alpha = 0 and d = 0), so those queued iterations leave
    // x and r unchanged and the host resumes at the check that decided.
    // Restart iterations and the last iteration are read synchronously.
    if (int r = ensure_readback()) return r;
    int pending = 0, pslot = 0, slot = 0;
    auto broken_at = [&](int at) {
        info->status = 1;
        info->iterations = at;
        info->residual_norm = host[S_RN];
        return OK;
    };
    for (int it = 1; it <= opt->max_iterations; ++it) {
        pcg_spmv_dot_kernel<<<SPMV_BLOCKS, RED_THREADS, 0, s>>>(*a, w.d, w.Ad, w.partials);
        AFSI_LAUNCH_CHECK();
        double* sums = w.partials + SPMV_BLOCKS;
        pcg_update_kernel<<<RED_BLOCKS, RED_THREADS, 0, s>>>(x, w.r, w.d, w.Ad, diagonal, n, w.state, w.partials,
                                                             sums);
        AFSI_LAUNCH_CHECK();
        pcg_next_direction_kernel<<<ge, RED_THREADS, 0, s>>>(w.d, w.r, diagonal, n, w.state, sums);
        AFSI_LAUNCH_CHECK();
        bool restart = it % opt->recompute_every == 0;
        const bool check = restart || it % opt->check_every == 0 || it == opt->max_iterations;
        if (check) {
            bool converged = false;
            int at = it;
            if (pending) {  // the previous check decides first
                AFSI_HIP(hipEventSynchronize(readback.done[pslot]));
                for (int k = 0; k < S_COUNT; ++k) host[k] = readback.host[pslot * S_COUNT + k];
                if (host[S_BROKEN] != 0.0) return broken_at(pending);
                if (host[S_RN] <= tol) {
                    converged = true;
                    at = pending;
                }
                pending = 0;
            }
            if (!converged) {
                if (restart || it == opt->max_iterations) {
                    if (int r = read_state()) return r;
                    if (host[S_BROKEN] != 0.0) return broken_at(it);
                    converged = host[S_RN] <= tol;
                } else {
                    AFSI_HIP(hipMemcpyAsync(readback.host + slot * S_COUNT, w.state, S_COUNT * sizeof(double),
                                            hipMemcpyDeviceToHost, s));
                    AFSI_HIP(hipEventRecord(readback.done[slot], s));
                    pending = it;
                    pslot = slot;
                    slot ^= 1;
                }
            }
            if (converged || restart || it == opt->max_iterations) {
                // true residual of the iterate of iteration `at` (frozen since)
                if (int r = true_residual()) return r;
                if (int r = read_state()) return r;
                true_norm = host[S_TRUE];
                info->iterations = at;
                info->residual_norm = true_norm;
                if (!std::isfinite(true_norm)) {
                    info->status = 3;
                    return OK;
                }
                if (true_norm <= tol) {
                    info->status = 0;
                    return OK;
                }
                it = at;  // queued iterations after `at` were no-ops
                pending = 0;
                restart = true;
            }
        }
        if (restart) {  // r.z of the (true) residual, active = 1, d = z
            pcg_dot_kernel<<<RED_BLOCKS, RED_THREADS, 0, s>>>(1, w.r, w.r, diagonal, n, w.partials);
            AFSI_LAUNCH_CHECK();
            if (int r = scalar(2)) return r;
            pcg_direction_kernel<<<ge, 256, 0, s>>>(w.d, w.r, diagonal, n, 1, w.state);
            AFSI_LAUNCH_CHECK();
        }
    }
    info->status = 2;
    info->iterations = opt->max_iterations;
    info->residual_norm = true_norm;
    return OK;
}

}  // extern "C"
