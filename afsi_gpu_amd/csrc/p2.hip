// P2 fused cell-block transfer (gfx950): the three operators of
// mac/csrc/p2_transfer.cpp:67-69 (kernels p2_transfer.cu:12-188).
//
// prepare: interaction points x_q = sum_a N_qa x[cells[e,a]] evaluated in
//   registers and turned directly into the shared face/centre Peskin tables
//   base (2,total,3) int64 and phi (2,total,3,4) of mac/shared_stencil.py:8-12
//   (lattice 0 = faces, lattice 1 = centres; component c reads lattice 0 along
//   c and lattice 1 elsewhere), with the support flag of
//   mac/compact_transfer.py:94-96 (2 <= (x-origin)/h < n-2, finite).
// spread: f_c += (phi_x phi_y phi_z) (W_eq sum_a N_qa F[cells[e,a]]_c) / V
//   without materialising interaction points or point forces.
// gather-assemble: rhs[cells[e,a]]_c += sum_q (N_qa W_eq) U_c(x_q).
//
// One workgroup (4 waves) per element. The element's stencil bounding box,
// per component, is an LDS patch: the spread accumulates into it with LDS
// FP64 atomics and flushes each non-zero entry once with a hardware global
// FP64 atomic; the gather stages it from HBM once and reads it from LDS.
// Table chunks of 64 points are staged in LDS with coalesced loads.
//   spread: one wave per point, lane = stencil tap (i,j,k) = (l>>4, l>>2&3, l&3):
//     the 64 lanes of one LDS atomic hit 64 distinct patch entries, so dense
//     points never collide inside an instruction.
//   gather: four lanes per point, lane part p takes the taps with i = p, and
//     the four partial sums are added in a fixed order.
// Elements whose patch exceeds P2_PATCH values fall back to direct global
// atomics (spread) or direct HBM reads (gather); both stay exact.
#include "common.hpp"

namespace afsi {
namespace {

constexpr int P2_NODES = 10;
constexpr int P2_THREADS = 256;
constexpr int P2_WAVES = P2_THREADS / WAVE;
constexpr int P2_CHUNK = 64;     // points staged per pass
constexpr int P2_PATCH = 1536;   // LDS patch capacity (values, three components)
constexpr int P2_BLOCK = 4;      // consecutive elements per spread workgroup
constexpr int P2_SPREAD_PATCH = 2304;
static_assert(P2_BLOCK * P2_NODES * 3 <= P2_THREADS, "one pass loads the block's nodal values");

struct Lattice {
    double origin[3];  // faces along their own axis; centres start at origin + h/2
    double h[3];
    double limit[3];   // n - 2
    int64_t n[3];
};

// ---------------------------------------------------------------------------
// prepare: one thread per interaction point, all three axes.
// Computed in the table type T, as the reference does (its origin and
// spacing tensors share the positions' dtype).
template <class T>
__global__ void __launch_bounds__(256) p2_prepare_kernel(const T* __restrict__ x,
                                                         const int64_t* __restrict__ cells,
                                                         const T* __restrict__ N, int64_t ne, int nq,
                                                         int64_t nodes, Lattice L, int64_t* __restrict__ base,
                                                         T* __restrict__ phi, int64_t total, int64_t offset,
                                                         int* __restrict__ invalid) {
    const int64_t count = ne * nq;
    for (int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < count;
         p += int64_t(gridDim.x) * blockDim.x) {
        const int64_t e = p / nq;
        const int q = int(p - e * nq);
        int64_t node[P2_NODES];
        bool bad_node = false;
#pragma unroll
        for (int a = 0; a < P2_NODES; ++a) {
            node[a] = cells[e * P2_NODES + a];
            bad_node = bad_node || node[a] < 0 || node[a] >= nodes;
        }
        if (bad_node) {
            raise_flag(invalid, F_INDEX);
            continue;
        }
        T pos[3] = {T(0), T(0), T(0)};
#pragma unroll
        for (int a = 0; a < P2_NODES; ++a) {
            const T w = N[q * P2_NODES + a];
            pos[0] += w * x[node[a] * 3];
            pos[1] += w * x[node[a] * 3 + 1];
            pos[2] += w * x[node[a] * 3 + 2];
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const T h = T(L.h[a]), o = T(L.origin[a]);
            const T physical = (pos[a] - o) / h;
            const bool valid = isfinite(physical) && physical >= T(2) && physical < T(L.limit[a]);
            if (!valid) raise_flag(invalid, F_SUPPORT);
#pragma unroll
            for (int l = 0; l < 2; ++l) {
                const T start = l ? o + T(0.5) * h : o;
                const T scaled = valid ? (pos[a] - start) / h : T(0);
                const T lower = floor(scaled - T(1));
                const T f = scaled - lower - T(1);
                T w[4];
                peskin_weights(f, w);
                const int64_t index = (l * total + offset + p) * 3 + a;
                base[index] = int64_t(lower);
                phi[index * 4] = w[0];
                phi[index * 4 + 1] = w[1];
                phi[index * 4 + 2] = w[2];
                phi[index * 4 + 3] = w[3];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Element patch: per component c the box start[c][a] .. start + ext[c][a].
struct Patch {
    int start[3][3];
    int ext[3][3];
    int off[4];
    int status;  // 0, F_SUPPORT (box outside the face arrays) or F_INDEX (node id)
};

template <class T>
struct Stage {
    T phi[2][P2_CHUNK * 12];   // [lattice][point * 12 + axis * 4 + k]
    int base[2][P2_CHUNK * 3]; // [lattice][point * 3 + axis]
};

__device__ inline int wave_min(int v) {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) v = min(v, __shfl_xor(v, m));
    return v;
}
__device__ inline int wave_max(int v) {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) v = max(v, __shfl_xor(v, m));
    return v;
}

constexpr int BASE_LIMIT = 1 << 28;

// Bounding box of the element's bases per lattice and axis, turned into the
// three component boxes. status = 0 when every box lies inside its face
// array, every base is representable and every node id of the element is in
// range. Uniform across the workgroup on return.
// With ``staged`` (every point of the element already in ``st``, bases
// range-checked by stage_chunk into ``bad``) the box is taken from LDS.
template <class T>
__device__ void plan_patch(Patch& s, int (*red)[12], const int64_t* __restrict__ base, int64_t first, int nq,
                           int64_t total, const int64_t n[3], const int64_t* __restrict__ cell, int ids,
                           int64_t nodes, const Stage<T>* staged, bool ok) {
    const int tid = threadIdx.x, lane = tid % WAVE, wave = tid / WAVE;
    int lo[6], hi[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) { lo[k] = BASE_LIMIT; hi[k] = -BASE_LIMIT; }
    for (int q = tid; q < nq; q += P2_THREADS) {
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            int64_t v;
            if (staged) v = staged->base[k / 3][q * 3 + k % 3];
            else v = base[((k / 3) * total + first + q) * 3 + k % 3];
            ok = ok && v > -BASE_LIMIT && v < BASE_LIMIT;
            lo[k] = min(lo[k], int(v));
            hi[k] = max(hi[k], int(v));
        }
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        lo[k] = wave_min(lo[k]);
        hi[k] = wave_max(hi[k]);
    }
    int bad = ok ? 0 : F_SUPPORT;
    if (tid < ids && (cell[tid] < 0 || cell[tid] >= nodes)) bad |= F_INDEX;
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) bad |= __shfl_xor(bad, m);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            red[wave][k] = lo[k];
            red[wave][6 + k] = hi[k];
        }
        red[P2_WAVES][wave] = bad;
    }
    __syncthreads();
    if (tid == 0) {
        int l6[6], h6[6];
        int any_bad = 0;
        for (int k = 0; k < 6; ++k) {
            l6[k] = red[0][k];
            h6[k] = red[0][6 + k];
            for (int w = 1; w < P2_WAVES; ++w) {
                l6[k] = min(l6[k], red[w][k]);
                h6[k] = max(h6[k], red[w][6 + k]);
            }
        }
        for (int w = 0; w < P2_WAVES; ++w) any_bad |= red[P2_WAVES][w];
        bool valid = any_bad == 0;
        s.status = any_bad;
        s.off[0] = 0;
        for (int c = 0; c < 3; ++c) {
            int size = 1;
            for (int a = 0; a < 3; ++a) {
                const int k = 3 * int(a != c) + a;  // lattice (a != c), axis a
                s.start[c][a] = l6[k];
                s.ext[c][a] = h6[k] - l6[k] + 4;
                const int64_t extent = n[a] + int(a == c);
                valid = valid && l6[k] >= 0 && int64_t(l6[k]) + s.ext[c][a] <= extent;
                size = valid ? size * s.ext[c][a] : 0;
            }
            s.off[c + 1] = s.off[c] + size;
        }
        if (!valid && !s.status) s.status = F_SUPPORT;
    }
    __syncthreads();
}

__device__ inline int64_t face_index(int c, int x, int y, int z, const int64_t n[3]) {
    return (int64_t(x) * (n[1] + int(c == 1)) + y) * (n[2] + int(c == 2)) + z;
}

// The patch geometry in scalar registers (every field is workgroup-uniform),
// with float reciprocals of the extents for the entry decode.
struct PatchRegs {
    int off[4];
    int start[3][3];
    int ext[3][3];
    float inv[3][3];
};

__device__ inline PatchRegs patch_regs(const Patch& s) {
    PatchRegs r;
#pragma unroll
    for (int c = 0; c < 4; ++c) r.off[c] = __builtin_amdgcn_readfirstlane(s.off[c]);
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            r.start[c][a] = __builtin_amdgcn_readfirstlane(s.start[c][a]);
            r.ext[c][a] = __builtin_amdgcn_readfirstlane(s.ext[c][a]);
            r.inv[c][a] = 1.0f / float(r.ext[c][a]);
        }
    return r;
}

// i / d for 0 <= i < 2^20 and 1 <= d: the float quotient is within one of
// the exact one, then corrected.
__device__ inline int small_div(int i, int d, float inv) {
    int q = int(float(i) * inv);
    if (q * d > i) --q;
    else if ((q + 1) * d <= i) ++q;
    return q;
}

template <class T>
__device__ inline T pick(int c, T a, T b, T d) {
    return c == 0 ? a : (c == 1 ? b : d);
}

// Patch entry i (< P2_PATCH) -> component and grid coordinates (selects, no
// dynamic register indexing).
__device__ inline int patch_decode(const PatchRegs& s, int i, int& x, int& y, int& z) {
    const int c = i >= s.off[2] ? 2 : (i >= s.off[1] ? 1 : 0);
    const int p = i - pick(c, s.off[0], s.off[1], s.off[2]);
    const int e2 = pick(c, s.ext[0][2], s.ext[1][2], s.ext[2][2]);
    const int e1 = pick(c, s.ext[0][1], s.ext[1][1], s.ext[2][1]);
    const int r = small_div(p, e2, pick(c, s.inv[0][2], s.inv[1][2], s.inv[2][2]));
    z = p - r * e2 + pick(c, s.start[0][2], s.start[1][2], s.start[2][2]);
    const int xr = small_div(r, e1, pick(c, s.inv[0][1], s.inv[1][1], s.inv[2][1]));
    y = r - xr * e1 + pick(c, s.start[0][1], s.start[1][1], s.start[2][1]);
    x = xr + pick(c, s.start[0][0], s.start[1][0], s.start[2][0]);
    return c;
}

// Coalesced copy of one chunk's tables into LDS; returns false when a base
// of this thread's share is out of the representable range.
template <class T>
__device__ inline bool stage_chunk(Stage<T>& st, const int64_t* __restrict__ base, const T* __restrict__ phi,
                                   int64_t first, int q0, int cnt, int64_t total) {
    const int tid = threadIdx.x;
    bool ok = true;
#pragma unroll
    for (int l = 0; l < 2; ++l) {
        const int64_t row = l * total + first + q0;
        const T* src = phi + row * 12;
        for (int i = tid; i < cnt * 12; i += P2_THREADS) st.phi[l][i] = src[i];
        const int64_t* bsrc = base + row * 3;
        for (int i = tid; i < cnt * 3; i += P2_THREADS) {
            const int64_t v = bsrc[i];
            ok = ok && v > -BASE_LIMIT && v < BASE_LIMIT;
            st.base[l][i] = int(v);
        }
    }
    return ok;
}

// ---------------------------------------------------------------------------
// Any T: products and the LDS patch in FP64, faces accumulated in T.
template <class T>
__global__ void __launch_bounds__(P2_THREADS) p2_spread_kernel(
    const T* __restrict__ value, const int64_t* __restrict__ cells, const T* __restrict__ N,
    const T* __restrict__ W, int64_t ne, int nq, const int64_t* __restrict__ base,
    const T* __restrict__ phi, int64_t total, int64_t offset, T* __restrict__ u0,
    T* __restrict__ u1, T* __restrict__ u2, int64_t nodes, Lattice L, double volume,
    int* __restrict__ flags) {
    __shared__ Patch s;
    __shared__ int red[P2_WAVES + 1][12];
    __shared__ double nodal[P2_BLOCK][P2_NODES * 3];
    __shared__ double force[P2_CHUNK * 3];
    __shared__ Stage<T> st;
    __shared__ double patch[P2_SPREAD_PATCH];

    const int tid = threadIdx.x, lane = tid % WAVE, wave = tid / WAVE;
    // this workgroup's elements e0 .. e0 + count: consecutive table rows
    const int64_t e0 = int64_t(blockIdx.x) * P2_BLOCK;
    const int count = int(min(int64_t(P2_BLOCK), ne - e0));
    const int points = count * nq;
    const int64_t first = offset + e0 * nq;
    const bool single = points <= P2_CHUNK;  // one chunk: stage first, plan from LDS
    bool ok = true;
    if (single) {
        ok = stage_chunk(st, base, phi, first, 0, points, total);
        __syncthreads();
    }
    plan_patch(s, red, base, first, points, total, L.n, cells + e0 * P2_NODES, count * P2_NODES, nodes,
               single ? &st : nullptr, ok);
    if (s.status) {
        if (tid == 0) raise_flag(flags, s.status);
        return;
    }
    const bool cached = s.off[3] <= P2_SPREAD_PATCH;
    if (tid < count * P2_NODES * 3) {
        const int k = tid / (P2_NODES * 3), r = tid % (P2_NODES * 3);
        nodal[k][r] = double(value[cells[(e0 + k) * P2_NODES + r / 3] * 3 + r % 3]);
    }
    if (cached)
        for (int i = tid; i < s.off[3]; i += P2_THREADS) patch[i] = 0.0;
    T* const fields[3] = {u0, u1, u2};
    const int ti = lane >> 4, tj = (lane >> 2) & 3, tk = lane & 3;
    const PatchRegs R = patch_regs(s);
    const int uwave = __builtin_amdgcn_readfirstlane(wave);
    int tapoff[3];  // this lane's tap inside each component box
#pragma unroll
    for (int c = 0; c < 3; ++c) tapoff[c] = (ti * R.ext[c][1] + tj) * R.ext[c][2] + tk;

    for (int g0 = 0; g0 < points; g0 += P2_CHUNK) {
        const int cnt = min(P2_CHUNK, points - g0);
        __syncthreads();  // previous chunk consumed; nodal and patch zero visible
        if (!single) stage_chunk(st, base, phi, first, g0, cnt, total);
        if (tid < cnt * 3) {
            const int g = g0 + tid / 3, c = tid % 3;
            const int k = g / nq, q = g - k * nq;
            const T* Nq = N + int64_t(q) * P2_NODES;
            double f = 0.0;
#pragma unroll
            for (int a = 0; a < P2_NODES; ++a) f += double(Nq[a]) * nodal[k][a * 3 + c];
            force[tid] = (f * double(W[(e0 + k) * nq + q])) / volume;
        }
        __syncthreads();
        for (int q = uwave; q < cnt; q += P2_WAVES) {
            const T* P0 = st.phi[0] + q * 12;
            const T* P1 = st.phi[1] + q * 12;
            const double fx0 = P0[ti], fx1 = P1[ti];
            const double fy0 = P0[4 + tj], fy1 = P1[4 + tj];
            const double fz0 = P0[8 + tk], fz1 = P1[8 + tk];
            int b[2][3];  // the point's bases, wave-uniform
#pragma unroll
            for (int l = 0; l < 2; ++l)
#pragma unroll
                for (int a = 0; a < 3; ++a) b[l][a] = __builtin_amdgcn_readfirstlane(st.base[l][q * 3 + a]);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double w = c == 0 ? (fx0 * fy1) * fz1 : (c == 1 ? (fx1 * fy0) * fz1 : (fx1 * fy1) * fz0);
                const double v = w * force[q * 3 + c];
                const int bx = b[int(c != 0)][0], by = b[int(c != 1)][1], bz = b[int(c != 2)][2];
                if (cached) {
                    const int point = R.off[c] + ((bx - R.start[c][0]) * R.ext[c][1] + (by - R.start[c][1])) *
                                      R.ext[c][2] + (bz - R.start[c][2]);
                    atomicAdd(patch + point + tapoff[c], v);
                } else {
                    unsafeAtomicAdd(fields[c] + face_index(c, bx + ti, by + tj, bz + tk, L.n), T(v));
                }
            }
        }
    }
    if (!cached) return;
    __syncthreads();
    for (int i = tid; i < R.off[3]; i += P2_THREADS) {
        const double v = patch[i];
        if (v == 0.0) continue;
        int x, y, z;
        const int c = patch_decode(R, i, x, y, z);
        unsafeAtomicAdd(fields[c] + face_index(c, x, y, z, L.n), T(v));
    }
}

// ---------------------------------------------------------------------------
// Any T: the patch, products and sums in FP64, rhs accumulated in T.
template <class T>
__global__ void __launch_bounds__(P2_THREADS) p2_gather_assemble_kernel(
    const T* __restrict__ u0, const T* __restrict__ u1, const T* __restrict__ u2,
    const int64_t* __restrict__ cells, const T* __restrict__ N, const T* __restrict__ W, int nq,
    const int64_t* __restrict__ base, const T* __restrict__ phi, int64_t total, int64_t offset,
    T* __restrict__ rhs, int64_t nodes, Lattice L, int* __restrict__ flags) {
    __shared__ Patch s;
    __shared__ int red[P2_WAVES + 1][12];
    __shared__ double part[4][P2_CHUNK * 3];  // [i tap][point * 3 + c]
    __shared__ double sums[30][8];
    __shared__ double nw[P2_CHUNK * P2_NODES];  // N_qa W_eq of the chunk's points
    __shared__ Stage<T> st;
    __shared__ double patch[P2_PATCH];

    const int tid = threadIdx.x;
    const int64_t e = blockIdx.x, first = offset + e * nq;
    const bool single = nq <= P2_CHUNK;  // one chunk: stage first, plan from LDS
    bool ok = true;
    if (single) {
        ok = stage_chunk(st, base, phi, first, 0, nq, total);
        __syncthreads();
    }
    plan_patch(s, red, base, first, nq, total, L.n, cells + e * P2_NODES, P2_NODES, nodes, single ? &st : nullptr,
               ok);
    if (s.status) {
        if (tid == 0) raise_flag(flags, s.status);
        return;
    }
    const bool cached = s.off[3] <= P2_PATCH;
    const T* const fields[3] = {u0, u1, u2};
    const PatchRegs R = patch_regs(s);
    if (cached)
        for (int i = tid; i < R.off[3]; i += P2_THREADS) {
            int x, y, z;
            const int c = patch_decode(R, i, x, y, z);
            patch[i] = double(fields[c][face_index(c, x, y, z, L.n)]);
        }
    // thread -> (point in chunk, tap slice i)
    const int pq = tid % P2_CHUNK, ti = tid / P2_CHUNK;
    // thread -> (node a, component c, residue r) of the node assembly
    const int aa = tid / 24, ac = (tid / 8) % 3, ar = tid % 8;
    double acc = 0.0;

    for (int q0 = 0; q0 < nq; q0 += P2_CHUNK) {
        const int cnt = min(P2_CHUNK, nq - q0);
        __syncthreads();
        if (!single) stage_chunk(st, base, phi, first, q0, cnt, total);
        for (int i = tid; i < cnt * P2_NODES; i += P2_THREADS)
            nw[i] = double(N[int64_t(q0) * P2_NODES + i]) * double(W[e * nq + q0 + i / P2_NODES]);
        __syncthreads();
        if (pq < cnt) {
            const T* P0 = st.phi[0] + pq * 12;
            const T* P1 = st.phi[1] + pq * 12;
            const int* B0 = st.base[0] + pq * 3;
            const int* B1 = st.base[1] + pq * 3;
            const double fx0 = P0[ti], fx1 = P1[ti];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double wx = c == 0 ? fx0 : fx1;
                const T* Py = (c == 1 ? P0 : P1) + 4;
                const T* Pz = (c == 2 ? P0 : P1) + 8;
                const int x = (c == 0 ? B0[0] : B1[0]) + ti;
                const int y0 = c == 1 ? B0[1] : B1[1];
                const int z0 = c == 2 ? B0[2] : B1[2];
                double got = 0.0;
                if (cached) {
                    const int ez = R.ext[c][2];
                    const double* row = patch + R.off[c] +
                                        ((x - R.start[c][0]) * R.ext[c][1] + (y0 - R.start[c][1])) * ez +
                                        (z0 - R.start[c][2]);
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const double wxy = wx * double(Py[j]);
                        const double* r = row + j * ez;
#pragma unroll
                        for (int k = 0; k < 4; ++k) got += r[k] * (wxy * double(Pz[k]));
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const double wxy = wx * double(Py[j]);
#pragma unroll
                        for (int k = 0; k < 4; ++k)
                            got += double(fields[c][face_index(c, x, y0 + j, z0 + k, L.n)]) * (wxy * double(Pz[k]));
                    }
                }
                part[ti][pq * 3 + c] = got;
            }
        }
        __syncthreads();
        if (tid < 240) {
            for (int q = ar; q < cnt; q += 8) {
                const double got = ((part[0][q * 3 + ac] + part[1][q * 3 + ac]) + part[2][q * 3 + ac]) +
                                   part[3][q * 3 + ac];
                acc += nw[q * P2_NODES + aa] * got;
            }
        }
    }
    if (tid < 240) sums[aa * 3 + ac][ar] = acc;
    __syncthreads();
    if (tid < 30) {
        double v = 0.0;
#pragma unroll
        for (int r = 0; r < 8; ++r) v += sums[tid][r];
        unsafeAtomicAdd(rhs + cells[e * P2_NODES + tid / 3] * 3 + tid % 3, T(v));
    }
}

int make_lattice(const afsi_grid3* g, Lattice& L) {
    AFSI_REQUIRE(g, "null grid");
    AFSI_REQUIRE(!g->periodic, "the P2 fused transfer requires a closed MAC box");
    for (int a = 0; a < 3; ++a) {
        AFSI_REQUIRE(g->n[a] > 0 && g->n[a] < BASE_LIMIT, "invalid P2 MAC grid");
        AFSI_REQUIRE(std::isfinite(g->h[a]) && g->h[a] > 0, "invalid P2 MAC grid");
        L.n[a] = g->n[a];
        L.h[a] = g->h[a];
        L.origin[a] = g->origin[a];
        L.limit[a] = double(g->n[a] - 2);
    }
    return OK;
}

int check_group(const afsi_p2_group* g, const afsi_p2_tables* t) {
    AFSI_REQUIRE(g && t, "null P2 group or tables");
    AFSI_REQUIRE(g->elements >= 0 && g->quad > 0 && g->quad < (1 << 20) && g->node_count > 0,
                 "P2 transfer requires ten basis functions");
    AFSI_REQUIRE(g->elements == 0 || (g->cells && g->values), "null P2 group arrays");
    AFSI_REQUIRE(t->base && t->phi, "P2 IB requires shared face/center tables");
    AFSI_REQUIRE(t->offset >= 0 && t->total >= 0 && t->offset <= t->total &&
                     g->elements * g->quad <= t->total - t->offset,
                 "P2 IB group exceeds table bounds");
    AFSI_REQUIRE(g->elements < (int64_t(1) << 31), "too many P2 elements in one group");
    return OK;
}

}  // namespace
}  // namespace afsi

namespace afsi {
namespace {

template <class T>
int p2_prepare(const T* x, const afsi_p2_group* g, const afsi_grid3* grid, afsi_p2_tables* t, int* invalid,
               void* stream) {
    if (int r = check_group(g, t)) return r;
    Lattice L;
    if (int r = make_lattice(grid, L)) return r;
    AFSI_REQUIRE(x, "null positions");
    const int64_t count = g->elements * g->quad;
    if (!count) return OK;
    p2_prepare_kernel<T><<<grid_for(count, 256), 256, 0, hipStream_t(stream)>>>(
        x, g->cells, static_cast<const T*>(g->values), g->elements, int(g->quad), g->node_count, L, t->base,
        static_cast<T*>(t->phi), t->total, t->offset, invalid);
    AFSI_LAUNCH_CHECK();
    return OK;
}

template <class T>
int p2_spread(const T* value, const afsi_p2_group* g, const afsi_p2_tables* t, const afsi_grid3* grid,
              double volume, T* u0, T* u1, T* u2, int* flags, void* stream) {
    if (int r = check_group(g, t)) return r;
    Lattice L;
    if (int r = make_lattice(grid, L)) return r;
    AFSI_REQUIRE(value && u0 && u1 && u2 && g->weights, "null P2 spread arrays");
    AFSI_REQUIRE(std::isfinite(volume) && volume > 0, "invalid P2 MAC grid");
    if (!g->elements) return OK;
    p2_spread_kernel<T><<<unsigned(cdiv(g->elements, P2_BLOCK)), P2_THREADS, 0, hipStream_t(stream)>>>(
        value, g->cells, static_cast<const T*>(g->values), static_cast<const T*>(g->weights), g->elements,
        int(g->quad), t->base, static_cast<const T*>(t->phi), t->total, t->offset, u0, u1, u2, g->node_count, L,
        volume, flags);
    AFSI_LAUNCH_CHECK();
    return OK;
}

template <class T>
int p2_gather_assemble(const T* u0, const T* u1, const T* u2, const afsi_p2_group* g, const afsi_p2_tables* t,
                       const afsi_grid3* grid, T* rhs, int* flags, void* stream) {
    if (int r = check_group(g, t)) return r;
    Lattice L;
    if (int r = make_lattice(grid, L)) return r;
    AFSI_REQUIRE(u0 && u1 && u2 && rhs && g->weights, "null P2 gather arrays");
    if (!g->elements) return OK;
    p2_gather_assemble_kernel<T><<<unsigned(g->elements), P2_THREADS, 0, hipStream_t(stream)>>>(
        u0, u1, u2, g->cells, static_cast<const T*>(g->values), static_cast<const T*>(g->weights), int(g->quad),
        t->base, static_cast<const T*>(t->phi), t->total, t->offset, rhs, g->node_count, L, flags);
    AFSI_LAUNCH_CHECK();
    return OK;
}

}  // namespace
}  // namespace afsi

using namespace afsi;

extern "C" {

int afsi_p2_prepare_f64(const double* x, const afsi_p2_group* g, const afsi_grid3* grid, afsi_p2_tables* t,
                        int* invalid, void* stream) {
    return p2_prepare<double>(x, g, grid, t, invalid, stream);
}
int afsi_p2_prepare_f32(const float* x, const afsi_p2_group* g, const afsi_grid3* grid, afsi_p2_tables* t,
                        int* invalid, void* stream) {
    return p2_prepare<float>(x, g, grid, t, invalid, stream);
}
int afsi_p2_spread_f64(const double* value, const afsi_p2_group* g, const afsi_p2_tables* t,
                       const afsi_grid3* grid, double volume, double* u0, double* u1, double* u2, int* flags,
                       void* stream) {
    return p2_spread<double>(value, g, t, grid, volume, u0, u1, u2, flags, stream);
}
int afsi_p2_spread_f32(const float* value, const afsi_p2_group* g, const afsi_p2_tables* t,
                       const afsi_grid3* grid, double volume, float* u0, float* u1, float* u2, int* flags,
                       void* stream) {
    return p2_spread<float>(value, g, t, grid, volume, u0, u1, u2, flags, stream);
}
int afsi_p2_gather_assemble_f64(const double* u0, const double* u1, const double* u2, const afsi_p2_group* g,
                                const afsi_p2_tables* t, const afsi_grid3* grid, double* rhs, int* flags,
                                void* stream) {
    return p2_gather_assemble<double>(u0, u1, u2, g, t, grid, rhs, flags, stream);
}
int afsi_p2_gather_assemble_f32(const float* u0, const float* u1, const float* u2, const afsi_p2_group* g,
                                const afsi_p2_tables* t, const afsi_grid3* grid, float* rhs, int* flags,
                                void* stream) {
    return p2_gather_assemble<float>(u0, u1, u2, g, t, grid, rhs, flags, stream);
}

}  // extern "C"
