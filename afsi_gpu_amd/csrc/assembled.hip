// CSR-assembled quadrature IB operator for P1 meshes (gfx950): the frozen
// transfer of mac/assembled_transfer.py:1-8,
//   B_c[a, f] = sum_e sum_q W_eq N_qa phi_c(f - x_eq)   (a = cells[e, local]),
// contracted per cell (mac/csrc/ib.cu:56-123) and reduced over cells in a
// bounded open-addressing table keyed by node * faces + face
// (mac/hash_transfer_assembly.py:13-70): SplitMix64 finaliser buckets,
// linear probing of at most 128 slots, empty key -1, overflow -> flag.
// Spread and interpolation are then CSR products (csr_spmv below).
//
// Contraction: one wave per cell. The cell's stencil box of component c
// (the bounding box of its points' bases, widened by the four taps) holds
// one FP64 accumulator per local node and site in LDS. Points are taken in
// order; for each point lane = tap (i,j,k) = (l>>4, l>>2&3, l&3) adds
// W phi_x phi_y phi_z N_a to its site with a plain LDS read-modify-write:
// the 64 taps of one point are 64 distinct sites, so no atomics are needed
// and the per-site sums are deterministic. Non-zero sites are then added to
// the table with global atomics (the only order-dependent step, across
// cells).
#include "common.hpp"

namespace afsi {
namespace {

constexpr int AS_NODES = 4;      // P1 cells: triangles (3) or tetrahedra (4)
constexpr int AS_BOX = 2048;     // LDS accumulators per cell (nodes x sites)
constexpr int HASH_PROBES = 128;

__device__ inline uint64_t splitmix_bucket(int64_t key, int64_t capacity) {
    uint64_t h = uint64_t(key);
    h = (h ^ (h >> 30)) * 0xbf58476d1ce4e5b9ULL;
    h = (h ^ (h >> 27)) * 0x94d049bb133111ebULL;
    return (h ^ (h >> 31)) & uint64_t(capacity - 1);
}

struct Table {
    int64_t* keys;
    double* values;
    int64_t capacity;  // power of two
    int* flag;
};

// values[slot(key)] += v; zero contributions are skipped; a negative key or
// HASH_PROBES occupied slots raise the flag (the caller retries larger).
__device__ inline void table_add(const Table& t, int64_t key, double v) {
    if (v == 0.0) return;
    if (key < 0) {
        atomicOr(t.flag, 1);
        return;
    }
    uint64_t slot = splitmix_bucket(key, t.capacity);
    for (int attempt = 0; attempt < HASH_PROBES; ++attempt, slot = (slot + 1) & uint64_t(t.capacity - 1)) {
        unsigned long long* k = reinterpret_cast<unsigned long long*>(t.keys + slot);
        unsigned long long seen = __hip_atomic_load(k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (seen == ~0ull) seen = atomicCAS(k, ~0ull, (unsigned long long)key);
        if (seen == ~0ull || seen == (unsigned long long)key) {
            unsafeAtomicAdd(t.values + slot, v);
            return;
        }
    }
    atomicOr(t.flag, 1);
}

__global__ void __launch_bounds__(256) table_accumulate_kernel(const int64_t* __restrict__ keys,
                                                               const double* __restrict__ values, int64_t count,
                                                               Table t) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
         i += int64_t(gridDim.x) * blockDim.x)
        table_add(t, keys[i], values[i]);
}

__device__ inline int wmin(int v) {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) v = min(v, __shfl_xor(v, m));
    return v;
}
__device__ inline int wmax(int v) {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) v = max(v, __shfl_xor(v, m));
    return v;
}

// One wave per cell e of component c. base (3,P,3) / phi (3,P,3,4) are the
// compact per-component tables of cell-major interaction points; the cells'
// points start at point `offset` of each lattice.
__global__ void __launch_bounds__(WAVE) contract_kernel(afsi_fe_geometry g, const int64_t* __restrict__ base,
                                                        const double* __restrict__ phi, int64_t P, int64_t offset,
                                                        int c, int64_t n0, int64_t n1, int64_t n2, Table t) {
    __shared__ double acc[AS_BOX];
    const int lane = threadIdx.x;
    const int64_t e = blockIdx.x;
    const int Q = g.quad, A = g.nodes_per_cell;
    const int64_t* cb = base + (int64_t(c) * P + offset + e * Q) * 3;
    const double* cp = phi + (int64_t(c) * P + offset + e * Q) * 12;
    int lo[3], hi[3];
    bool ok = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) { lo[a] = 1 << 30; hi[a] = -(1 << 30); }
    for (int q = lane; q < Q; q += WAVE) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int64_t v = cb[q * 3 + a];
            ok = ok && v > -(1 << 28) && v < (1 << 28);
            lo[a] = min(lo[a], int(v));
            hi[a] = max(hi[a], int(v));
        }
    }
    int w[3];
    const int64_t extent[3] = {n0, n1, n2};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        lo[a] = wmin(lo[a]);
        hi[a] = wmax(hi[a]);
        w[a] = hi[a] - lo[a] + 4;
    }
    int bad = ok ? 0 : 1;
    if (lane < A) {
        const int64_t node = g.cells[e * A + lane];
        if (node < 0 || node >= g.node_count) bad = 1;
    }
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) bad |= __shfl_xor(bad, m);
#pragma unroll
    for (int a = 0; a < 3; ++a)
        if (lo[a] < 0 || int64_t(lo[a]) + w[a] > extent[a]) bad = 1;
    if (bad) {
        if (lane == 0) atomicOr(t.flag, 2);
        return;
    }
    const int vol = w[0] * w[1] * w[2];
    const int64_t nf = n0 * n1 * n2;
    const int ti = lane >> 4, tj = (lane >> 2) & 3, tk = lane & 3;
    if (A * vol > AS_BOX) {  // box beyond LDS: every tap goes to the table
        for (int q = 0; q < Q; ++q) {
            const int64_t* b = cb + q * 3;
            const double* f = cp + q * 12;
            const double k = ((f[ti] * f[4 + tj]) * f[8 + tk]) * g.weights[e * Q + q];
            const int64_t face = ((b[0] + ti) * n1 + (b[1] + tj)) * n2 + (b[2] + tk);
            for (int a = 0; a < A; ++a)
                table_add(t, g.cells[e * A + a] * nf + face, k * g.values[q * A + a]);
        }
        return;
    }
    for (int i = lane; i < A * vol; i += WAVE) acc[i] = 0.0;
    __syncthreads();
    for (int q = 0; q < Q; ++q) {
        const int64_t* b = cb + q * 3;
        const double* f = cp + q * 12;
        const double k = ((f[ti] * f[4 + tj]) * f[8 + tk]) * g.weights[e * Q + q];
        const int pos = ((int(b[0]) + ti - lo[0]) * w[1] + (int(b[1]) + tj - lo[1])) * w[2] + (int(b[2]) + tk - lo[2]);
        for (int a = 0; a < A; ++a) acc[a * vol + pos] += k * g.values[q * A + a];
        __syncthreads();  // the next point's taps may revisit these sites
    }
    for (int i = lane; i < A * vol; i += WAVE) {
        const double v = acc[i];
        if (v == 0.0) continue;
        const int a = i / vol;
        int r = i - a * vol;
        const int z = r % w[2] + lo[2];
        r /= w[2];
        const int y = r % w[1] + lo[1];
        const int x = r / w[1] + lo[0];
        table_add(t, g.cells[e * A + a] * nf + (int64_t(x) * n1 + y) * n2 + z, v);
    }
}

// y[r * ys] = scale * sum_k vals[k] x[cols[k] * xs], SPMV lanes per row.
constexpr int SPMV1_LANES = 8;
__global__ void __launch_bounds__(256) csr_spmv_kernel(afsi_csr a, const double* __restrict__ x, int64_t xs,
                                                       double* __restrict__ y, int64_t ys, double scale) {
    const int sub = threadIdx.x % SPMV1_LANES;
    const int64_t groups = int64_t(gridDim.x) * blockDim.x / SPMV1_LANES;
    for (int64_t r = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / SPMV1_LANES; r < a.rows; r += groups) {
        double s = 0.0;
        const int32_t end = a.row_ptr[r + 1];
        for (int32_t k = a.row_ptr[r] + sub; k < end; k += SPMV1_LANES) s += a.vals[k] * x[int64_t(a.cols[k]) * xs];
#pragma unroll
        for (int m = SPMV1_LANES / 2; m > 0; m >>= 1) s += __shfl_xor(s, m);
        if (sub == 0) y[r * ys] = s * scale;
    }
}

int check_table(const afsi_hash_table* t) {
    AFSI_REQUIRE(t && t->keys && t->values && t->flag, "null hash table");
    AFSI_REQUIRE(t->capacity >= 1 && (t->capacity & (t->capacity - 1)) == 0, "power-of-two hash capacity required");
    return OK;
}

}  // namespace
}  // namespace afsi

using namespace afsi;

extern "C" {

int afsi_hash_accumulate_f64(const int64_t* keys, const double* values, int64_t count, const afsi_hash_table* t,
                             void* stream) {
    if (int r = check_table(t)) return r;
    AFSI_REQUIRE(count >= 0 && (count == 0 || (keys && values)), "invalid hash input");
    if (!count) return OK;
    table_accumulate_kernel<<<grid_for(count, 256, 16384), 256, 0, hipStream_t(stream)>>>(
        keys, values, count, Table{t->keys, t->values, t->capacity, t->flag});
    AFSI_LAUNCH_CHECK();
    return OK;
}

int afsi_p1_contract_f64(const afsi_fe_geometry* g, const int64_t* base, const double* phi, int64_t lattice_points,
                         int64_t offset, int component, const afsi_grid3* grid, const afsi_hash_table* t,
                         void* stream) {
    if (int r = check_table(t)) return r;
    AFSI_REQUIRE(g && g->cells && g->values && g->weights && base && phi && grid, "null contraction argument");
    AFSI_REQUIRE(offset >= 0 && lattice_points >= 0 && g->elements * g->quad <= lattice_points - offset,
                 "the cells' interaction points exceed the stencil tables");
    AFSI_REQUIRE(g->nodes_per_cell >= 1 && g->nodes_per_cell <= AS_NODES, "P1 contraction takes cells of at most %d nodes",
                 AS_NODES);
    AFSI_REQUIRE(g->quad >= 1 && g->elements >= 0 && g->node_count > 0, "invalid FE geometry sizes");
    AFSI_REQUIRE(component >= 0 && component < 3, "component must be 0, 1 or 2");
    AFSI_REQUIRE(!grid->periodic, "the assembled operator requires a closed MAC box");
    int64_t n[3];
    for (int a = 0; a < 3; ++a) n[a] = grid->n[a] + (a == component ? 1 : 0);
    AFSI_REQUIRE(n[0] > 0 && n[1] > 0 && n[2] > 0 && n[0] * n[1] * n[2] < (int64_t(1) << 31),
                 "invalid MAC grid for the assembled operator");
    AFSI_REQUIRE(g->node_count < (int64_t(1) << 62) / (n[0] * n[1] * n[2]), "node x face keys exceed int64");
    if (!g->elements) return OK;
    AFSI_REQUIRE(g->elements < (int64_t(1) << 31), "too many cells");
    contract_kernel<<<unsigned(g->elements), WAVE, 0, hipStream_t(stream)>>>(
        *g, base, phi, lattice_points, offset, component, n[0], n[1], n[2],
        Table{t->keys, t->values, t->capacity, t->flag});
    AFSI_LAUNCH_CHECK();
    return OK;
}

int afsi_csr_spmv_f64(const afsi_csr* a, const double* x, int64_t x_stride, double* y, int64_t y_stride,
                      double scale, void* stream) {
    AFSI_REQUIRE(a && a->row_ptr && a->cols && a->vals && x && y && a->rows > 0, "invalid CSR product arguments");
    AFSI_REQUIRE(x_stride >= 1 && y_stride >= 1, "CSR product strides must be positive");
    csr_spmv_kernel<<<grid_for(a->rows * SPMV1_LANES, 256, 16384), 256, 0, hipStream_t(stream)>>>(
        *a, x, x_stride, y, y_stride, scale);
    AFSI_LAUNCH_CHECK();
    return OK;
}

}  // extern "C"
