// Point binning: integer cell keys, stable LSD radix sort, CSR offsets.
//
// The cell of a point on axis a is floor((x_a-origin_a)/h_a - 1) + 1: the
// reference's face-lattice stencil base (mac/compact_transfer.py:21-23)
// plus one, evaluated with the same IEEE operations (sub, div, sub, floor),
// so keys are bit-identical to the oracle's. Keys are tile-major: tiles of
// tx*ty*tz cells, tiles and cells inside a tile both x-major (z fastest), so
// every tile's points are one contiguous CSR segment.
//
// Sort: 8-bit-digit LSD radix passes. Each pass: per-block digit histogram
// (LDS atomics), per-digit exclusive scan over blocks, and a stable scatter
// that ranks keys inside a block with wave64 ballots: the lanes holding the
// same digit are found by AND-ing eight __ballot masks, their order is the
// lane order, and waves/rounds are combined in index order through LDS
// counters, so equal keys keep their input order (a stable sort).
#include "binshare.hpp"

namespace afsi {

constexpr int SORT_THREADS = 256;
constexpr int SORT_ROUNDS = 8;  // keys per thread per block
constexpr int SORT_TILE = SORT_THREADS * SORT_ROUNDS;
constexpr int RADIX = 256;
constexpr int SORT_WAVES = SORT_THREADS / WAVE;

static int64_t tile_count(const afsi_grid3& g, const int32_t t[3], int a) {
    return cdiv(g.n[a], t[a]);
}

static int64_t key_space(const afsi_grid3& g, const int32_t t[3]) {
    return tile_count(g, t, 0) * tile_count(g, t, 1) * tile_count(g, t, 2) *
           int64_t(t[0]) * t[1] * t[2];
}


// Cell index of one coordinate. The key space is below 2^31 (checked by
// every entry point), so cells and keys take 32-bit integer arithmetic; the
// 64-bit remainder runs only for a periodic point outside the domain.
__device__ inline uint32_t cell_of(double xa, double o, double h, int64_t n, bool periodic,
                                   int* flags) {
    const double fl = floor((xa - o) / h - 1.0);
    if (!isfinite(fl)) {
        raise_flag(flags, F_NONFINITE);
        return 0;
    }
    const int64_t c = int64_t(fl) + 1;
    if (c >= 0 && c < n) return uint32_t(c);
    if (periodic) return uint32_t(wrap(c, n));
    raise_flag(flags, F_SUPPORT);
    return c < 0 ? 0u : uint32_t(n - 1);  // keep every later access in bounds
}

// Tile-major key of point p of one instance.
__device__ inline uint32_t point_key(const double* __restrict__ x, int64_t p, const afsi_grid3& g,
                                     const TileGeom& tg, int* flags) {
    uint32_t c[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double xa = x[3 * p + a];
        if (!g.periodic) {
            const double s0 = (xa - g.origin[a]) / g.h[a];
            if (!(s0 >= 2.0) || !(s0 < double(g.n[a]) - 2.0)) raise_flag(flags, F_SUPPORT);
        }
        c[a] = cell_of(xa, g.origin[a], g.h[a], g.n[a], g.periodic != 0, flags);
    }
    const uint32_t t0 = uint32_t(tg.t[0]), t1 = uint32_t(tg.t[1]), t2 = uint32_t(tg.t[2]);
    const uint32_t tid = ((c[0] / t0) * uint32_t(tg.nt[1]) + c[1] / t1) * uint32_t(tg.nt[2]) + c[2] / t2;
    const uint32_t local = ((c[0] % t0) * t1 + c[1] % t1) * t2 + c[2] % t2;
    return tid * (t0 * t1 * t2) + local;
}

__global__ void keys_kernel(const double* __restrict__ x, int64_t points, afsi_grid3 g,
                            TileGeom tg, uint32_t* __restrict__ keys,
                            uint32_t* __restrict__ vals, int* flags) {
    for (int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < points;
         p += int64_t(gridDim.x) * blockDim.x) {
        keys[p] = point_key(x, p, g, tg, flags);
        vals[p] = uint32_t(p);
    }
}


__global__ void keys_batch_kernel(const BinDesc* __restrict__ desc, int n, int64_t total,
                                  uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    for (int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < total;
         j += int64_t(gridDim.x) * blockDim.x) {
        int lo = 0, hi = n;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (desc[mid].begin <= j) lo = mid;
            else hi = mid;
        }
        const BinDesc& d = desc[lo];
        keys[j] = uint32_t(d.key_offset) + point_key(d.x, j - d.begin, d.g, d.tg, d.flags);
        vals[j] = uint32_t(j);
    }
}

// Per-block digit counts, one LDS histogram per wavefront (atomics contend
// only within a wavefront), summed per digit at the end.
__global__ void __launch_bounds__(SORT_THREADS) radix_hist_kernel(const uint32_t* __restrict__ keys,
                                                                 int64_t n, int shift,
                                                                 uint32_t* __restrict__ hist,
                                                                 int64_t nblocks) {
    __shared__ uint32_t h[SORT_WAVES][RADIX];
    const int wave = threadIdx.x / WAVE;
#pragma unroll
    for (int w = 0; w < SORT_WAVES; ++w) h[w][threadIdx.x] = 0;
    __syncthreads();
    const int64_t start = int64_t(blockIdx.x) * SORT_TILE;
    uint32_t k[SORT_ROUNDS];
#pragma unroll
    for (int r = 0; r < SORT_ROUNDS; ++r) {
        const int64_t i = start + int64_t(r) * SORT_THREADS + threadIdx.x;
        k[r] = i < n ? keys[i] : 0u;
    }
#pragma unroll
    for (int r = 0; r < SORT_ROUNDS; ++r) {
        const int64_t i = start + int64_t(r) * SORT_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&h[wave][(k[r] >> shift) & (RADIX - 1)], 1u);
    }
    __syncthreads();
    uint32_t c = 0;
#pragma unroll
    for (int w = 0; w < SORT_WAVES; ++w) c += h[w][threadIdx.x];
    hist[int64_t(threadIdx.x) * nblocks + blockIdx.x] = c;
}

// Exclusive scan of hist[d][0..nblocks) for digit d = blockIdx.x; total[d].
// Each step covers SORT_THREADS * SCAN_PER consecutive entries: a thread sums
// its SCAN_PER entries, a wavefront scan of those sums (shuffles) and the
// wave totals in LDS give every thread its offset, and the entries are
// rewritten in place. Two barriers per step.
constexpr int SCAN_PER = 8;

__global__ void __launch_bounds__(SORT_THREADS) radix_scan_kernel(uint32_t* __restrict__ hist,
                                                                 int64_t nblocks,
                                                                 uint32_t* __restrict__ total) {
    __shared__ uint32_t wave_tot[SORT_WAVES];
    uint32_t* row = hist + int64_t(blockIdx.x) * nblocks;
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
    uint32_t running = 0;
    for (int64_t base = 0; base < nblocks; base += SORT_THREADS * SCAN_PER) {
        const int64_t i0 = base + int64_t(threadIdx.x) * SCAN_PER;
        uint32_t v[SCAN_PER];
        uint32_t sum = 0;
#pragma unroll
        for (int q = 0; q < SCAN_PER; ++q) {
            v[q] = i0 + q < nblocks ? row[i0 + q] : 0u;
            sum += v[q];
        }
        uint32_t inc = sum;
#pragma unroll
        for (int d = 1; d < WAVE; d <<= 1) {
            const uint32_t up = __shfl_up(inc, d, WAVE);
            if (lane >= d) inc += up;
        }
        if (lane == WAVE - 1) wave_tot[wave] = inc;
        __syncthreads();
        uint32_t run = running + inc - sum;
#pragma unroll
        for (int w = 0; w < SORT_WAVES; ++w) {
            if (w < wave) run += wave_tot[w];
            running += wave_tot[w];
        }
        __syncthreads();  // wave_tot is rewritten by the next step
#pragma unroll
        for (int q = 0; q < SCAN_PER; ++q) {
            if (i0 + q < nblocks) row[i0 + q] = run;
            run += v[q];
        }
    }
    if (threadIdx.x == 0) total[blockIdx.x] = running;
}

// Exclusive scan over the SORT_THREADS threads of one value each (wavefront
// shuffles, wave totals through LDS). Every thread must call it.
__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t* wave_tot) {
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
    uint32_t inc = v;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        const uint32_t up = __shfl_up(inc, d, WAVE);
        if (lane >= d) inc += up;
    }
    if (lane == WAVE - 1) wave_tot[wave] = inc;
    __syncthreads();
    uint32_t before = 0;
#pragma unroll
    for (int w = 0; w < SORT_WAVES; ++w)
        if (w < wave) before += wave_tot[w];
    __syncthreads();
    return before + inc - v;
}

// One tile of SORT_TILE keys per block, stable. The tile is first ordered
// by digit in LDS: ranks come from per-round wavefront ballots in input
// order (round r holds keys start + r * SORT_THREADS + tid), offset by the
// tile's own digit starts. It is then written out in that order, so
// consecutive threads store consecutive positions of one digit's run
// instead of one isolated word per key.
__global__ void __launch_bounds__(SORT_THREADS) radix_scatter_kernel(
    const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals, int64_t n, int shift,
    const uint32_t* __restrict__ hist, const uint32_t* __restrict__ total, int64_t nblocks,
    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out) {
    __shared__ uint32_t gbase[RADIX];            // global position of local index 0 for digit d
    __shared__ uint32_t run[RADIX];              // next local position of digit d
    __shared__ uint32_t cnt[SORT_WAVES][RADIX];  // per-wave counts, then positions
    __shared__ uint32_t wave_tot[SORT_WAVES];
    __shared__ uint32_t sk[SORT_TILE], sv[SORT_TILE];
    const int tid = threadIdx.x;
    const int lane = tid & (WAVE - 1), wave = tid / WAVE;
    const int64_t start = int64_t(blockIdx.x) * SORT_TILE;
    const int m = int(n - start < SORT_TILE ? n - start : SORT_TILE);
    uint32_t key[SORT_ROUNDS], val[SORT_ROUNDS];
    run[tid] = 0;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < SORT_ROUNDS; ++r) {
        const int e = r * SORT_THREADS + tid;
        key[r] = e < m ? keys[start + e] : 0u;
        val[r] = e < m ? vals[start + e] : 0u;
        if (e < m) atomicAdd(&run[(key[r] >> shift) & (RADIX - 1)], 1u);
    }
    __syncthreads();
    const uint32_t local_count = run[tid];
    const uint32_t local_start = block_excl_scan(local_count, wave_tot);
    const uint32_t digit_start = block_excl_scan(total[tid], wave_tot);
    gbase[tid] = digit_start + hist[int64_t(tid) * nblocks + blockIdx.x] - local_start;
    run[tid] = local_start;
    const uint64_t lt = (uint64_t(1) << lane) - 1;
#pragma unroll
    for (int r = 0; r < SORT_ROUNDS; ++r) {
        const bool valid = r * SORT_THREADS + tid < m;
        const uint32_t digit = (key[r] >> shift) & (RADIX - 1);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const uint64_t q = __ballot((digit >> b) & 1u);
            peers &= ((digit >> b) & 1u) ? q : ~q;
        }
#pragma unroll
        for (int w = 0; w < SORT_WAVES; ++w) cnt[w][tid] = 0u;
        __syncthreads();
        // The lowest peer lane publishes the wave's count for this digit.
        if (valid && (peers & lt) == 0) cnt[wave][digit] = uint32_t(__popcll(peers));
        __syncthreads();
        {   // thread tid owns digit tid: prefix over waves, advance the run.
            uint32_t at = run[tid];
#pragma unroll
            for (int w = 0; w < SORT_WAVES; ++w) {
                const uint32_t c = cnt[w][tid];
                cnt[w][tid] = at;
                at += c;
            }
            run[tid] = at;
        }
        __syncthreads();
        if (valid) {
            const uint32_t pos = cnt[wave][digit] + uint32_t(__popcll(peers & lt));
            sk[pos] = key[r];
            sv[pos] = val[r];
        }
        __syncthreads();  // cnt is cleared by the next round
    }
    for (int e = tid; e < m; e += SORT_THREADS) {
        const uint32_t k = sk[e];
        const uint32_t pos = gbase[(k >> shift) & (RADIX - 1)] + uint32_t(e);
        keys_out[pos] = k;
        vals_out[pos] = sv[e];
    }
}

// cell_ptr[k] = lower_bound(sorted keys, k) for k in [0, key_space].
// cell_ptr[c] = lower_bound(keys, c) for c in [0, space], in two passes.
// Mark: one thread per block of 256 cells binary-searches the block's first
// point, cell_ptr[256 b] = lower_bound(keys, 256 b) (one short round of
// independent searches, bounded work per thread).
// Fill: a workgroup takes one block [c0, c0 + 256), reads its point range
// [cell_ptr[c0], cell_ptr[c0 + 256]) (n past the end), stages that range of
// keys in LDS when it fits, and searches each cell there. A block reads its
// neighbour's start entry while that block rewrites it with the same value.
constexpr int CELL_PTR_STAGE = 2048;

__device__ inline int64_t lower_bound_keys(const uint32_t* __restrict__ keys, int64_t lo, int64_t hi,
                                           int64_t k) {
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (int64_t(keys[mid]) < k) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ void cell_mark_kernel(const uint32_t* __restrict__ keys, int64_t n, int64_t space,
                                 int32_t* __restrict__ ptr) {
    const int64_t nblk = (space >> 8) + 1;  // blocks whose first cell is <= space
    for (int64_t b = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; b < nblk;
         b += int64_t(gridDim.x) * blockDim.x)
        ptr[b << 8] = int32_t(lower_bound_keys(keys, 0, n, b << 8));
}

__global__ void __launch_bounds__(256) cell_ptr_kernel(const uint32_t* __restrict__ keys, int64_t n,
                                                      int64_t space, int32_t* __restrict__ ptr) {
    __shared__ uint32_t sk[CELL_PTR_STAGE];
    for (int64_t c0 = int64_t(blockIdx.x) * 256; c0 <= space; c0 += int64_t(gridDim.x) * 256) {
        const int64_t lo = ptr[c0], hi = c0 + 256 <= space ? int64_t(ptr[c0 + 256]) : n;
        const int64_t m = hi - lo;
        const int64_t c = c0 + threadIdx.x;
        __syncthreads();
        if (m <= CELL_PTR_STAGE) {
            for (int j = threadIdx.x; j < m; j += 256) sk[j] = keys[lo + j];
            __syncthreads();
            if (c <= space) {
                int a = 0, e = int(m);
                while (a < e) {
                    const int mid = (a + e) >> 1;
                    if (int64_t(sk[mid]) < c) a = mid + 1;
                    else e = mid;
                }
                ptr[c] = int32_t(lo + a);
            }
        } else if (c <= space) {
            ptr[c] = int32_t(lower_bound_keys(keys, lo, hi, c));
        }
    }
}

static void launch_cell_ptr(const uint32_t* keys, int64_t n, int64_t space, int32_t* ptr, hipStream_t s) {
    cell_mark_kernel<<<grid_for((space >> 8) + 1, 256, 8192), 256, 0, s>>>(keys, n, space, ptr);
    cell_ptr_kernel<<<grid_for(space + 1, 256, 8192), 256, 0, s>>>(keys, n, space, ptr);
}

static size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

static int64_t sort_blocks(int64_t n) { return cdiv(n, SORT_TILE); }

size_t bin_workspace_bytes(int64_t points) {
    const int64_t nb = sort_blocks(points) > 0 ? sort_blocks(points) : 1;
    return 4 * align256(size_t(points) * 4)        // keys/vals ping-pong buffers
           + align256(size_t(nb) * RADIX * 4)      // hist
           + align256(RADIX * 4);                  // totals
}

// LSD radix sort of (keys, vals) in k0/v0 (ping-pong with k1/v1; the last
// pass writes keys_sorted/perm), then cell_ptr over [0, space]. Keys lie in
// [0, space), or in [0, space] when `sentinel` (a key equal to space marks
// an entry that sorts last and lies outside every CSR range).
static int sort_and_index(int64_t points, int64_t space, uint32_t* k0, uint32_t* v0, uint32_t* k1,
                          uint32_t* v1, uint32_t* hist, uint32_t* total, uint32_t* keys_sorted,
                          int32_t* perm, int32_t* cell_ptr, hipStream_t s, bool sentinel = false) {
    const int64_t nb = sort_blocks(points);
    const int64_t bound = sentinel ? space + 1 : space;  // keys < bound
    int bits = 0;
    while ((int64_t(1) << bits) < bound) ++bits;
    int passes = (bits + 7) / 8;
    if (passes < 1) passes = 1;
    uint32_t* kin = k0;
    uint32_t* vin = v0;
    for (int pass = 0; pass < passes; ++pass) {
        const bool last = pass == passes - 1;
        uint32_t* kout = last ? keys_sorted : (kin == k0 ? k1 : k0);
        uint32_t* vout = last ? reinterpret_cast<uint32_t*>(perm) : (vin == v0 ? v1 : v0);
        radix_hist_kernel<<<unsigned(nb), SORT_THREADS, 0, s>>>(kin, points, 8 * pass, hist, nb);
        AFSI_LAUNCH_CHECK();
        radix_scan_kernel<<<RADIX, SORT_THREADS, 0, s>>>(hist, nb, total);
        AFSI_LAUNCH_CHECK();
        radix_scatter_kernel<<<unsigned(nb), SORT_THREADS, 0, s>>>(kin, vin, points, 8 * pass,
                                                                   hist, total, nb, kout, vout);
        AFSI_LAUNCH_CHECK();
        kin = kout;
        vin = vout;
    }
    launch_cell_ptr(keys_sorted, points, space, cell_ptr, s);
    AFSI_LAUNCH_CHECK();
    return OK;
}

// Workspace slices: keys/vals ping-pong, per-block histograms, totals.
struct SortSpace {
    uint32_t *k0, *v0, *k1, *v1, *hist, *total;
};

static SortSpace sort_space(void* workspace, int64_t points) {
    char* ws = static_cast<char*>(workspace);
    const size_t kb = align256(size_t(points) * 4);
    SortSpace w;
    w.k0 = reinterpret_cast<uint32_t*>(ws);
    w.v0 = reinterpret_cast<uint32_t*>(ws + kb);
    w.k1 = reinterpret_cast<uint32_t*>(ws + 2 * kb);
    w.v1 = reinterpret_cast<uint32_t*>(ws + 3 * kb);
    w.hist = reinterpret_cast<uint32_t*>(ws + 4 * kb);
    w.total = reinterpret_cast<uint32_t*>(ws + 4 * kb + align256(size_t(sort_blocks(points)) * RADIX * 4));
    return w;
}

// Keys/values already in the first two slices of the sort workspace
// (sort_key_slots): stable sort, then ptr = lower bounds over [0, space].
int sort_pairs(int64_t n, int64_t space, void* workspace, uint32_t* keys_sorted, int32_t* perm,
               int32_t* ptr, hipStream_t s) {
    AFSI_REQUIRE(n < (int64_t(1) << 31) && space < (int64_t(1) << 31), "sort exceeds 2^31 keys");
    if (n == 0) {
        launch_cell_ptr(keys_sorted, 0, space, ptr, s);
        AFSI_LAUNCH_CHECK();
        return OK;
    }
    const SortSpace w = sort_space(workspace, n);
    return sort_and_index(n, space, w.k0, w.v0, w.k1, w.v1, w.hist, w.total, keys_sorted, perm, ptr, s,
                          true);
}

void sort_key_slots(void* workspace, int64_t n, uint32_t** keys, uint32_t** vals) {
    const SortSpace w = sort_space(workspace, n);
    *keys = w.k0;
    *vals = w.v0;
}

int64_t bin_key_space(const afsi_grid3& g, const int32_t tile[3]) { return key_space(g, tile); }

TileGeom bin_tile_geom(const afsi_grid3& g, const int32_t tile[3]) {
    TileGeom tg;
    for (int a = 0; a < 3; ++a) {
        tg.t[a] = tile[a];
        tg.nt[a] = tile_count(g, tile, a);
    }
    return tg;
}

int bin_batch(const BinDesc* desc_device, int count, int64_t total_points, int64_t space,
              uint32_t* keys_sorted, int32_t* perm, int32_t* cell_ptr, void* sort_workspace,
              hipStream_t s) {
    AFSI_REQUIRE(total_points < (int64_t(1) << 31) && space < (int64_t(1) << 31),
                 "batch exceeds 2^31 points or keys");
    if (total_points == 0) {
        launch_cell_ptr(keys_sorted, 0, space, cell_ptr, s);
        AFSI_LAUNCH_CHECK();
        return OK;
    }
    const SortSpace w = sort_space(sort_workspace, total_points);
    keys_batch_kernel<<<grid_for(total_points, 256, 8192), 256, 0, s>>>(desc_device, count, total_points,
                                                                         w.k0, w.v0);
    AFSI_LAUNCH_CHECK();
    return sort_and_index(total_points, space, w.k0, w.v0, w.k1, w.v1, w.hist, w.total, keys_sorted,
                          perm, cell_ptr, s);
}

}  // namespace afsi

using namespace afsi;

extern "C" {

int64_t afsi_bin_key_space(const afsi_grid3* grid, const int32_t tile[3]) {
    if (!grid || !tile || tile[0] <= 0 || tile[1] <= 0 || tile[2] <= 0) return -1;
    return key_space(*grid, tile);
}

size_t afsi_bin_workspace_bytes(int64_t points) { return bin_workspace_bytes(points); }

int afsi_bin_points(const double* x, int64_t points, const afsi_grid3* grid,
                    const int32_t tile[3], uint32_t* keys_sorted, int32_t* perm,
                    int32_t* cell_ptr, void* workspace, size_t workspace_bytes, int* flags,
                    void* stream) {
    AFSI_REQUIRE(grid && tile && points >= 0, "invalid binning arguments");
    AFSI_REQUIRE(tile[0] > 0 && tile[1] > 0 && tile[2] > 0, "tile sizes must be positive");
    AFSI_REQUIRE(grid->n[0] > 0 && grid->n[1] > 0 && grid->n[2] > 0, "invalid grid");
    AFSI_REQUIRE(points < (int64_t(1) << 31), "at most 2^31-1 points per instance");
    const int64_t space = key_space(*grid, tile);
    AFSI_REQUIRE(space < (int64_t(1) << 31), "cell key space exceeds 2^31");
    AFSI_REQUIRE(cell_ptr, "null cell_ptr");
    const hipStream_t s = hipStream_t(stream);
    if (points == 0) {
        launch_cell_ptr(keys_sorted, 0, space, cell_ptr, s);
        AFSI_LAUNCH_CHECK();
        return OK;
    }
    AFSI_REQUIRE(x && keys_sorted && perm && workspace, "null pointer");
    AFSI_REQUIRE(workspace_bytes >= bin_workspace_bytes(points), "binning workspace too small");
    char* ws = static_cast<char*>(workspace);
    const size_t kb = align256(size_t(points) * 4);
    uint32_t* k0 = reinterpret_cast<uint32_t*>(ws);
    uint32_t* v0 = reinterpret_cast<uint32_t*>(ws + kb);
    uint32_t* k1 = reinterpret_cast<uint32_t*>(ws + 2 * kb);
    uint32_t* v1 = reinterpret_cast<uint32_t*>(ws + 3 * kb);
    uint32_t* hist = reinterpret_cast<uint32_t*>(ws + 4 * kb);
    const int64_t nb = sort_blocks(points);
    uint32_t* total = reinterpret_cast<uint32_t*>(ws + 4 * kb + align256(size_t(nb) * RADIX * 4));

    TileGeom tg;
    for (int a = 0; a < 3; ++a) {
        tg.t[a] = tile[a];
        tg.nt[a] = tile_count(*grid, tile, a);
    }
    keys_kernel<<<grid_for(points, 256, 8192), 256, 0, s>>>(x, points, *grid, tg, k0, v0, flags);
    AFSI_LAUNCH_CHECK();
    return sort_and_index(points, space, k0, v0, k1, v1, hist, total, keys_sorted, perm, cell_ptr, s);
    return OK;
}

}  // extern "C"
