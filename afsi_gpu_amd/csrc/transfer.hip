// Stencil preparation and the reference-compatible compact / indexed IB
// transfers (op layouts of mac/csrc/transfer.cpp:34-53: compact stencils
// base (P,3) int64 + phi (P,3,4) per component, face arrays flat C-order,
// z fastest; indexed stencils ids (P,K) int64 + weights (P,K), fields (N,C)).
//
// Design (gfx950):
// * Gathers: one thread per point (compact); one lane per link and a
//   butterfly over the point's lanes when the indexed stencil width is a
//   power of two up to 64, else one thread per (point, component)
//   (indexed). A compact point's 64 taps are summed separably from its
//   12 one-dimensional weights (x rows of y rows of 4 contiguous z values,
//   fused multiply-adds), so a thread issues 16 row reads of 4 doubles
//   instead of 64 scattered ones.
// * Compact spread: the points are stably sorted by their stencil base into
//   8x8x8 base tiles (the binning's radix passes); one work item per (tile,
//   chunk of <= 512 points) accumulates its points' 64 links in an FP64 LDS
//   box of the tile's 11x11x11 faces and adds each nonzero entry to HBM
//   with one atomic, so a face gets one global update per touching work
//   item instead of one per link (like the reference's CUDA op, the order of
//   those updates is not fixed).
// * Indexed spread: a gather, deterministic and free of atomics: the links
//   are stably sorted by target and every output entry sums its links in
//   link order.
// Out-of-range stencils (the reference's device assert,
// mac/csrc/transfer.cu:19,41) raise F_INDEX: their in-range taps are kept
// by the gathers, and the spreads drop the point (compact) or link
// (indexed); callers turn the flag into an error.
#include "binshare.hpp"
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

template <class T>
__global__ void fill_kernel(T* x, int64_t count, T value) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
         i += int64_t(gridDim.x) * blockDim.x)
        x[i] = value;
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

// ---------------------------------------------------------------------------
// compact gather: one thread per point

template <class T>
__global__ void __launch_bounds__(256) compact_gather_kernel(const T* __restrict__ field,
                                                             const int64_t* __restrict__ base,
                                                             const T* __restrict__ phi, T* __restrict__ out,
                                                             int64_t ostride, int64_t points, int64_t nx,
                                                             int64_t ny, int64_t nz, int* flags) {
    for (int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < points;
         p += int64_t(gridDim.x) * blockDim.x) {
        const int64_t bx = base[3 * p], by = base[3 * p + 1], bz = base[3 * p + 2];
        T w[3][4];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int k = 0; k < 4; ++k) w[a][k] = phi[12 * p + 4 * a + k];
        const bool inside = bx >= 0 && bx + 3 < nx && by >= 0 && by + 3 < ny && bz >= 0 && bz + 3 < nz;
        T sum = T(0);
        if (inside) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                T sy = T(0);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const T* row = field + ((bx + i) * ny + (by + j)) * nz + bz;
                    T sz = row[0] * w[2][0];
#pragma unroll
                    for (int k = 1; k < 4; ++k) sz = __builtin_fma(row[k], w[2][k], sz);
                    sy = __builtin_fma(w[1][j], sz, sy);
                }
                sum = __builtin_fma(w[0][i], sy, sum);
            }
        } else {
            raise_flag(flags, F_INDEX);
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j)
                    for (int k = 0; k < 4; ++k) {
                        const int64_t xi = bx + i, yj = by + j, zk = bz + k;
                        if (xi < 0 || xi >= nx || yj < 0 || yj >= ny || zk < 0 || zk >= nz) continue;
                        sum += ((w[0][i] * w[1][j]) * w[2][k]) * field[(xi * ny + yj) * nz + zk];
                    }
        }
        out[p * ostride] = sum;
    }
}

// ---------------------------------------------------------------------------
// compact spread: sort points by stencil base, then one thread per face

constexpr int CP_T = 8;                        // base tile edge
constexpr int CP_TV = CP_T * CP_T * CP_T;       // bases per tile
constexpr int CP_BOX = CP_T + 3;                // faces per box axis
constexpr int CP_CHUNK = 512;                   // points per work item
constexpr int CP_THREADS = 256;

// Base b (per axis in [0, n-4]) -> tile-major key over the valid base
// positions (tiles of 8^3 bases, bases inside a tile x-major). A stencil
// reaching outside the face array gets the sentinel key `space` (sorted
// last, never visited) and raises F_INDEX.
__global__ void compact_keys_kernel(const int64_t* __restrict__ base, int64_t points, int64_t nx,
                                    int64_t ny, int64_t nz, int64_t space, uint32_t* __restrict__ keys,
                                    uint32_t* __restrict__ vals, int* flags) {
    const int64_t nbx = nx - 3, nby = ny - 3, nbz = nz - 3;
    const int64_t nty = cdiv(nby, CP_T), ntz = cdiv(nbz, CP_T);
    for (int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < points;
         p += int64_t(gridDim.x) * blockDim.x) {
        const int64_t bx = base[3 * p], by = base[3 * p + 1], bz = base[3 * p + 2];
        const bool ok = bx >= 0 && bx < nbx && by >= 0 && by < nby && bz >= 0 && bz < nbz;
        if (!ok) raise_flag(flags, F_INDEX);
        const int64_t tile = ((bx / CP_T) * nty + by / CP_T) * ntz + bz / CP_T;
        const int64_t local = ((bx % CP_T) * CP_T + by % CP_T) * CP_T + bz % CP_T;
        keys[p] = uint32_t(ok ? tile * CP_TV + local : space);
        vals[p] = uint32_t(p);
    }
}

// Work items per tile: exclusive scan of ceil(points / CP_CHUNK) over the
// tiles, by one workgroup.
__global__ void __launch_bounds__(1024) compact_chunks_kernel(const int32_t* __restrict__ ptr, int64_t tiles,
                                                              int32_t* __restrict__ start) {
    __shared__ int32_t s[1024];
    int32_t running = 0;
    for (int64_t b = 0; b < tiles; b += 1024) {
        const int64_t t = b + threadIdx.x;
        const int32_t v = t < tiles ? (ptr[(t + 1) * CP_TV] - ptr[t * CP_TV] + CP_CHUNK - 1) / CP_CHUNK : 0;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int d = 1; d < 1024; d <<= 1) {
            const int32_t add = int(threadIdx.x) >= d ? s[threadIdx.x - d] : 0;
            __syncthreads();
            s[threadIdx.x] += add;
            __syncthreads();
        }
        if (t < tiles) start[t] = running + s[threadIdx.x] - v;
        running += s[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) start[tiles] = running;
}

// One work item: the points of a chunk of one base tile, accumulated into the
// tile's LDS box with ds_add_f64 (a lane walks its (j,k) links from a
// lane-dependent rotation, so the lanes of one dense base hit 16 addresses
// at a time), then one atomic per nonzero box entry.
template <class T>
__global__ void __launch_bounds__(CP_THREADS) compact_tile_spread_kernel(
    const T* __restrict__ force, int64_t fstride, const int64_t* __restrict__ base, const T* __restrict__ phi,
    const int32_t* __restrict__ perm, const int32_t* __restrict__ ptr, const int32_t* __restrict__ start,
    int64_t tiles, int64_t nx, int64_t ny, int64_t nz, double inv_volume, T* __restrict__ out) {
    __shared__ double box[CP_BOX * CP_BOX * CP_BOX];
    const int64_t nty = cdiv(ny - 3, CP_T), ntz = cdiv(nz - 3, CP_T);
    const int total = start[tiles];
    const int rj = (threadIdx.x >> 2) & 3, rk = threadIdx.x & 3;
    for (int w = blockIdx.x; w < total; w += gridDim.x) {
        int lo = 0, hi = int(tiles);
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (start[mid] <= w) lo = mid;
            else hi = mid;
        }
        const int64_t tile = lo;
        const int chunk = w - start[lo];
        const int beg = ptr[tile * CP_TV] + chunk * CP_CHUNK;
        const int end = min(ptr[(tile + 1) * CP_TV], beg + CP_CHUNK);
        const int64_t ox = (tile / (nty * ntz)) * CP_T, oy = ((tile / ntz) % nty) * CP_T, oz = (tile % ntz) * CP_T;
        for (int e = threadIdx.x; e < CP_BOX * CP_BOX * CP_BOX; e += CP_THREADS) box[e] = 0.0;
        __syncthreads();
        for (int j = beg + int(threadIdx.x); j < end; j += CP_THREADS) {
            const int64_t p = perm[j];
            const int lx = int(base[3 * p] - ox), ly = int(base[3 * p + 1] - oy), lz = int(base[3 * p + 2] - oz);
            double w3[3][4];
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int k = 0; k < 4; ++k) w3[a][k] = double(phi[12 * p + 4 * a + k]);
            const double fv = double(force[p * fstride]) * inv_volume;
            double wyr[4], wzr[4];
            int oyr[4], ozr[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int jq = (q + rj) & 3, kq = (q + rk) & 3;
                wyr[q] = jq == 0 ? w3[1][0] : (jq == 1 ? w3[1][1] : (jq == 2 ? w3[1][2] : w3[1][3]));
                wzr[q] = kq == 0 ? w3[2][0] : (kq == 1 ? w3[2][1] : (kq == 2 ? w3[2][2] : w3[2][3]));
                oyr[q] = jq * CP_BOX;
                ozr[q] = kq;
            }
            const int b0 = (lx * CP_BOX + ly) * CP_BOX + lz;
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) {
                    const double wxy = w3[0][i] * wyr[jj];
                    const int row = b0 + i * CP_BOX * CP_BOX + oyr[jj];
#pragma unroll
                    for (int k = 0; k < 4; ++k) atomicAdd(box + row + ozr[k], (wxy * wzr[k]) * fv);
                }
        }
        __syncthreads();
        for (int e = threadIdx.x; e < CP_BOX * CP_BOX * CP_BOX; e += CP_THREADS) {
            const double v = box[e];
            if (v == 0.0) continue;
            const int64_t fx = ox + e / (CP_BOX * CP_BOX), fy = oy + (e / CP_BOX) % CP_BOX, fz = oz + e % CP_BOX;
            if (fx < nx && fy < ny && fz < nz) unsafeAtomicAdd(out + (fx * ny + fy) * nz + fz, T(v));
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// indexed gather: one thread per (point, component), links in order

template <class T>
__global__ void __launch_bounds__(256) indexed_gather_kernel(const T* __restrict__ field,
                                                             const int64_t* __restrict__ ids,
                                                             const T* __restrict__ w, T* __restrict__ out,
                                                             int64_t points, int64_t neighbors,
                                                             int64_t comps, int64_t nodes, int* flags) {
    for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < points * comps;
         t += int64_t(gridDim.x) * blockDim.x) {
        const int64_t p = t / comps, c = t % comps;
        const int64_t* row = ids + p * neighbors;
        const T* wr = w + p * neighbors;
        T acc = T(0);
        for (int64_t k = 0; k < neighbors; ++k) {
            const int64_t j = row[k];
            if (j < 0 || j >= nodes) {
                raise_flag(flags, F_INDEX);
                continue;
            }
            acc += field[j * comps + c] * wr[k];
        }
        out[t] = acc;
    }
}

// Indexed gather with one lane per link, for G = neighbors a power of two
// dividing the wavefront: the ids and weights of consecutive links are read
// by consecutive lanes (fully coalesced; the one-thread-per-point kernel above
// reads them with a 8*neighbors-byte stride), each lane forms its link's
// products, and the G lanes of a point add them with an xor butterfly. The
// launch loop advances whole workgroups and links = points * G, so a point's
// group is either entirely live or entirely idle and every lane takes part
// in every shuffle.
template <class T, int G>
__global__ void __launch_bounds__(256) indexed_gather_group_kernel(const T* __restrict__ field,
                                                                   const int64_t* __restrict__ ids,
                                                                   const T* __restrict__ w, T* __restrict__ out,
                                                                   int64_t links, int64_t comps, int64_t nodes,
                                                                   int* flags) {
    static_assert(G > 1 && G <= WAVE && (G & (G - 1)) == 0, "group must be a power of two within a wavefront");
    for (int64_t l0 = int64_t(blockIdx.x) * 256; l0 < links; l0 += int64_t(gridDim.x) * 256) {
        const int64_t l = l0 + threadIdx.x;
        const bool live = l < links;
        const int64_t j = live ? ids[l] : 0;
        const T wl = live ? w[l] : T(0);
        const bool ok = j >= 0 && j < nodes;
        if (live && !ok) raise_flag(flags, F_INDEX);
        for (int64_t c = 0; c < comps; ++c) {
            T v = live && ok ? field[j * comps + c] * wl : T(0);
#pragma unroll
            for (int s = G / 2; s > 0; s >>= 1) v += __shfl_xor(v, s);
            if (live && (threadIdx.x & (G - 1)) == 0) out[(l / G) * comps + c] = v;
        }
    }
}

// indexed spread: sort links by target node, one thread per (node, comp)
__global__ void link_keys_kernel(const int64_t* __restrict__ ids, int64_t links, int64_t size,
                                 uint32_t* __restrict__ keys, uint32_t* __restrict__ vals, int* flags) {
    for (int64_t l = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; l < links;
         l += int64_t(gridDim.x) * blockDim.x) {
        const int64_t j = ids[l];
        const bool ok = j >= 0 && j < size;
        if (!ok) raise_flag(flags, F_INDEX);
        keys[l] = uint32_t(ok ? j : size);
        vals[l] = uint32_t(l);
    }
}

constexpr int SEG_SPAN = 512;  // sorted links per wavefront in the segmented sums

// Per-node sums over the links sorted by node, as a segmented reduction of
// the flat sorted list (so a node with thousands of links is summed by many
// lanes, and every lane has work whatever the distribution). Wavefront gw
// owns links [gw * SEG_SPAN, +SEG_SPAN); it walks them 64 at a time with a
// segmented inclusive scan (Hillis-Steele over lanes, heads where the node
// changes) plus a carry between steps. A node whose links lie inside one
// span is written by that wavefront; the first and last node of a span
// that continue across its ends leave partial sums (head, tail), which
// seg_combine_kernel adds in span order. Links with key = size (out of range)
// are skipped. Every sum is formed in a fixed order, so the result is
// deterministic.
__device__ inline bool seg_single(const uint32_t* keys, int64_t links, int64_t gw, uint32_t& first,
                                  uint32_t& last, bool& from_prev, bool& cont_next) {
    const int64_t j0 = gw * SEG_SPAN, j1 = j0 + SEG_SPAN < links ? j0 + SEG_SPAN : links;
    first = keys[j0];
    last = keys[j1 - 1];
    from_prev = j0 > 0 && keys[j0 - 1] == first;
    cont_next = j1 < links && keys[j1] == last;
    return first == last;
}

template <class T>
__global__ void __launch_bounds__(256) seg_sum_kernel(const T* __restrict__ field, const T* __restrict__ w,
                                                      const uint32_t* __restrict__ keys,
                                                      const int32_t* __restrict__ perm, int64_t links,
                                                      int64_t size, int64_t neighbors, int64_t comps, int64_t c,
                                                      double volume, T* __restrict__ out,
                                                      double* __restrict__ part) {
    const int lane = threadIdx.x & 63;
    const int64_t gw = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    if (gw * SEG_SPAN >= links) return;  // wave-uniform
    const T vol = T(volume);
    uint32_t first, last;
    bool from_prev, cont_next;
    seg_single(keys, links, gw, first, last, from_prev, cont_next);
    const int64_t j0 = gw * SEG_SPAN, j1 = j0 + SEG_SPAN < links ? j0 + SEG_SPAN : links;
    T carry = T(0);
    for (int64_t b = j0; b < j1; b += 64) {
        const int64_t j = b + lane;
        const bool in = j < j1;
        const uint32_t k = in ? keys[j] : 0xffffffffu;
        T v = T(0);
        if (in && k < uint32_t(size)) {
            const int64_t l = perm[j];
            v = (field[(l / neighbors) * comps + c] * w[l]) / vol;
        }
        // head: the node starts here (within the span; the carry covers the
        // segment continuing from the previous step or span)
        const bool head = in && (j == j0 || k != keys[j - 1]);
        T sum = v;
        int h = head ? 1 : 0;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const T up = __shfl_up(sum, d, 64);
            const int hup = __shfl_up(h, d, 64);
            if (lane >= d) {
                if (!h) sum = up + sum;
                h = h | hup;
            }
        }
        if (!h) sum = carry + sum;  // no head from the step start to here: continues the carry
        const bool end = in && (j == j1 - 1 || k != keys[j + 1]);
        if (end && k < uint32_t(size)) {
            const bool is_first = k == first && from_prev;
            const bool is_last = k == last && cont_next;
            if (is_first) part[2 * gw] = double(sum);
            if (is_last) part[2 * gw + 1] = double(sum);
            if (!is_first && !is_last) out[int64_t(k) * comps + c] = sum;
        }
        // carry into the next step: lane 63's running sum unless its node ends there
        const T s63 = __shfl(sum, 63, 64);
        const bool e63 = __shfl(end ? 1 : 0, 63, 64) != 0;
        carry = e63 ? T(0) : s63;
    }
}

// Nodes whose links cross span ends: the span where the node starts adds its
// tail and the heads of the following spans, in order.
template <class T>
__global__ void __launch_bounds__(256) seg_combine_kernel(const uint32_t* __restrict__ keys, int64_t links,
                                                          int64_t size, int64_t comps, int64_t c,
                                                          const double* __restrict__ part, T* __restrict__ out) {
    const int64_t spans = cdiv(links, SEG_SPAN);
    for (int64_t gw = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; gw < spans;
         gw += int64_t(gridDim.x) * blockDim.x) {
        uint32_t first, last;
        bool from_prev, cont_next;
        const bool single = seg_single(keys, links, gw, first, last, from_prev, cont_next);
        if (!cont_next || (single && from_prev) || last >= uint32_t(size)) continue;
        T total = T(part[2 * gw + 1]);
        for (int64_t k = gw + 1; k < spans; ++k) {
            uint32_t f2, l2;
            bool p2, n2;
            const bool s2 = seg_single(keys, links, k, f2, l2, p2, n2);
            total = total + T(part[2 * k]);
            if (!(s2 && n2)) break;
        }
        out[int64_t(last) * comps + c] = total;
    }
}

// Workspace of a sort-based spread over n items and `space` targets: the
// radix sort's scratch, then sorted keys, permutation and CSR offsets.
struct SpreadSpace {
    size_t sort_bytes, keys_off, perm_off, ptr_off, total, start_off, part_off;
};

static size_t alignb(size_t b) { return (b + 255) & ~size_t(255); }

static SpreadSpace spread_space(int64_t n, int64_t space) {
    SpreadSpace w{};
    w.sort_bytes = alignb(bin_workspace_bytes(n > 0 ? n : 1));
    w.keys_off = w.sort_bytes;
    w.perm_off = w.keys_off + alignb(size_t(n > 0 ? n : 1) * 4);
    w.ptr_off = w.perm_off + alignb(size_t(n > 0 ? n : 1) * 4);
    w.total = w.ptr_off + alignb(size_t(space + 1) * 4);
    w.start_off = w.total;
    w.total += alignb(size_t(space / CP_TV + 1) * 4);
    w.part_off = w.total;  // head and tail partial sums of each SEG_SPAN span
    w.total += alignb(size_t(2 * cdiv(n > 0 ? n : 1, SEG_SPAN)) * sizeof(double));
    return w;
}

// Key space of the compact spread: whole 8^3 tiles over the valid bases.
static int64_t compact_tiles(int64_t nx, int64_t ny, int64_t nz) {
    return nx > 3 && ny > 3 && nz > 3 ? cdiv(nx - 3, CP_T) * cdiv(ny - 3, CP_T) * cdiv(nz - 3, CP_T) : 0;
}
static int64_t compact_space(int64_t nx, int64_t ny, int64_t nz) { return compact_tiles(nx, ny, nz) * CP_TV; }

template <class T>
static int compact_gather(const T* field, const int64_t* base, const T* phi, int64_t points,
                          int64_t nx, int64_t ny, int64_t nz, T* out, int64_t stride,
                          int* flags, void* stream) {
    AFSI_REQUIRE(points >= 0 && nx > 0 && ny > 0 && nz > 0 && stride > 0,
                 "invalid compact IB field/base");
    if (points == 0) return OK;
    AFSI_REQUIRE(field && base && phi && out, "null pointer");
    compact_gather_kernel<T><<<grid_for(points, 256, 16384), 256, 0, hipStream_t(stream)>>>(
        field, base, phi, out, stride, points, nx, ny, nz, flags);
    AFSI_LAUNCH_CHECK();
    return OK;
}

template <class T>
static int compact_spread(const T* force, int64_t fstride, const int64_t* base, const T* phi,
                          int64_t points, int64_t nx, int64_t ny, int64_t nz, double volume,
                          T* out, void* workspace, size_t workspace_bytes, int* flags, void* stream) {
    AFSI_REQUIRE(points >= 0 && nx > 0 && ny > 0 && nz > 0 && fstride > 0 &&
                     std::isfinite(volume) && volume > 0,
                 "invalid compact grid/volume");
    AFSI_REQUIRE(out, "null pointer");
    AFSI_REQUIRE(points < (int64_t(1) << 31) && nx * ny * nz < (int64_t(1) << 31),
                 "compact spread exceeds 2^31 points or faces");
    const hipStream_t s = hipStream_t(stream);
    const int64_t space = compact_space(nx, ny, nz), tiles = compact_tiles(nx, ny, nz);
    AFSI_REQUIRE(space < (int64_t(1) << 31), "compact spread exceeds 2^31 stencil bases");
    if (int r = launch_fill<T>(out, nx * ny * nz, T(0), s)) return r;
    if (points == 0) return OK;
    AFSI_REQUIRE(force && base && phi && workspace, "null pointer");
    const SpreadSpace L = spread_space(points, space);
    AFSI_REQUIRE(workspace_bytes >= L.total, "compact spread workspace too small");
    char* ws = static_cast<char*>(workspace);
    uint32_t *k0, *v0;
    sort_key_slots(ws, points, &k0, &v0);
    compact_keys_kernel<<<grid_for(points, 256, 8192), 256, 0, s>>>(base, points, nx, ny, nz, space, k0, v0,
                                                                     flags);
    AFSI_LAUNCH_CHECK();
    uint32_t* keys = reinterpret_cast<uint32_t*>(ws + L.keys_off);
    int32_t* perm = reinterpret_cast<int32_t*>(ws + L.perm_off);
    int32_t* ptr = reinterpret_cast<int32_t*>(ws + L.ptr_off);
    int32_t* start = reinterpret_cast<int32_t*>(ws + L.start_off);
    if (int r = sort_pairs(points, space, ws, keys, perm, ptr, s)) return r;
    if (tiles == 0) return OK;
    compact_chunks_kernel<<<1, 1024, 0, s>>>(ptr, tiles, start);
    AFSI_LAUNCH_CHECK();
    const int64_t upper = tiles + points / CP_CHUNK + 1;
    compact_tile_spread_kernel<T><<<unsigned(upper < 8192 ? upper : 8192), CP_THREADS, 0, s>>>(
        force, fstride, base, phi, perm, ptr, start, tiles, nx, ny, nz, 1.0 / volume, out);
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
    const int64_t links = points * neighbors;
    const unsigned blocks = grid_for(links, 256, 16384);
    switch (neighbors) {
    case 2: indexed_gather_group_kernel<T, 2><<<blocks, 256, 0, s>>>(field, ids, w, out, links, comps, nodes, flags); break;
    case 4: indexed_gather_group_kernel<T, 4><<<blocks, 256, 0, s>>>(field, ids, w, out, links, comps, nodes, flags); break;
    case 8: indexed_gather_group_kernel<T, 8><<<blocks, 256, 0, s>>>(field, ids, w, out, links, comps, nodes, flags); break;
    case 16: indexed_gather_group_kernel<T, 16><<<blocks, 256, 0, s>>>(field, ids, w, out, links, comps, nodes, flags); break;
    case 32: indexed_gather_group_kernel<T, 32><<<blocks, 256, 0, s>>>(field, ids, w, out, links, comps, nodes, flags); break;
    case 64: indexed_gather_group_kernel<T, 64><<<blocks, 256, 0, s>>>(field, ids, w, out, links, comps, nodes, flags); break;
    default:
        indexed_gather_kernel<T><<<grid_for(points * comps, 256, 16384), 256, 0, s>>>(
            field, ids, w, out, points, neighbors, comps, nodes, flags);
    }
    AFSI_LAUNCH_CHECK();
    return OK;
}

template <class T>
static int indexed_spread(const T* field, int64_t points, int64_t comps, const int64_t* ids,
                          const T* w, int64_t neighbors, int64_t size, double volume, T* out,
                          void* workspace, size_t workspace_bytes, int* flags, void* stream) {
    AFSI_REQUIRE(size > 0 && std::isfinite(volume) && volume > 0, "invalid IB output size/volume");
    AFSI_REQUIRE(comps > 0 && neighbors > 0 && points >= 0, "IB indexed stencil shape mismatch");
    AFSI_REQUIRE(out, "null pointer");
    const int64_t links = points * neighbors;
    AFSI_REQUIRE(links < (int64_t(1) << 31) && size < (int64_t(1) << 31), "indexed spread exceeds 2^31 links or nodes");
    const hipStream_t s = hipStream_t(stream);
    if (points == 0) return launch_fill<T>(out, size * comps, T(0), s);
    AFSI_REQUIRE(field && ids && w && workspace, "null pointer");
    const SpreadSpace L = spread_space(links, size);
    AFSI_REQUIRE(workspace_bytes >= L.total, "indexed spread workspace too small");
    char* ws = static_cast<char*>(workspace);
    uint32_t *k0, *v0;
    sort_key_slots(ws, links, &k0, &v0);
    link_keys_kernel<<<grid_for(links, 256, 8192), 256, 0, s>>>(ids, links, size, k0, v0, flags);
    AFSI_LAUNCH_CHECK();
    uint32_t* keys = reinterpret_cast<uint32_t*>(ws + L.keys_off);
    int32_t* perm = reinterpret_cast<int32_t*>(ws + L.perm_off);
    int32_t* ptr = reinterpret_cast<int32_t*>(ws + L.ptr_off);
    if (int r = sort_pairs(links, size, ws, keys, perm, ptr, s)) return r;
    // nodes without links stay 0; every other node is written exactly once
    if (int r = launch_fill<T>(out, size * comps, T(0), s)) return r;
    double* part = reinterpret_cast<double*>(ws + L.part_off);
    const int64_t spans = cdiv(links, SEG_SPAN);
    for (int64_t c = 0; c < comps; ++c) {
        seg_sum_kernel<T><<<unsigned(cdiv(spans, 4)), 256, 0, s>>>(field, w, keys, perm, links, size, neighbors,
                                                                  comps, c, volume, out, part);
        AFSI_LAUNCH_CHECK();
        seg_combine_kernel<T><<<grid_for(spans, 256, 4096), 256, 0, s>>>(keys, links, size, comps, c, part, out);
        AFSI_LAUNCH_CHECK();
    }
    return OK;
}

}  // namespace afsi

using namespace afsi;

extern "C" {

const char* afsi_last_error(void) { return afsi::g_error; }
int afsi_abi_version(void) { return 2; }

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
    return compact_gather<double>(field, base, phi, points, nx, ny, nz, out, out_stride, flags, stream);
}
int afsi_compact_gather_f32(const float* field, const int64_t* base, const float* phi,
                            int64_t points, int64_t nx, int64_t ny, int64_t nz, float* out,
                            int64_t out_stride, int* flags, void* stream) {
    return compact_gather<float>(field, base, phi, points, nx, ny, nz, out, out_stride, flags, stream);
}

size_t afsi_compact_spread_workspace_bytes(int64_t points, int64_t nx, int64_t ny, int64_t nz) {
    if (points < 0 || nx <= 0 || ny <= 0 || nz <= 0) return 0;
    return spread_space(points, compact_space(nx, ny, nz)).total;
}
int afsi_compact_spread_f64(const double* force, int64_t force_stride, const int64_t* base,
                            const double* phi, int64_t points, int64_t nx, int64_t ny,
                            int64_t nz, double volume, double* out, void* workspace,
                            size_t workspace_bytes, int* flags, void* stream) {
    return compact_spread<double>(force, force_stride, base, phi, points, nx, ny, nz, volume, out,
                                  workspace, workspace_bytes, flags, stream);
}
int afsi_compact_spread_f32(const float* force, int64_t force_stride, const int64_t* base,
                            const float* phi, int64_t points, int64_t nx, int64_t ny, int64_t nz,
                            double volume, float* out, void* workspace, size_t workspace_bytes,
                            int* flags, void* stream) {
    return compact_spread<float>(force, force_stride, base, phi, points, nx, ny, nz, volume, out,
                                 workspace, workspace_bytes, flags, stream);
}

int afsi_indexed_gather_f64(const double* field, int64_t nodes, int64_t comps,
                            const int64_t* ids, const double* weights, int64_t points,
                            int64_t neighbors, double* out, int* flags, void* stream) {
    return indexed_gather<double>(field, nodes, comps, ids, weights, points, neighbors, out, flags, stream);
}
int afsi_indexed_gather_f32(const float* field, int64_t nodes, int64_t comps, const int64_t* ids,
                            const float* weights, int64_t points, int64_t neighbors, float* out,
                            int* flags, void* stream) {
    return indexed_gather<float>(field, nodes, comps, ids, weights, points, neighbors, out, flags, stream);
}

size_t afsi_indexed_spread_workspace_bytes(int64_t points, int64_t neighbors, int64_t size) {
    if (points < 0 || neighbors <= 0 || size <= 0) return 0;
    return spread_space(points * neighbors, size).total;
}
int afsi_indexed_spread_f64(const double* field, int64_t points, int64_t comps,
                            const int64_t* ids, const double* weights, int64_t neighbors,
                            int64_t size, double volume, double* out, void* workspace,
                            size_t workspace_bytes, int* flags, void* stream) {
    return indexed_spread<double>(field, points, comps, ids, weights, neighbors, size, volume, out,
                                  workspace, workspace_bytes, flags, stream);
}
int afsi_indexed_spread_f32(const float* field, int64_t points, int64_t comps, const int64_t* ids,
                            const float* weights, int64_t neighbors, int64_t size, double volume,
                            float* out, void* workspace, size_t workspace_bytes, int* flags,
                            void* stream) {
    return indexed_spread<float>(field, points, comps, ids, weights, neighbors, size, volume, out,
                                 workspace, workspace_bytes, flags, stream);
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
