// Host-side interface between the binning and tiled-transfer translation
// units (used by the batched transfer entry points).
#pragma once
#include "common.hpp"

namespace afsi {

struct TileGeom {
    int32_t t[3];
    int64_t nt[3];
};

// Batched keys: global point j belongs to the instance whose [begin,
// begin + points) holds it; its key is key_offset + the instance's
// tile-major cell key, its value is j.
struct BinDesc {
    const double* x;
    int* flags;
    afsi_grid3 g;
    TileGeom tg;
    int64_t points;
    int64_t begin;       // first global point
    int64_t key_offset;  // first global key
};

int64_t bin_key_space(const afsi_grid3& g, const int32_t tile[3]);
TileGeom bin_tile_geom(const afsi_grid3& g, const int32_t tile[3]);
size_t bin_workspace_bytes(int64_t points);
// Generic stable sort of n (key < space, value) pairs through the radix
// passes of the binning: sort_key_slots names where the caller writes the
// unsorted pairs inside a workspace of bin_workspace_bytes(n); sort_pairs
// sorts them into keys_sorted / perm and writes ptr[k] = lower_bound(k) for
// k in [0, space]. Keys may equal space (entries that sort last and belong
// to no CSR range).
void sort_key_slots(void* workspace, int64_t n, uint32_t** keys, uint32_t** vals);
int sort_pairs(int64_t n, int64_t space, void* workspace, uint32_t* keys_sorted, int32_t* perm,
               int32_t* ptr, hipStream_t s);
// chunk_start[t] = sum over tiles before t of ceil(count_t / chunk), with
// count_t = cell_ptr[(t + 1) * tvol] - cell_ptr[t * tvol], and
// chunk_start[ntiles] the total (one workgroup).
int launch_chunk_scan(const int32_t* cell_ptr, int64_t ntiles, int64_t tvol, int chunk, int32_t* chunk_start,
                      hipStream_t s);
// keys/vals of all points of a batch into the sort workspace, then the
// stable radix sort and cell_ptr over [0, key_space].
int bin_batch(const BinDesc* desc_device, int count, int64_t total_points, int64_t key_space,
              uint32_t* keys_sorted, int32_t* perm, int32_t* cell_ptr, void* sort_workspace,
              hipStream_t s);

}  // namespace afsi
