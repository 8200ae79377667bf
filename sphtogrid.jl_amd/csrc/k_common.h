// k_common.h -- dimension-independent stages shared by the 2D and 3D
// deposits (class partition, tile pair lists, work items).
#pragma once

#include "s2g_internal.h"

namespace s2g {

constexpr int kPartBlock = 256;
// large-list entries per counting-sort chunk (one LDS histogram per chunk)
constexpr int kPairChunk = 4096;
// per-workgroup LDS bound of the tile histograms (ints)
constexpr int kMaxTiles = 16384;

int64_t partition_blocks(int64_t n);
// counts: 3*partition_blocks(n)+1 int64; afterwards counts[c*nb] is the first
// index of class c in the concatenation and counts[3*nb] the grand total
void launch_partition(const int32_t *cls, int64_t n, int64_t *counts, int64_t *scan_tmp,
                      int32_t *l0, int32_t *l1, int32_t *l2, hipStream_t s);
// fill == false: count + scan into cnt (ntiles*nchunks+1 int64, last = total)
// fill == true : fill pairs from the scanned cnt
template <typename R>
void launch_pairs(const int32_t *large, int64_t nl, const R *rec, int nt, int ntc, int ntiles,
                  int nchunks, int64_t *cnt, int64_t *scan_tmp, int32_t *pairs_or_null,
                  bool fill, hipStream_t s);
// fill == false: nitems (ntiles+1, last = total) ; fill == true: items + tile_items
void launch_items(const int64_t *off, int ntiles, int nchunks, int64_t npairs, int64_t G,
                  int64_t *nitems, int64_t *scan_tmp, int4 *items_or_null, int64_t *tile_items,
                  bool fill, hipStream_t s);

}  // namespace s2g
