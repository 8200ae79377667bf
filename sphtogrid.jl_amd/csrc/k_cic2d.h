// k_cic2d.h -- launchers of the 2D deposit kernels (k_cic2d.hip).
#pragma once

#include "s2g_internal.h"

namespace s2g {

// large-list entries per counting-sort chunk (one LDS histogram per chunk)
constexpr int kPairChunk = 4096;
// per-workgroup LDS bound of the tile histograms (ints)
constexpr int kMaxTiles2D = 16384;

struct Map2DArgs {
    int dtype;
    const void *pos, *hsml, *m, *rho, *binq, *w;
    int nq;
    int64_t n, npix;
    double len2pix;
    int kernel;
    uint32_t flags;
};

struct GatherArgs {
    int dtype, kernel, np, plane0, nq, ntj;
    int64_t ntiles, nitems, npix;
    const int4 *items;
    const int32_t *pairs, *large;
    const Rec2 *rec;
    const void *binq;
    double *partial;
    const int64_t *tile_items;
    double *image;
};

struct SmallArgs {
    int dtype, kernel, nq;
    int64_t ns, npix;
    const int32_t *list;
    const Rec2 *rec;
    const void *binq;
    double *image;
};

void launch_2d_setup(const Map2DArgs &a, Rec2 *rec, hipStream_t s);
int64_t partition_blocks(int64_t n);
// counts: 3*partition_blocks(n)+1 int64; after the call counts[c*nb] is the
// first index of class c and counts[3*nb] the grand total
void launch_2d_partition(const Rec2 *rec, int64_t n, int64_t *counts, int64_t *scan_tmp,
                         int32_t *l0, int32_t *l1, int32_t *l2, hipStream_t s);
void launch_2d_norm(int kernel, const int32_t *list, int64_t ne, Rec2 *rec, hipStream_t s);
// fill == false: count + scan into cnt (ntiles*nchunks+1 int64, last = total)
// fill == true : fill pairs from the scanned cnt
void launch_2d_pairs(const int32_t *large, int64_t nl, const Rec2 *rec, int ntj, int ntiles,
                     int nchunks, int64_t *cnt, int64_t *scan_tmp, int32_t *pairs_or_null,
                     bool fill, hipStream_t s);
// fill == false: nitems (ntiles+1, last = total) ; fill == true: items + tile_items
void launch_2d_items(const int64_t *off, int ntiles, int nchunks, int64_t npairs, int64_t G,
                     int64_t *nitems, int64_t *scan_tmp, int4 *items_or_null,
                     int64_t *tile_items, bool fill, hipStream_t s);
void launch_2d_gather(const GatherArgs &g, hipStream_t s);
void launch_2d_finalize(const GatherArgs &g, hipStream_t s);
void launch_2d_small(const SmallArgs &a, hipStream_t s);
void launch_2d_reduce(const double *image, int64_t npix, int nq, int reduce, double *out,
                      hipStream_t s);

}  // namespace s2g
