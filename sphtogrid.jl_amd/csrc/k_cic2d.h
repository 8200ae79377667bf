// k_cic2d.h -- launchers of the 2D deposit kernels (k_cic2d.hip).
#pragma once

#include "k_common.h"

namespace s2g {

// One map of a 2D deposit pass.  A pass deposits any number of maps that
// share the input dtype, the kernel and nq; their particles are concatenated
// ([pbase, pbase + n) in pass order) and so are their tile grids.  The device
// copy of this table is read by every stage.
struct Map2D {
    const void *pos, *hsml, *m, *rho, *binq, *w;
    double *image;    // nq + 1 planes of npix^2, deposit order (cic_2D.jl:120-123)
    double *out;      // reduced map (reduce_image_2D layout) or null
    double len2pix;
    int64_t npix;
    int64_t n;
    int64_t pbase;    // first particle of the map in the pass
    int32_t tbase;    // first gather tile of the map in the pass
    int32_t ntj;      // gather tiles per side
    int32_t sbase;    // first small-footprint tile of the map in the pass
    int32_t nts;      // small-footprint tiles per side
    uint32_t flags;
    int32_t small_side;  // box side up to which a particle is small (<= kSmallSide2D)
};

// small-footprint path: particles are bucketed by the kST x kST tile holding
// their footprint's first pixel; a work item accumulates its particles in a
// kSW x kSW LDS window (the tile plus the largest footprint) per plane
constexpr int kSW = kST + kSmallSide2D;

// host-side arguments of one map (Map2D without the pass offsets)
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
    int dtype, kernel, np, plane0, nq, nmaps;
    int fp32acc;  // accumulate each item's pixels in FP32 (S2G_FLAG_FP32_ACCUM)
    int64_t ntiles, nitems;  // nitems: the item count, or its upper bound with d_nitems
    const int64_t *d_nitems = nullptr;  // the item count on the device (replayed passes)
    const int *d_flag = nullptr;        // nonzero: skip the gather (capacity exceeded)
    const int4 *items;
    const int32_t *pairs;  // particle index per (tile, particle) pair
    const Rec2 *rec;
    const Map2D *maps;
    double *partial;
    const int64_t *tile_items;  // null: no large particle in the pass
    int add = 0;  // finalize adds its sums to the image instead of storing them
};

// setup over all particles of the pass: records and the class array
void launch_2d_setup(int dtype, const Map2D *maps, int nmaps, int64_t ntot, int nq, int kernel,
                     Rec2 *rec, int32_t *cls, hipStream_t s);
// ne: the count, or with d_bounds its upper bound (the count then being
// d_bounds[3] - d_bounds[2] on the device)
void launch_2d_norm(int kernel, const int32_t *list, int64_t ne, const Map2D *maps, Rec2 *rec,
                    hipStream_t s, const int64_t *d_bounds = nullptr);
void launch_2d_gather(const GatherArgs &g, hipStream_t s);
void launch_2d_finalize(const GatherArgs &g, hipStream_t s);
// keys (global small tile) and values (particle) of the small list
void launch_2d_small_keys(const int32_t *list, int64_t ns, const Rec2 *rec, const Map2D *maps,
                          uint32_t *keys, int32_t *vals, hipStream_t s);
// off[t] = first sorted entry of tile t, t in [0, ntiles]
void launch_tile_offsets(const uint32_t *keys, int64_t ns, int ntiles, int64_t *off,
                         hipStream_t s);
// one block per work item (grid = max_items; blocks past *nitems exit)
void launch_2d_small_tile(int dtype, int kernel, int np, const int4 *items,
                          const int64_t *nitems, int64_t max_items, const int32_t *plist,
                          const Rec2 *rec, const Map2D *maps, int nq, int plane0, hipStream_t s);
// reduce_image_2D of every map of the pass that has an output
void launch_2d_reduce_maps(const Map2D *maps, int nmaps, int ntiles, int nq, hipStream_t s);
// reduce_image_2D of one flat image
void launch_2d_reduce(const double *image, int64_t npix, int nq, int reduce, double *out,
                      hipStream_t s);

}  // namespace s2g
