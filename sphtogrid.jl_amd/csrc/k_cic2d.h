// k_cic2d.h -- launchers of the 2D deposit kernels (k_cic2d.hip).
#pragma once

#include "k_common.h"

namespace s2g {


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

// setup writes the per-particle records and the class array
void launch_2d_setup(const Map2DArgs &a, Rec2 *rec, int32_t *cls, hipStream_t s);
void launch_2d_norm(int kernel, const int32_t *list, int64_t ne, Rec2 *rec, hipStream_t s);
void launch_2d_gather(const GatherArgs &g, hipStream_t s);
void launch_2d_finalize(const GatherArgs &g, hipStream_t s);
void launch_2d_small(const SmallArgs &a, hipStream_t s);
void launch_2d_reduce(const double *image, int64_t npix, int nq, int reduce, double *out,
                      hipStream_t s);

}  // namespace s2g
