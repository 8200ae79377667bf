// k_cic3d.h -- launchers of the 3D deposit kernels (k_cic3d.hip).
#pragma once

#include "k_cic2d.h"

namespace s2g {

using Map3DArgs = Map2DArgs;  // same fields; binq is scalar (nq == 1)

struct Gather3Args {
    int dtype, kernel, np, plane0, nq, nb;
    int64_t ntiles, nitems, npix;
    const int4 *items;
    const int32_t *pairs;  // particle index per (tile, particle) pair
    const Rec3 *rec;
    const void *binq;
    double *partial;
    const int64_t *tile_items;
    double *image;
    int add = 0;  // finalize adds its sums to the image instead of storing them
};

struct Small3Args {
    int dtype, kernel;
    int64_t ns, npix;
    const int32_t *list;
    const Rec3 *rec;
    const void *binq;
    double *image;
};

void launch_3d_setup(const Map3DArgs &a, Rec3 *rec, int32_t *cls, hipStream_t s);
// explicit normalisation sums (flags: S2G_FLAG_EXACT_NORM turns off the
// border-cut closed form)
void launch_3d_norm(int kernel, const int32_t *list, int64_t ne, int64_t npix, uint32_t flags,
                    Rec3 *rec, hipStream_t s);
void launch_3d_gather(const Gather3Args &g, hipStream_t s);
void launch_3d_finalize(const Gather3Args &g, hipStream_t s);
void launch_3d_small(const Small3Args &a, hipStream_t s);
void launch_3d_reduce(const double *image, int64_t npix, int reduce, double *out, hipStream_t s);

}  // namespace s2g
