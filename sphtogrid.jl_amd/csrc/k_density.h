// k_density.h -- mass_density kernels (k_density.hip).
#pragma once

#include "s2g_internal.h"

namespace s2g {

// domain of the Morton keys of the neighbour search
struct MdGrid {
    double lo[3];       // lower corner (0 in a periodic box)
    double ext[3];      // extent per axis (the box size when periodic)
    double inv_ext[3];
    double box[3];      // periodic box size per axis (unused otherwise)
    double min_ext;     // narrowest extent
    int periodic;
};

// Dehnen & Aly (2012, MNRAS 425, 1068) eqs. 18-19, SPHKernels.jl
// bias_correction (called at mass_density.jl:86): the Wendland kernels'
// self-contribution bias eps * m * W(0), eps = eps100 * (N / 100)^(-alpha),
// removed when that leaves a positive density.  W(0) = norm * h_inv^dim
// (shape(0) = 1).  Other kernels are returned unchanged.
__host__ __device__ inline double md_bias_correction(int kernel, double rho, double m,
                                                     double w0, int nneigh)
{
    double e100, alpha;
    switch (kernel) {
    case S2G_KERNEL_WENDLAND_C2: e100 = 0.0294; alpha = 0.977; break;
    case S2G_KERNEL_WENDLAND_C4: e100 = 0.01342; alpha = 1.579; break;
    case S2G_KERNEL_WENDLAND_C6: e100 = 0.0116; alpha = 2.236; break;
    default: return rho;
    }
    const double corr = e100 * pow((double)nneigh * 0.01, -alpha) * m * w0;
    return rho - corr > 0.0 ? rho - corr : rho;
}

int64_t md_prep_blocks(int64_t n);
// largest Nneighbors the search supports
int md_max_neighbours();
// Morton bits per axis
int md_bits();
// FP64 (x, y, z, m), wrapped into a periodic box; bbox: 6 doubles per block
void launch_md_prep(int dtype, const void *pos, const void *mass, int64_t n, const MdGrid &g,
                    double4 *P, double *bbox, hipStream_t s);
void launch_md_keys(const double4 *P, int64_t n, const MdGrid &g, uint64_t *keys, int32_t *vals,
                    hipStream_t s);
void launch_md_sorted(const double4 *P, const int32_t *idx, int64_t n, double4 *S,
                      hipStream_t s);
// offsets of the 8^lt cells of level lt in the sorted keys (8^lt + 1 entries)
void launch_md_table(const uint64_t *K, int64_t n, int lt, int32_t *tab, hipStream_t s);
void launch_md_knn(int kernel, const double4 *S, const uint64_t *SK, const int32_t *order,
                   const int32_t *tab, int lt, int64_t n, const MdGrid &g, int k1, int dim,
                   int nneigh, double norm, double *rho, double *hsml, hipStream_t s);

}  // namespace s2g
