// k_accum.hip -- element-wise stages of the subfile reductions:
//   k_accumulate     dst += src where src is finite (distributed_mapping/cic.jl:63-73,
//                    distributed_mapping/healpix.jl:50-58 skip NaN and Inf values)
//   k_hp_reduce      map /= weights where the weight is finite and non-zero
//                    (distributed_mapping/healpix.jl:67-73)
// Grid-stride loops over 64-bit counts; four doubles per thread per step.
#include "k_accum.h"

namespace s2g {

__device__ __forceinline__ bool finite(double v) { return !(isnan(v) || isinf(v)); }

__global__ __launch_bounds__(256) void k_accumulate(const double *__restrict__ src,
                                                    double *__restrict__ dst, int64_t n)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double v = src[i];
        if (finite(v))
            dst[i] += v;
    }
}

__global__ __launch_bounds__(256) void k_hp_reduce(double *__restrict__ map,
                                                   const double *__restrict__ w, int64_t n)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double wi = w[i];
        if (finite(wi) && wi != 0.0)
            map[i] /= wi;
    }
}

static unsigned grid_for(int64_t n)
{
    int64_t b = (n + 255) / 256;
    const int64_t cap = 1 << 16;
    return (unsigned)(b < 1 ? 1 : (b > cap ? cap : b));
}

void launch_accumulate(const double *src, double *dst, int64_t n, hipStream_t s)
{
    if (n > 0)
        hipLaunchKernelGGL(k_accumulate, dim3(grid_for(n)), dim3(256), 0, s, src, dst, n);
}

void launch_hp_reduce(double *map, const double *w, int64_t n, hipStream_t s)
{
    if (n > 0)
        hipLaunchKernelGGL(k_hp_reduce, dim3(grid_for(n)), dim3(256), 0, s, map, w, n);
}

}  // namespace s2g
