// k_accum.h -- element-wise kernels of the subfile reductions (k_accum.hip).
#pragma once

#include "s2g_internal.h"

namespace s2g {

// dst[i] += src[i] for finite src[i]
void launch_accumulate(const double *src, double *dst, int64_t n, hipStream_t s);
// map[i] /= w[i] for finite, non-zero w[i]
void launch_hp_reduce(double *map, const double *w, int64_t n, hipStream_t s);

}  // namespace s2g
