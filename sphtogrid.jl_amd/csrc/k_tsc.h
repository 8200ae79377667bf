// k_tsc.h -- launcher of the TSC grid deposit.
#pragma once

#include "s2g_internal.h"

namespace s2g {

void launch_tsc(int dtype, const void *pos, const void *binq, int64_t n, int64_t npix,
                double len2pix, int periodic, double *grid, hipStream_t s);

}  // namespace s2g
