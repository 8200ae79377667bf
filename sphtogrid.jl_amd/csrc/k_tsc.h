// k_tsc.h -- launcher of the TSC grid deposit.
#pragma once

#include "s2g_internal.h"

namespace s2g {

// key bits of launch_tsc_keys for an npix^3 grid
int tsc_key_bits(int64_t npix);
// (brick key, particle) pairs of the n particles, to be sorted by key
void launch_tsc_keys(int dtype, const void *pos, int64_t n, int64_t npix, double len2pix,
                     int periodic, uint32_t *keys, int32_t *vals, hipStream_t s);
// the deposit, over the particles in `order` (null: in input order)
void launch_tsc(int dtype, const void *pos, const void *binq, int64_t n, int64_t npix,
                double len2pix, int periodic, const int32_t *order, double *grid, hipStream_t s);

}  // namespace s2g
