// k_tsc.h -- launchers of the TSC grid deposit (a brick gather, k_tsc.hip).
#pragma once

#include "s2g_internal.h"

namespace s2g {

// brick of the gather: kTX x kTY x kTZ cells (z fastest in memory), one wave
// per brick; a 3-cell stencil meets on average (1 + 2/4)(1 + 2/8)^2 = 2.34
// bricks.  Small bricks keep the wave's LDS (two FP64 planes of 256 cells and
// the queue, 8 KB) small enough for 20 waves per CU: on 2^22 particles at
// 512^3, 8x8x16 bricks took 3.80 ms, 8x8x8 2.17, 8x8x4 2.19, 4x8x8 2.05 and
// 4x8x16 2.87 ms.
constexpr int kTX = 4, kTY = 8, kTZ = 8;

// bricks of an npix^3 grid (the sort keys run over [0, nbricks])
int64_t tsc_nbricks(int64_t npix);
// key bits of launch_tsc_keys for an npix^3 grid
int tsc_key_bits(int64_t npix);
// (brick key, particle) pairs of the n particles, to be sorted by key
void launch_tsc_keys(int dtype, const void *pos, int64_t n, int64_t npix, double len2pix,
                     int periodic, uint32_t *keys, int32_t *vals, hipStream_t s);
// off[b] = first sorted particle of brick b, b in [0, nbricks] (off[nbricks] =
// the particles that are not dropped)
void launch_tsc_offsets(const uint32_t *sorted_keys, int64_t n, int64_t npix, int64_t *off,
                        hipStream_t s);
// sorted grid coordinates + quantity and packed nearest cells
void launch_tsc_prep(int dtype, const void *pos, const void *binq, const int32_t *order,
                     const int64_t *off, int64_t n, int64_t npix, double len2pix, int periodic,
                     double4 *gq, uint64_t *cell, hipStream_t s);
// every cell of both planes (grid: 2 x npix^3 doubles)
void launch_tsc_gather(const double4 *gq, const uint64_t *cell, const int64_t *off,
                       int64_t npix, int periodic, double *grid, hipStream_t s);

}  // namespace s2g
