// k_common.h -- dimension-independent stages shared by the 2D and 3D
// deposits (class partition, tile pair lists, work items).
#pragma once

#include "s2g_internal.h"

namespace s2g {

constexpr int kPartBlock = 256;
// largest usual count of large-list entries per counting-sort chunk (one LDS
// histogram per chunk, one wave per chunk in the fill); the host sizes each
// map's chunks from 256 up to this, larger only to bound the counters
constexpr int kPairChunk = 4096;
// per-workgroup LDS bound of the tile histograms (ints): tiles of one map
constexpr int kMaxTiles = 16384;

int64_t partition_blocks(int64_t n);
// counts: 3*partition_blocks(n)+1 int64; afterwards counts[c*nb] is the first
// index of class c in the concatenation and counts[3*nb] the grand total
void launch_partition(const int32_t *cls, int64_t n, int64_t *counts, int64_t *scan_tmp,
                      int32_t *l0, int32_t *l1, int32_t *l2, hipStream_t s);
// After launch_partition over the particles of several maps: out[0..3] =
// class offsets counts[0], counts[nb], counts[2nb], counts[3nb]; out[4 + m] =
// first entry of map m in the large list (particles of map m are
// [pbase[m], pbase[m+1]), and the partition is stable), out[4 + nmaps] = size
// of the large list; out[5 + nmaps + m]: the same for the small list, and
// out[5 + 2 nmaps] its size (out: 2 nmaps + 6 entries).
void launch_class_bounds(const int64_t *counts, int64_t nb, const int32_t *small,
                         const int32_t *large, const int64_t *pbase, int nmaps, int64_t *out,
                         hipStream_t s);

// Tile grid of one map in the (tile, particle) counting sort.  Local tile id
// t = (a * nt + b) * ntc + c; global tile id tbase + t.  The map's count block
// is cnt[cbase + t * nch + k] for its chunks k < nch.
struct TileMap {
    int64_t cbase;
    int32_t tbase, ntiles;
    int32_t nt, ntc;
    int32_t nch;
    int32_t csz;  // entries per chunk when the chunk ranges are taken on the device
};
// chunk k of map `map`: list entries lbase[map] + k csz .. (clipped to the
// map's list), with lbase the per-map list starts on the device
struct PairChunk {
    int32_t map, k;
};

// fill == false: count + scan into cnt (ncnt int64, the last is the total)
// fill == true : fill pairs (particle indices) from the scanned cnt
// lbase: per-map large-list starts on the device (nmaps + 1 entries); a map's
// segment is taken in the order f_j = e_{(j r) mod L} (L its length, r the
// integer nearest to L / phi made coprime to L), so every tile list follows
// that order; cap: entries of pairs (slots past it are not written)
template <typename R>
void launch_pairs(const int32_t *large, const R *rec, const PairChunk *chunks, int nchunks,
                  const TileMap *tmaps, int max_tiles, int64_t *cnt, int64_t ncnt,
                  int64_t *scan_tmp, int32_t *pairs_or_null, bool fill, hipStream_t s,
                  const int64_t *lbase, int64_t cap = INT64_MAX);
// small-footprint buckets as per-tile lists (fill == false: counts per
// (tile, chunk) + scan into cnt; fill == true: out = the particles, tile t of
// map m at cnt[cbase + t nch] ..), chunk ranges from lbase as for the pairs
void launch_anchor_lists(const int32_t *list, const Rec2 *rec, const PairChunk *chunks,
                         int nchunks, const TileMap *tmaps, int max_tiles, int64_t *cnt,
                         int64_t ncnt, int64_t *scan_tmp, int32_t *out, bool fill,
                         const int64_t *lbase, hipStream_t s);
// flag = (a map's large list outgrew nch * csz entries) | 2 * (pair total >
// pair_cap), for a replayed pass sized at creation
void launch_plan_check(const int64_t *lbase, int nmaps, const TileMap *tmaps,
                       const int64_t *pair_total, int64_t pair_cap, int *flag, hipStream_t s);
// fill == false: nitems (ntiles+1, last = total) ; fill == true: items + tile_items.
// items[i] = (global tile, first pair, count, map).  A tile's list is cut into
// items of at most G entries, G = ceil(*total / target) capped at gmax (if
// gmax > 0), raised to ceil(*total / tmax) (if tmax > 0; then at most
// ntiles + tmax items) and to gmin; G = gmin if total is null.  *total is
// read on the device.
void launch_items(const int64_t *off, const TileMap *tmaps, int nmaps, int ntiles,
                  int64_t gmin, int64_t target, const int64_t *total, int64_t *nitems,
                  int64_t *scan_tmp, int4 *items_or_null, int64_t *tile_items, bool fill,
                  hipStream_t s, int64_t gmax = 0, int64_t tmax = 0);

// index of the last table entry whose key is <= v (keys ascending, key[0] <= v)
template <typename E, typename K, typename F>
__device__ __forceinline__ int find_entry(const E *tab, int n, K v, F key)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (key(tab[mid]) <= v)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

}  // namespace s2g
