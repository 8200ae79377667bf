// k_common.hip -- dimension-independent stages of the deposits:
//   * stable (particle-order) partition of the particle classes into lists
//   * the (tile, particle) counting sort into per-tile lists
//   * splitting tile lists into work items
// All offsets come from deterministic exclusive scans, so the lists are
// bit-reproducible.
#include "k_common.h"

namespace s2g {

// ---------------------------------------------------------------------------
// Stable partition into three particle-order lists from a class array.
//   pass 1: per-block class counts          (k_part_count)
//   scan  : exclusive over [class][block] (+ a sentinel: the grand total)
//   pass 2: ballot ranks inside the block    (k_part_scatter)
// list 0 = small, list 1 = large (An + Ex), list 2 = exact (Ex only)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void part_flags(int cls, int &f0, int &f1, int &f2)
{
    f0 = cls == kClsSmall;
    f1 = cls == kClsLargeAn || cls == kClsLargeEx;
    f2 = cls == kClsLargeEx;
}

__global__ __launch_bounds__(kPartBlock) void k_part_count(const int32_t *__restrict__ cls,
                                                           int64_t n,
                                                           int64_t *__restrict__ counts)
{
    __shared__ int s[3][kPartBlock / kWave];
    const int64_t p = (int64_t)blockIdx.x * kPartBlock + threadIdx.x;
    const int c = p < n ? cls[p] : kClsSkip;
    int f0, f1, f2;
    part_flags(c, f0, f1, f2);
    const int wv = threadIdx.x / kWave, ln = threadIdx.x % kWave;
    const uint64_t b0 = __ballot(f0), b1 = __ballot(f1), b2 = __ballot(f2);
    if (ln == 0) {
        s[0][wv] = __popcll(b0);
        s[1][wv] = __popcll(b1);
        s[2][wv] = __popcll(b2);
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        int t = 0;
        for (int k = 0; k < kPartBlock / kWave; ++k)
            t += s[threadIdx.x][k];
        counts[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = t;
    }
    if (blockIdx.x == 0 && threadIdx.x == 3)
        counts[3 * (int64_t)gridDim.x] = 0;  // becomes the grand total after the scan
}

__global__ __launch_bounds__(kPartBlock) void k_part_scatter(
    const int32_t *__restrict__ cls, int64_t n, const int64_t *__restrict__ offs,
    int32_t *__restrict__ l0, int32_t *__restrict__ l1, int32_t *__restrict__ l2)
{
    __shared__ int s[3][kPartBlock / kWave];
    const int64_t p = (int64_t)blockIdx.x * kPartBlock + threadIdx.x;
    const int c = p < n ? cls[p] : kClsSkip;
    int f[3];
    part_flags(c, f[0], f[1], f[2]);
    const int wv = threadIdx.x / kWave, ln = threadIdx.x % kWave;
    const uint64_t below = (ln == 0) ? 0ull : (~0ull >> (64 - ln));
    const uint64_t b[3] = {__ballot(f[0]), __ballot(f[1]), __ballot(f[2])};
    if (ln == 0)
        for (int k = 0; k < 3; ++k)
            s[k][wv] = __popcll(b[k]);
    __syncthreads();
    int32_t *lists[3] = {l0, l1, l2};
    for (int k = 0; k < 3; ++k) {
        if (!f[k])
            continue;
        // the scan runs over [class][block]: subtract the class base so the
        // position is local to list k (each list holds at most n entries)
        int64_t pos = offs[(int64_t)k * gridDim.x + blockIdx.x] - offs[(int64_t)k * gridDim.x];
        for (int w = 0; w < wv; ++w)
            pos += s[k][w];
        pos += __popcll(b[k] & below);
        lists[k][pos] = (int32_t)p;
    }
}

int64_t partition_blocks(int64_t n) { return (n + kPartBlock - 1) / kPartBlock; }

void launch_partition(const int32_t *cls, int64_t n, int64_t *counts, int64_t *scan_tmp,
                      int32_t *l0, int32_t *l1, int32_t *l2, hipStream_t s)
{
    if (n == 0)
        return;
    const unsigned nb = (unsigned)partition_blocks(n);
    hipLaunchKernelGGL(k_part_count, dim3(nb), dim3(kPartBlock), 0, s, cls, n, counts);
    exclusive_scan_i64(counts, 3 * (int64_t)nb + 1, scan_tmp, s);
    hipLaunchKernelGGL(k_part_scatter, dim3(nb), dim3(kPartBlock), 0, s, cls, n, counts, l0,
                       l1, l2);
}

// first entry of a particle-order list (len entries) whose particle index is
// >= key
__device__ __forceinline__ int64_t list_lower_bound(const int32_t *__restrict__ list, int64_t len,
                                                    int64_t key)
{
    int64_t lo = 0, hi = len;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if ((int64_t)list[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// class offsets and the per-map ranges of the large and small lists
__global__ void k_class_bounds(const int64_t *__restrict__ counts, int64_t nb,
                               const int32_t *__restrict__ small,
                               const int32_t *__restrict__ large,
                               const int64_t *__restrict__ pbase, int nmaps,
                               int64_t *__restrict__ out)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t ns = counts[nb] - counts[0];
    const int64_t nl = counts[2 * nb] - counts[nb];
    if (m < 4)
        out[m] = counts[(int64_t)m * nb];
    if (m > nmaps)
        return;
    if (m == nmaps) {
        out[4 + m] = nl;
        out[5 + nmaps + m] = ns;
        return;
    }
    out[4 + m] = list_lower_bound(large, nl, pbase[m]);
    out[5 + nmaps + m] = list_lower_bound(small, ns, pbase[m]);
}

void launch_class_bounds(const int64_t *counts, int64_t nb, const int32_t *small,
                         const int32_t *large, const int64_t *pbase, int nmaps, int64_t *out,
                         hipStream_t s)
{
    const int nt = nmaps + 1 > 4 ? nmaps + 1 : 4;
    hipLaunchKernelGGL(k_class_bounds, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s,
                       counts, nb, small, large, pbase, nmaps, out);
}

// ---------------------------------------------------------------------------
// (tile, particle) pair lists: a counting sort with deterministic order.
// Chunks are map-local runs of large-list entries (sized per map); map m's
// counts are tile-major [tile][chunk] at cnt[cbase], so one exclusive scan
// over all maps yields every list offset, ordered by (map, tile, particle).
// Local tiles are t = (a * nt + b) * ntc + c over the map's tile grid
// (ntc = 1 in 2D).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void tile_rect(const Rec2 &r, int &a0, int &a1, int &b0, int &b1,
                                          int &c0, int &c1)
{
    a0 = r.i0 / kGT;
    a1 = r.i1 / kGT;
    b0 = r.j0 / kGT;
    b1 = r.j1 / kGT;
    c0 = c1 = 0;
}

__device__ __forceinline__ void tile_rect(const Rec3 &r, int &a0, int &a1, int &b0, int &b1,
                                          int &c0, int &c1)
{
    a0 = r.i0 / kBT;
    a1 = r.i1 / kBT;
    b0 = r.j0 / kBT;
    b1 = r.j1 / kBT;
    c0 = r.k0 / kBT;
    c1 = r.k1 / kBT;
}

// Order in which the tile lists take a map's large particles.  The map's
// large-list segment e_0 .. e_{L-1} (particle order) is read as f_j =
// e_{(j r) mod L}, r the integer nearest to L / phi (golden ratio), raised to
// the first value coprime to L, so every entry comes once and consecutive
// entries of a tile's list lie far apart in the input.  Snapshots written in
// Peano-Hilbert order would otherwise give a gather batch of consecutive
// entries footprints that cover the same part of the tile, and the block's
// waves that own the other parts idle until its next barrier (2^24 cluster
// set on 4096^2, particles reordered by the 64-pixel tile of the centre:
// gather 282.6 ms in particle order, 272.8 ms read this way, against 270.7
// ms for the set in its generated order).
__device__ __forceinline__ int64_t large_stride(int64_t L)
{
    if (L <= 2)
        return 1;
    int64_t r = (int64_t)((double)L * 0.6180339887498949 + 0.5);
    for (;; ++r) {
        int64_t a = L, b = r;
        while (b) {
            const int64_t t = a % b;
            a = b;
            b = t;
        }
        if (a == 1)
            return r;
    }
}

// A chunk's list entries [l0, l1): entries lbase[m] + k csz .. of map m
// (lbase: the per-map list starts, on the device; csz = TileMap::csz),
// clipped to the map's list, so that the chunk table does not depend on the
// class counts.
__device__ __forceinline__ void chunk_range(const PairChunk &ch, const TileMap &tm,
                                            const int64_t *__restrict__ lbase, int64_t &l0,
                                            int64_t &l1)
{
    const int64_t a = lbase[ch.map] + (int64_t)ch.k * tm.csz, e = lbase[ch.map + 1];
    l0 = a;
    l1 = a + tm.csz < e ? a + tm.csz : e;
    if (l1 < l0)
        l1 = l0;
}

// the small-footprint bucket of a particle: the kST x kST tile holding the
// first pixel of its footprint (one tile, in the map's small-tile grid)
struct AnchorTile {
    const Rec2 &r;
};
__device__ __forceinline__ void tile_rect(const AnchorTile &t, int &a0, int &a1, int &b0,
                                          int &b1, int &c0, int &c1)
{
    a0 = a1 = t.r.i0 / kST;
    b0 = b1 = t.r.j0 / kST;
    c0 = c1 = 0;
}

template <typename R, bool Anchor>
__device__ __forceinline__ void list_rect(const R &r, int &a0, int &a1, int &b0, int &b1, int &c0,
                                          int &c1)
{
    if constexpr (Anchor)
        tile_rect(AnchorTile{r}, a0, a1, b0, b1, c0, c1);
    else
        tile_rect(r, a0, a1, b0, b1, c0, c1);
}

template <typename R, bool Anchor = false>
__global__ __launch_bounds__(256) void k_pair_count(const int32_t *__restrict__ large,
                                                    const R *__restrict__ rec,
                                                    const PairChunk *__restrict__ chunks,
                                                    const TileMap *__restrict__ tmaps,
                                                    const int64_t *__restrict__ lbase,
                                                    int64_t ncnt, int64_t *__restrict__ cnt)
{
    extern __shared__ int hist[];
    const PairChunk ch = chunks[blockIdx.x];
    const TileMap tm = tmaps[ch.map];
    int64_t cl0, cl1;
    chunk_range(ch, tm, lbase, cl0, cl1);
    // the pair lists read the map's segment in large_stride order (the
    // anchor buckets in list order)
    const int64_t mb = lbase[ch.map], L = lbase[ch.map + 1] - mb;
    const int64_t str = Anchor ? 1 : large_stride(L);
    for (int t = threadIdx.x; t < tm.ntiles; t += blockDim.x)
        hist[t] = 0;
    __syncthreads();
    for (int64_t l = cl0 + threadIdx.x; l < cl1; l += blockDim.x) {
        const R r = rec[large[Anchor ? l : mb + (l - mb) * str % L]];
        int a0, a1, b0, b1, c0, c1;
        list_rect<R, Anchor>(r, a0, a1, b0, b1, c0, c1);
        for (int a = a0; a <= a1; ++a)
            for (int b = b0; b <= b1; ++b)
                for (int c = c0; c <= c1; ++c)
                    atomicAdd(&hist[(a * tm.nt + b) * tm.ntc + c], 1);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < tm.ntiles; t += blockDim.x)
        cnt[tm.cbase + (int64_t)t * tm.nch + ch.k] = hist[t];
    if (blockIdx.x == 0 && threadIdx.x == 0)
        cnt[ncnt - 1] = 0;  // becomes the pair total after the scan
}

// one wave per chunk: sequential over the chunk's particles, lanes over the
// particle's tiles (distinct per lane), so every list follows the map's
// large_stride order.
// The records of 64 particles are loaded at once (a lane each) and each
// particle's tile rectangle is broadcast by readlane, so the loads' latency
// is paid once per 64 particles.
// pairs: cap entries; a slot at or past cap is not written (a replayed pass
// whose pair total outgrew its capacity; the caller sees the total)
template <typename R>
__global__ __launch_bounds__(kWave) void k_pair_fill(const int32_t *__restrict__ large,
                                                     const R *__restrict__ rec,
                                                     const PairChunk *__restrict__ chunks,
                                                     const TileMap *__restrict__ tmaps,
                                                     const int64_t *__restrict__ lbase,
                                                     const int64_t *__restrict__ off,
                                                     int32_t *__restrict__ pairs, int64_t cap)
{
    extern __shared__ int cursor[];
    const PairChunk ch = chunks[blockIdx.x];
    const TileMap tm = tmaps[ch.map];
    int64_t cl0, cl1;
    chunk_range(ch, tm, lbase, cl0, cl1);
    const int64_t mb = lbase[ch.map], L = lbase[ch.map + 1] - mb;
    const int64_t str = large_stride(L);
    for (int t = threadIdx.x; t < tm.ntiles; t += kWave)
        cursor[t] = (int)off[tm.cbase + (int64_t)t * tm.nch + ch.k];
    __syncthreads();
    for (int64_t lb = cl0; lb < cl1; lb += kWave) {
        const int n = cl1 - lb < kWave ? (int)(cl1 - lb) : kWave;
        int32_t pl = 0;
        int ra0 = 0, ra1 = -1, rb0 = 0, rb1 = -1, rc0 = 0, rc1 = -1;
        if ((int)threadIdx.x < n) {
            pl = large[mb + (lb + threadIdx.x - mb) * str % L];
            tile_rect(rec[pl], ra0, ra1, rb0, rb1, rc0, rc1);
        }
        for (int e = 0; e < n; ++e) {
            const int32_t p = __builtin_amdgcn_readlane(pl, e);
            const int a0 = __builtin_amdgcn_readlane(ra0, e), a1 = __builtin_amdgcn_readlane(ra1, e);
            const int b0 = __builtin_amdgcn_readlane(rb0, e), b1 = __builtin_amdgcn_readlane(rb1, e);
            const int c0 = __builtin_amdgcn_readlane(rc0, e), c1 = __builtin_amdgcn_readlane(rc1, e);
            const int wb = b1 - b0 + 1, wc = c1 - c0 + 1;
            const int ntr = (a1 - a0 + 1) * wb * wc;
            for (int k = threadIdx.x; k < ntr; k += kWave) {
                const int a = a0 + k / (wb * wc);
                const int rem = k % (wb * wc);
                const int b = b0 + rem / wc, c = c0 + rem % wc;
                const int slot = atomicAdd(&cursor[(a * tm.nt + b) * tm.ntc + c], 1);
                if (slot < cap)
                    pairs[slot] = p;
            }
        }
    }
}

template <typename R>
void launch_pairs(const int32_t *large, const R *rec, const PairChunk *chunks, int nchunks,
                  const TileMap *tmaps, int max_tiles, int64_t *cnt, int64_t ncnt,
                  int64_t *scan_tmp, int32_t *pairs_or_null, bool fill, hipStream_t s,
                  const int64_t *lbase, int64_t cap)
{
    const size_t lds = sizeof(int) * (size_t)(max_tiles > 0 ? max_tiles : 1);
    if (!fill) {
        hipLaunchKernelGGL((k_pair_count<R, false>), dim3((unsigned)nchunks), dim3(256), lds, s, large,
                           rec, chunks, tmaps, lbase, ncnt, cnt);
        exclusive_scan_i64(cnt, ncnt, scan_tmp, s);
    } else {
        hipLaunchKernelGGL(k_pair_fill<R>, dim3((unsigned)nchunks), dim3(kWave), lds, s, large,
                           rec, chunks, tmaps, lbase, cnt, pairs_or_null, cap);
    }
}

template void launch_pairs<Rec2>(const int32_t *, const Rec2 *, const PairChunk *, int,
                                 const TileMap *, int, int64_t *, int64_t, int64_t *, int32_t *,
                                 bool, hipStream_t, const int64_t *, int64_t);
template void launch_pairs<Rec3>(const int32_t *, const Rec3 *, const PairChunk *, int,
                                 const TileMap *, int, int64_t *, int64_t, int64_t *, int32_t *,
                                 bool, hipStream_t, const int64_t *, int64_t);

// Small-footprint buckets as lists over the maps' small-tile grids: counts
// per (tile, chunk) as for the pair lists (k_pair_count with the anchor
// tile), one exclusive scan, then one block per chunk placing its particles
// at LDS cursors.  Within a chunk the order in a tile follows the LDS
// atomics, which only reorders the small scatter's additions.
__global__ __launch_bounds__(256) void k_anchor_fill(const int32_t *__restrict__ list,
                                                     const Rec2 *__restrict__ rec,
                                                     const PairChunk *__restrict__ chunks,
                                                     const TileMap *__restrict__ tmaps,
                                                     const int64_t *__restrict__ lbase,
                                                     const int64_t *__restrict__ off,
                                                     int32_t *__restrict__ out)
{
    extern __shared__ int cursor[];
    const PairChunk ch = chunks[blockIdx.x];
    const TileMap tm = tmaps[ch.map];
    int64_t cl0, cl1;
    chunk_range(ch, tm, lbase, cl0, cl1);
    for (int t = threadIdx.x; t < tm.ntiles; t += blockDim.x)
        cursor[t] = (int)off[tm.cbase + (int64_t)t * tm.nch + ch.k];
    __syncthreads();
    for (int64_t l = cl0 + threadIdx.x; l < cl1; l += blockDim.x) {
        const int32_t p = list[l];
        const Rec2 r = rec[p];
        const int slot = atomicAdd(&cursor[(r.i0 / kST) * tm.nt + r.j0 / kST], 1);
        out[slot] = p;
    }
}

void launch_anchor_lists(const int32_t *list, const Rec2 *rec, const PairChunk *chunks,
                         int nchunks, const TileMap *tmaps, int max_tiles, int64_t *cnt,
                         int64_t ncnt, int64_t *scan_tmp, int32_t *out, bool fill,
                         const int64_t *lbase, hipStream_t s)
{
    const size_t lds = sizeof(int) * (size_t)(max_tiles > 0 ? max_tiles : 1);
    if (!fill) {
        hipLaunchKernelGGL((k_pair_count<Rec2, true>), dim3((unsigned)nchunks), dim3(256), lds, s,
                           list, rec, chunks, tmaps, lbase, ncnt, cnt);
        exclusive_scan_i64(cnt, ncnt, scan_tmp, s);
    } else {
        hipLaunchKernelGGL(k_anchor_fill, dim3((unsigned)nchunks), dim3(256), lds, s, list, rec,
                           chunks, tmaps, lbase, cnt, out);
    }
}

// Capacity flag of a replayed pass (written whole, one block): bit 0, a
// map's large list is longer than its chunks cover (nch * csz entries; the
// rest of the list was neither counted nor placed); bit 1, the pair total
// passed the pair capacity (the pairs past it were not written).
__global__ __launch_bounds__(256) void k_plan_check(const int64_t *__restrict__ lbase, int nmaps,
                                                    const TileMap *__restrict__ tmaps,
                                                    const int64_t *__restrict__ pair_total,
                                                    int64_t pair_cap, int *__restrict__ flag)
{
    __shared__ int f;
    if (threadIdx.x == 0)
        f = 0;
    __syncthreads();
    for (int m = threadIdx.x; m < nmaps; m += blockDim.x)
        if (lbase[m + 1] - lbase[m] > (int64_t)tmaps[m].nch * tmaps[m].csz)
            atomicOr(&f, 1);
    __syncthreads();
    if (threadIdx.x == 0)
        *flag = f | (*pair_total > pair_cap ? 2 : 0);
}

void launch_plan_check(const int64_t *lbase, int nmaps, const TileMap *tmaps,
                       const int64_t *pair_total, int64_t pair_cap, int *flag, hipStream_t s)
{
    hipLaunchKernelGGL(k_plan_check, dim3(1), dim3(256), 0, s, lbase, nmaps, tmaps, pair_total,
                       pair_cap, flag);
}

// ---------------------------------------------------------------------------
// Work items: tile t's list [start, start+count) is split into ceil(count/G)
// items (global tile, first pair, count, map).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void tile_list(const int64_t *__restrict__ off,
                                          const TileMap *__restrict__ tmaps, int nmaps, int t,
                                          int &m, int64_t &s, int64_t &e)
{
    m = find_entry(tmaps, nmaps, t, [](const TileMap &x) { return x.tbase; });
    const TileMap tm = tmaps[m];
    const int64_t tl = t - tm.tbase;
    // a map without chunks has an empty block: both ends read the same entry
    s = off[tm.cbase + tl * tm.nch];
    e = off[tm.cbase + (tl + 1) * tm.nch];
}

// item size with the total read on the device (so no host round trip is
// needed between the count and the items), or gmin when total is null:
// ceil(total / target), at most gmax (gmax > 0: a large pass is cut into more
// items than target, which shortens the launch's tail), at least
// ceil(total / tmax) (tmax > 0: the items allocated, ntiles + tmax, suffice
// whatever the total) and at least gmin
__device__ __forceinline__ int64_t item_size(int64_t gmin, int64_t target, int64_t gmax,
                                             int64_t tmax, const int64_t *__restrict__ total)
{
    if (!total)
        return gmin;
    const int64_t t = *total;
    int64_t g = (t + target - 1) / target;
    if (gmax > 0 && g > gmax)
        g = gmax;
    if (tmax > 0) {
        const int64_t gb = (t + tmax - 1) / tmax;
        g = g > gb ? g : gb;
    }
    return g > gmin ? g : gmin;
}

__global__ void k_item_count(const int64_t *__restrict__ off, const TileMap *__restrict__ tmaps,
                             int nmaps, int ntiles, int64_t gmin, int64_t target, int64_t gmax,
                             int64_t tmax, const int64_t *__restrict__ total,
                             int64_t *__restrict__ nitems)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == ntiles)
        nitems[t] = 0;  // becomes the item total after the scan
    if (t >= ntiles)
        return;
    const int64_t G = item_size(gmin, target, gmax, tmax, total);
    int m;
    int64_t s, e;
    tile_list(off, tmaps, nmaps, t, m, s, e);
    nitems[t] = (e - s + G - 1) / G;
}

__global__ void k_item_fill(const int64_t *__restrict__ off, const TileMap *__restrict__ tmaps,
                            int nmaps, int ntiles, int64_t gmin, int64_t target, int64_t gmax,
                            int64_t tmax, const int64_t *__restrict__ total,
                            const int64_t *__restrict__ ibase,
                            int4 *__restrict__ items, int64_t *__restrict__ tile_items)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles)
        return;
    const int64_t G = item_size(gmin, target, gmax, tmax, total);
    int m;
    int64_t s, e;
    tile_list(off, tmaps, nmaps, t, m, s, e);
    const int64_t b = ibase[t];
    int64_t k = 0;
    for (int64_t q = s; q < e; q += G, ++k) {
        const int64_t len = e - q < G ? e - q : G;
        items[b + k] = make_int4(t, (int)q, (int)len, m);
    }
    tile_items[2 * t] = b;
    tile_items[2 * t + 1] = k;
}

void launch_items(const int64_t *off, const TileMap *tmaps, int nmaps, int ntiles,
                  int64_t gmin, int64_t target, const int64_t *total, int64_t *nitems,
                  int64_t *scan_tmp, int4 *items_or_null, int64_t *tile_items, bool fill,
                  hipStream_t s, int64_t gmax, int64_t tmax)
{
    const unsigned nb = (unsigned)((ntiles + 1 + 255) / 256);
    if (!fill) {
        hipLaunchKernelGGL(k_item_count, dim3(nb), dim3(256), 0, s, off, tmaps, nmaps, ntiles,
                           gmin, target, gmax, tmax, total, nitems);
        exclusive_scan_i64(nitems, ntiles + 1, scan_tmp, s);
    } else {
        hipLaunchKernelGGL(k_item_fill, dim3(nb), dim3(256), 0, s, off, tmaps, nmaps, ntiles,
                           gmin, target, gmax, tmax, total, nitems, items_or_null, tile_items);
    }
}

}  // namespace s2g
