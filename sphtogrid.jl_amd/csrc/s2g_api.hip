// s2g_api.hip -- the C ABI of libsph2grid (include/sph2grid.h): context,
// staging of host arrays, orchestration of the device pipelines.
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "k_accum.h"
#include "k_cic3d.h"
#include "k_density.h"
#include "k_healpix.h"
#include "k_sort.h"
#include "k_tsc.h"
#include "s2g_internal.h"

using namespace s2g;

// Fixed device work space of s2g_tsc3d_device_noalloc: one copy per GPU the
// library runs on, held until the process ends.  Host code never uses this
// name as an address (it is a host-side stand-in there); it reads the device
// address of each GPU's copy with hipGetSymbolAddress once, in s2g_create,
// and passes that address to the kernels.
constexpr size_t kScratchBytes = (size_t)80 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// small_items_per_cu 16: the autotuner's choice for both the plain 256^2 map and the
// batched 64-map set (against 8: 0.334 vs 0.355 ms, 91.1 vs 92.9 ms; the
// cluster set 260.5 vs 261.2 ms)
constexpr s2g_tuning kDefaultTuning = {kSmallSide2D, 256, 16, 64, 16, 4, 4, 4096};

struct s2g_ctx {
    int device = 0;
    int ncu = 256;
    hipStream_t stream = nullptr;
    // side stream of the deposits (non-blocking, default priority): the
    // explicit normalisation and the small-footprint scatter run there beside
    // the pair lists and the gather, ordered by the fork / join events
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, normed = nullptr, smalled = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    static constexpr int kNumStageEvents = 80;
    hipEvent_t stage[kNumStageEvents] = {};
    std::string err;
    // staged host inputs
    DevBuf in[7];
    // 2D / 3D pipelines
    DevBuf rec, cls, counts, scan_tmp, lists[3], cnt, pairs, nitems, items, tile_items, partial;
    // per-pass tables: map table + particle bases, class bounds, tile maps + chunks
    DevBuf tab, bounds, tiles;
    // small-footprint path: sort keys / values (double-buffered), sort scratch,
    // tile offsets, work items
    DevBuf skeys, svals, sort_tmp, soff, snitems, sitems, stile, sscan_tmp;
    // mass_density: particles, bounding-box partials, particles in cell order
    DevBuf md_P, md_bbox, md_S, md_tab;
    PinnedBuf h_tab, h_bounds, h_tiles;
    // all-sky deposit
    DevBuf hp_rec, hp_nunits, hp_units, hp_part, hp_ring, hp_list, hp_rtab;
    // TSC: brick offsets, sorted grid coordinates + quantity, packed cells
    DevBuf tsc_off, tsc_gq, tsc_cell;
    // outputs
    DevBuf image, out;
    // pinned host scratch for small device->host reads
    int64_t *h_small = nullptr;
    // device address of `scratch` on this context's GPU (null if unavailable)
    void *scratch = nullptr;
    std::string scratch_err;
    // work split of the 2D deposit
    s2g_tuning tune = kDefaultTuning;
};

namespace {

const char *kVersion = "sph2grid 0.1.0 gfx950";

// per-GPU device address of `scratch`, looked up once per device
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_addr[kMaxDevices];
std::string g_scratch_err[kMaxDevices];

// With `device` current: the device address of its copy of `scratch`.
void *scratch_address(int device, std::string *err)
{
    if (device < 0 || device >= kMaxDevices) {
        *err = "device index beyond the work-space table";
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[device] && g_scratch_err[device].empty()) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e == hipSuccess && ptr)
            g_scratch_addr[device] = ptr;
        else
            g_scratch_err[device] = std::string("hipGetSymbolAddress: ") + hipGetErrorString(e);
    }
    *err = g_scratch_err[device];
    return g_scratch_addr[device];
}

template <typename F>
int guarded(s2g_ctx *c, F &&f)
{
    if (!c)
        return S2G_EINVAL;
    try {
        c->err.clear();
        int dev = -1;
        if (hipGetDevice(&dev) != hipSuccess || dev != c->device)
            S2G_HIP(hipSetDevice(c->device));
        return f();
    } catch (const Error &e) {
        c->err = e.msg;
        return e.code;
    } catch (const std::exception &e) {
        c->err = e.what();
        return S2G_EINVAL;
    } catch (...) {
        c->err = "unknown error";
        return S2G_EINVAL;
    }
}

size_t elem_size(int dtype) { return dtype == S2G_F32 ? 4 : 8; }

void check_common(int dtype, int kernel, int kernel_dim, int64_t n, const s2g_params *p)
{
    if (dtype != S2G_F32 && dtype != S2G_F64)
        throw Error(S2G_EINVAL, "dtype must be S2G_F32 or S2G_F64");
    if (kernel < S2G_KERNEL_CUBIC || kernel > S2G_KERNEL_WENDLAND_C8)
        throw Error(S2G_EUNSUPPORTED, "unsupported kernel id " + std::to_string(kernel));
    if (kernel_dim != 2 && kernel_dim != 3)
        throw Error(S2G_EINVAL, "kernel_dim must be 2 or 3");
    if (n < 0 || n > (int64_t)INT32_MAX)
        throw Error(S2G_EINVAL, "particle count out of range [0, 2^31)");
    if (!p)
        throw Error(S2G_EINVAL, "params is NULL");
    if (p->npixels < 1 || p->npixels > 65536)
        throw Error(S2G_EINVAL, "npixels out of range [1, 65536]");
    if (!(p->len2pix > 0.0) || !std::isfinite(p->len2pix))
        throw Error(S2G_EINVAL, "len2pix must be positive and finite");
}

// D2H of a few int64 values at device addresses (synchronises the stream)
void read_i64(s2g_ctx *c, hipStream_t s, std::initializer_list<const int64_t *> src,
              int64_t *dst)
{
    int k = 0;
    for (const int64_t *p : src)
        S2G_HIP(hipMemcpyAsync(&c->h_small[k++], p, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    S2G_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < k; ++i)
        dst[i] = c->h_small[i];
}

float elapsed(hipEvent_t a, hipEvent_t b)
{
    float ms = 0.f;
    S2G_HIP(hipEventElapsedTime(&ms, a, b));
    return ms;
}

// ---------------------------------------------------------------------------
// Shared large-particle stage: per-tile pair lists and work items.
// ---------------------------------------------------------------------------
struct TilePlan {
    int64_t npairs = 0, nitems = 0;
    int32_t *pairs = nullptr;
    int4 *items = nullptr;
    int64_t *tile_items = nullptr;
};

// tm: tile grid of every map of the pass (cbase / nch / csz filled here from
// the per-map large-list bounds lbase[0..nmaps]; d_lbase: the same on the
// device)
// Work items of <= G pairs, G = pairs / (ncu * per_cu) capped at gmax (> 0)
// and at least gmin: enough items to fill every CU several times, and for a
// large pass items of at most gmax pairs, so that the launch's tail is short
// (2^24 cluster set on 4096^2: 16 items per CU, 64k pairs each, gather 272.6
// ms; 4096-pair items, about 260 per CU, 246.2 ms).
TilePlan build_tile_lists(s2g_ctx *c, const int32_t *large, const void *rec, bool rec3,
                          std::vector<TileMap> &tm, const std::vector<int64_t> &lbase,
                          const int64_t *d_lbase, int ntiles, hipStream_t s, double *ms_sort,
                          int64_t gmin = 64, int64_t per_cu = 16, int64_t gmax = 0)
{
    hipEvent_t *E = c->stage;
    TilePlan tp;
    const int nmaps = (int)tm.size();
    std::vector<PairChunk> chunks;
    int64_t cb = 0;
    int max_tiles = 1;
    for (int m = 0; m < nmaps; ++m) {
        if (tm[m].ntiles > kMaxTiles)
            throw Error(S2G_EUNSUPPORTED, "grid too large for the tile lists (" +
                                              std::to_string(tm[m].ntiles) + " tiles > " +
                                              std::to_string(kMaxTiles) + ")");
        max_tiles = tm[m].ntiles > max_tiles ? tm[m].ntiles : max_tiles;
        const int64_t l0 = lbase[m], l1 = lbase[m + 1];
        // chunk size: about 32 chunks (one wave each in the fill) per CU,
        // 64..kPairChunk entries, grown while the map's counters (tiles x
        // chunks) would pass 2^25
        int64_t csz = (l1 - l0 + 32 * (int64_t)c->ncu - 1) / (32 * (int64_t)c->ncu);
        csz = csz < 64 ? 64 : csz > kPairChunk ? kPairChunk : (csz + 63) / 64 * 64;
        while ((int64_t)tm[m].ntiles * ((l1 - l0 + csz - 1) / csz) > (1 << 25))
            csz *= 2;
        const int nch = (int)((l1 - l0 + csz - 1) / csz);
        tm[m].cbase = cb;
        tm[m].nch = nch;
        tm[m].csz = (int32_t)csz;
        cb += (int64_t)tm[m].ntiles * nch;
        for (int k = 0; k < nch; ++k)
            chunks.push_back(PairChunk{m, k});
    }
    const int nchunks = (int)chunks.size();
    const int64_t ncnt = cb + 1;
    // one upload of the tile maps and the chunk table
    const size_t btm = sizeof(TileMap) * nmaps, bch = sizeof(PairChunk) * nchunks;
    char *h = (char *)c->h_tiles.get(btm + bch);
    std::memcpy(h, tm.data(), btm);
    std::memcpy(h + btm, chunks.data(), bch);
    char *d = (char *)c->tiles.get(btm + bch);
    S2G_HIP(hipMemcpyAsync(d, h, btm + bch, hipMemcpyHostToDevice, s));
    const TileMap *d_tm = (const TileMap *)d;
    const PairChunk *d_ch = (const PairChunk *)(d + btm);

    int64_t *cnt = (int64_t *)c->cnt.get(sizeof(int64_t) * ncnt);
    // the scans of the pair counts (ncnt) and of the item counts (ntiles + 1;
    // larger when maps of the pass have no large particle) share it
    const int64_t nscan = ncnt > (int64_t)ntiles + 1 ? ncnt : (int64_t)ntiles + 1;
    int64_t *tmp = (int64_t *)c->scan_tmp.get(sizeof(int64_t) * scan_tmp_elems(nscan));
    S2G_HIP(hipEventRecord(E[4], s));
    if (rec3)
        launch_pairs<Rec3>(large, (const Rec3 *)rec, d_ch, nchunks, d_tm, max_tiles, cnt, ncnt,
                           tmp, nullptr, false, s, d_lbase);
    else
        launch_pairs<Rec2>(large, (const Rec2 *)rec, d_ch, nchunks, d_tm, max_tiles, cnt, ncnt,
                           tmp, nullptr, false, s, d_lbase);
    // item counts from the scanned pair counts, the item size from the pair
    // total on the device: one read-back gives both totals
    const int64_t target = (int64_t)c->ncu * per_cu;
    const int64_t *d_npairs = cnt + ncnt - 1;
    int64_t *nit = (int64_t *)c->nitems.get(sizeof(int64_t) * (ntiles + 1));
    launch_items(cnt, d_tm, nmaps, ntiles, gmin, target, d_npairs, nit, tmp, nullptr, nullptr,
                 false, s, gmax);
    S2G_HIP(hipEventRecord(E[5], s));
    int64_t tot[2];
    read_i64(c, s, {d_npairs, nit + ntiles}, tot);
    tp.npairs = tot[0];
    tp.nitems = tot[1];
    if (tp.npairs >= (int64_t)INT32_MAX)
        throw Error(S2G_EUNSUPPORTED, "more than 2^31 (tile, particle) pairs");
    tp.pairs = (int32_t *)c->pairs.get(sizeof(int32_t) * (tp.npairs ? tp.npairs : 1));
    tp.items = (int4 *)c->items.get(sizeof(int4) * (tp.nitems ? tp.nitems : 1));
    tp.tile_items = (int64_t *)c->tile_items.get(sizeof(int64_t) * 2 * ntiles);
    S2G_HIP(hipEventRecord(E[6], s));
    if (rec3)
        launch_pairs<Rec3>(large, (const Rec3 *)rec, d_ch, nchunks, d_tm, max_tiles, cnt, ncnt,
                           tmp, tp.pairs, true, s, d_lbase);
    else
        launch_pairs<Rec2>(large, (const Rec2 *)rec, d_ch, nchunks, d_tm, max_tiles, cnt, ncnt,
                           tmp, tp.pairs, true, s, d_lbase);
    launch_items(cnt, d_tm, nmaps, ntiles, gmin, target, d_npairs, nit, tmp, tp.items,
                 tp.tile_items, true, s, gmax);
    S2G_HIP(hipEventRecord(E[7], s));
    S2G_HIP(hipEventSynchronize(E[7]));
    *ms_sort = elapsed(E[4], E[5]) + elapsed(E[6], E[7]);
    return tp;
}

struct ClassLists {
    int64_t ns = 0, nl = 0, ne = 0;
    int32_t *l0 = nullptr, *l1 = nullptr, *l2 = nullptr;
    std::vector<int64_t> lbase;  // per-map first large-list entry (+ the total)
    const int64_t *d_lbase = nullptr;  // the same on the device
};

// stable class partition of the pass's particles and one read-back of the
// class sizes and of the per-map large-list bounds
ClassLists partition_classes(s2g_ctx *c, const int32_t *cls, int64_t n, const int64_t *d_pbase,
                             int nmaps, hipStream_t s)
{
    ClassLists cl;
    const int64_t nb = partition_blocks(n);
    int64_t *counts = (int64_t *)c->counts.get(sizeof(int64_t) * (3 * nb + 1));
    cl.l0 = (int32_t *)c->lists[0].get(sizeof(int32_t) * n);
    cl.l1 = (int32_t *)c->lists[1].get(sizeof(int32_t) * n);
    cl.l2 = (int32_t *)c->lists[2].get(sizeof(int32_t) * n);
    int64_t *tmp = (int64_t *)c->scan_tmp.get(sizeof(int64_t) * scan_tmp_elems(3 * nb + 1));
    launch_partition(cls, n, counts, tmp, cl.l0, cl.l1, cl.l2, s);
    int64_t *bounds = (int64_t *)c->bounds.get(sizeof(int64_t) * (2 * nmaps + 6));
    launch_class_bounds(counts, nb, cl.l0, cl.l1, d_pbase, nmaps, bounds, s);
    S2G_HIP(hipEventRecord(c->stage[1], s));
    int64_t *h = (int64_t *)c->h_bounds.get(sizeof(int64_t) * (nmaps + 5));
    S2G_HIP(hipMemcpyAsync(h, bounds, sizeof(int64_t) * (nmaps + 5), hipMemcpyDeviceToHost, s));
    S2G_HIP(hipStreamSynchronize(s));
    cl.ns = h[1] - h[0];
    cl.nl = h[2] - h[1];
    cl.ne = h[3] - h[2];
    cl.lbase.assign(h + 4, h + 5 + nmaps);
    cl.d_lbase = bounds + 4;
    return cl;
}

void fill_class_stats(s2g_stats *st, int64_t n, const ClassLists &cl)
{
    if (!st)
        return;
    st->n_small = cl.ns;
    st->n_large = cl.nl;
    st->n_large_exact = cl.ne;
    st->n_skipped = n - cl.ns - cl.nl;
}

// ---------------------------------------------------------------------------
// 2D deposit pass over one or more maps on device pointers.  All maps share
// dtype, kernel and nq.  Each map's flat image (nq + 1 planes) is written in
// full; maps with an output also get reduce_image_2D.
// ---------------------------------------------------------------------------
struct MapIn {
    Map2DArgs a;
    double *image;
    double *out;
};

// The pass's tables: per-map particle and tile offsets and the grids of the
// small-footprint tiles (one bucket per tile), staged in pinned memory and
// uploaded on `s`.
struct PassTables {
    const Map2D *d_maps = nullptr;
    const int64_t *d_pbase = nullptr;
    const TileMap *d_stm = nullptr;
    std::vector<TileMap> tm;
    int64_t ntot = 0;
    int ntiles = 0, nstiles = 0;
    bool any_out = false;
};

PassTables pass_tables(s2g_ctx *c, const MapIn *mi, int nmaps, hipStream_t s)
{
    const size_t btab = sizeof(Map2D) * nmaps, bpb = sizeof(int64_t) * (nmaps + 1);
    const size_t bst = sizeof(TileMap) * nmaps;
    char *h = (char *)c->h_tab.get(btab + bpb + bst);
    Map2D *tab = (Map2D *)h;
    int64_t *pb = (int64_t *)(h + btab);
    TileMap *stm = (TileMap *)(h + btab + bpb);
    PassTables pt;
    pt.tm.resize(nmaps);
    int64_t ntot = 0, ttot = 0, stot = 0;
    for (int m = 0; m < nmaps; ++m) {
        const Map2DArgs &a = mi[m].a;
        Map2D &e = tab[m];
        e.pos = a.pos;
        e.hsml = a.hsml;
        e.m = a.m;
        e.rho = a.rho;
        e.binq = a.binq;
        e.w = a.w;
        e.image = mi[m].image;
        e.out = mi[m].out;
        e.len2pix = a.len2pix;
        e.npix = a.npix;
        e.n = a.n;
        e.pbase = ntot;
        e.flags = a.flags;
        e.small_side = c->tune.small_side;
        const int ntj = (int)((a.npix + kGT - 1) / kGT);
        e.ntj = ntj;
        e.tbase = (int32_t)ttot;
        const int nts = (int)((a.npix + kST - 1) / kST);
        e.nts = nts;
        e.sbase = (int32_t)stot;
        pb[m] = ntot;
        pt.tm[m] = TileMap{0, (int32_t)ttot, ntj * ntj, ntj, 1, 0, 0};
        // bucket t of the small path spans sorted entries [off[t], off[t+1])
        stm[m] = TileMap{stot, (int32_t)stot, nts * nts, nts, 1, 1, 0};
        ntot += a.n;
        ttot += (int64_t)ntj * ntj;
        stot += (int64_t)nts * nts;
        pt.any_out = pt.any_out || mi[m].out;
        if (ntot > (int64_t)INT32_MAX || stot > (int64_t)INT32_MAX)
            throw Error(S2G_EUNSUPPORTED, "deposit pass too large (2^31 particles or tiles)");
    }
    pb[nmaps] = ntot;
    char *d = (char *)c->tab.get(btab + bpb + bst);
    pt.d_maps = (const Map2D *)d;
    pt.d_pbase = (const int64_t *)(d + btab);
    pt.d_stm = (const TileMap *)(d + btab + bpb);
    pt.ntot = ntot;
    pt.ntiles = (int)ttot;
    pt.nstiles = (int)stot;
    S2G_HIP(hipMemcpyAsync(d, h, btab + bpb + bst, hipMemcpyHostToDevice, s));
    return pt;
}

// The small list bucketed by the kST x kST tile holding each footprint's
// first pixel: a stable radix sort of (tile key, particle), so particles stay
// in particle order inside a tile.  Returns the sorted particles; off[t] is
// the first entry of tile t (t in [0, nstiles]).
const int32_t *bucket_small(s2g_ctx *c, const ClassLists &cl, const Rec2 *rec,
                            const Map2D *d_maps, int nstiles, int64_t **off, hipStream_t s)
{
    const int64_t ns = cl.ns;
    uint32_t *keys = (uint32_t *)c->skeys.get(sizeof(uint32_t) * 2 * ns);
    int32_t *vals = (int32_t *)c->svals.get(sizeof(int32_t) * 2 * ns);
    launch_2d_small_keys(cl.l0, ns, rec, d_maps, keys, vals, s);
    int end_bit = 1;
    while (end_bit < 32 && ((int64_t)1 << end_bit) < (int64_t)nstiles)
        ++end_bit;
    const size_t tb = sort_pairs_tmp_bytes(ns, end_bit);
    void *stmp = c->sort_tmp.get(tb ? tb : 1);
    sort_pairs(stmp, tb, keys, keys + ns, vals, vals + ns, ns, end_bit, s);
    *off = (int64_t *)c->soff.get(sizeof(int64_t) * (nstiles + 1));
    launch_tile_offsets(keys + ns, ns, nstiles, *off, s);
    return vals + ns;
}

void deposit2d_pass(s2g_ctx *c, const MapIn *mi, int nmaps, hipStream_t s, s2g_stats *st)
{
    const int dtype = mi[0].a.dtype, kernel = mi[0].a.kernel, nq = mi[0].a.nq;
    const int nplanes = nq + 1;
    // FP32 per-item sums in the gather (all maps of a pass share the flag)
    const int fp32acc = (mi[0].a.flags & S2G_FLAG_FP32_ACCUM) ? 1 : 0;
    if (st) {
        std::memset(st, 0, sizeof(*st));
    }
    hipEvent_t *E = c->stage;
    S2G_HIP(hipEventRecord(E[0], s));
    PassTables pt = pass_tables(c, mi, nmaps, s);
    const Map2D *d_maps = pt.d_maps;
    const int64_t *d_pbase = pt.d_pbase;
    const TileMap *d_stm = pt.d_stm;
    std::vector<TileMap> &tm = pt.tm;
    const int64_t ntot = pt.ntot;
    const int ntiles = pt.ntiles, nstiles = pt.nstiles;
    const bool any_out = pt.any_out;
    if (st)
        st->n_particles = ntot;

    ClassLists cl;
    Rec2 *rec = nullptr;
    if (ntot > 0) {
        rec = (Rec2 *)c->rec.get(sizeof(Rec2) * ntot);
        int32_t *cls = (int32_t *)c->cls.get(sizeof(int32_t) * ntot);
        launch_2d_setup(dtype, d_maps, nmaps, ntot, nq, kernel, rec, cls, s);
        cl = partition_classes(c, cls, ntot, d_pbase, nmaps, s);
        fill_class_stats(st, ntot, cl);
    } else {
        S2G_HIP(hipEventRecord(E[1], s));
    }

    // Side stream: the explicit normalisation, then the small-footprint
    // scatter, beside the pair lists and the gather on s.  The gather waits
    // for the normalisation (it reads the records' scale); the small scatter
    // adds into images the side stream zeroed first, and the first finalize
    // waits for it and adds the large particles' sums (2^24 cluster set on
    // 4096^2: the 5.7 ms normalisation runs beside the 5.7 ms pair lists and
    // the 6.4 ms small scatter beside the gather; 288 -> 284 ms per step).
    hipStream_t s2 = c->side;
    S2G_HIP(hipEventRecord(c->fork, s));
    S2G_HIP(hipStreamWaitEvent(s2, c->fork, 0));
    S2G_HIP(hipEventRecord(E[2], s2));
    launch_2d_norm(kernel, cl.l2, cl.ne, d_maps, rec, s2);
    S2G_HIP(hipEventRecord(E[3], s2));
    S2G_HIP(hipEventRecord(c->normed, s2));
    const int eS = s2g_ctx::kNumStageEvents - 3;
    S2G_HIP(hipEventRecord(E[eS], s2));
    const bool add = cl.ns > 0;  // finalize adds to the small scatter's image
    if (add) {
        for (int m = 0; m < nmaps; ++m) {
            const int64_t nd = mi[m].a.npix * mi[m].a.npix;
            S2G_HIP(hipMemsetAsync(mi[m].image, 0, sizeof(double) * nplanes * nd, s2));
        }
        const int64_t ns = cl.ns;
        int64_t *soff = nullptr;
        const int32_t *sorted = bucket_small(c, cl, rec, d_maps, nstiles, &soff, s2);
        // work items of <= Gs particles; the item count stays on the device and
        // the grid is its upper bound nstiles + ns / Gs
        const int64_t starget = (int64_t)c->ncu * c->tune.small_items_per_cu;
        int64_t Gs = (ns + starget - 1) / starget;
        if (Gs < c->tune.small_item_min)
            Gs = c->tune.small_item_min;
        const int64_t max_items = nstiles + (ns + Gs - 1) / Gs;
        if (max_items > ((int64_t)1 << 31) / 1024)  // blocks x threads (up to 1024)
            throw Error(S2G_EUNSUPPORTED, "too many small-footprint work items");
        int64_t *snit = (int64_t *)c->snitems.get(sizeof(int64_t) * (nstiles + 1));
        int4 *sitems = (int4 *)c->sitems.get(sizeof(int4) * max_items);
        int64_t *stile = (int64_t *)c->stile.get(sizeof(int64_t) * 2 * nstiles);
        int64_t *tmp =
            (int64_t *)c->sscan_tmp.get(sizeof(int64_t) * scan_tmp_elems(nstiles + 1));
        launch_items(soff, d_stm, nmaps, nstiles, Gs, 1, nullptr, snit, tmp, nullptr, nullptr,
                     false, s2);
        launch_items(soff, d_stm, nmaps, nstiles, Gs, 1, nullptr, snit, tmp, sitems, stile,
                     true, s2);
        const int sstep = c->tune.small_planes;
        for (int plane0 = 0; plane0 < nplanes; plane0 += sstep) {
            const int np = nplanes - plane0 < sstep ? nplanes - plane0 : sstep;
            launch_2d_small_tile(dtype, kernel, np, sitems, snit + nstiles, max_items, sorted,
                                 rec, d_maps, nq, plane0, s2);
        }
    }
    S2G_HIP(hipEventRecord(E[eS + 1], s2));
    S2G_HIP(hipEventRecord(c->smalled, s2));

    double ms_sort = 0.0, ms_gather = 0.0, ms_fin = 0.0;
    int64_t launches = 0;
    TilePlan tp;
    if (cl.nl > 0) {
        tp = build_tile_lists(c, cl.l1, rec, false, tm, cl.lbase, cl.d_lbase, ntiles, s,
                              &ms_sort, c->tune.gather_item_min, c->tune.gather_items_per_cu,
                              c->tune.gather_item_max);
        if (st)
            st->n_pairs = tp.npairs;
    }
    S2G_HIP(hipStreamWaitEvent(s, c->normed, 0));
    int ev = 10;
    const int gstep = c->tune.gather_planes;
    for (int plane0 = 0; plane0 < nplanes; plane0 += gstep) {
        const int np = nplanes - plane0 < gstep ? nplanes - plane0 : gstep;
        GatherArgs g;
        g.dtype = dtype;
        g.kernel = kernel;
        g.fp32acc = fp32acc;
        g.np = np;
        g.plane0 = plane0;
        g.nq = nq;
        g.nmaps = nmaps;
        g.ntiles = ntiles;
        g.nitems = tp.nitems;
        g.items = tp.items;
        g.pairs = tp.pairs;
        g.rec = rec;
        g.maps = d_maps;
        g.partial = tp.nitems > 0 ? (double *)c->partial.get(sizeof(double) * tp.nitems * np *
                                                             kGT * kGT)
                                  : nullptr;
        g.tile_items = tp.tile_items;
        g.add = add;
        // events per plane group: gather start / end, finalize start (after
        // the join with the side stream) / end
        const bool timed = ev + 3 < s2g_ctx::kNumStageEvents - 3;
        if (timed)
            S2G_HIP(hipEventRecord(E[ev], s));
        if (tp.nitems > 0) {
            launch_2d_gather(g, s);
            ++launches;
        }
        if (timed)
            S2G_HIP(hipEventRecord(E[ev + 1], s));
        if (plane0 == 0)
            S2G_HIP(hipStreamWaitEvent(s, c->smalled, 0));
        if (timed)
            S2G_HIP(hipEventRecord(E[ev + 2], s));
        launch_2d_finalize(g, s);
        if (timed) {
            S2G_HIP(hipEventRecord(E[ev + 3], s));
            ev += 4;
        }
    }

    if (any_out)
        launch_2d_reduce_maps(d_maps, nmaps, ntiles, nq, s);
    S2G_HIP(hipGetLastError());
    S2G_HIP(hipEventRecord(E[eS + 2], s));
    S2G_HIP(hipEventSynchronize(E[eS + 2]));
    if (st) {
        for (int k = 10; k < ev; k += 4) {
            ms_gather += elapsed(E[k], E[k + 1]);
            ms_fin += elapsed(E[k + 2], E[k + 3]);
        }
        st->ms_device = elapsed(E[0], E[eS + 2]);
        st->ms_setup = elapsed(E[0], E[1]);
        st->ms_norm = elapsed(E[2], E[3]);
        st->ms_sort = ms_sort;
        st->ms_gather = ms_gather;
        st->ms_finalize = ms_fin;
        st->ms_small = elapsed(E[eS], E[eS + 1]);
        st->n_gather_launches = launches;
    }
}

// one map, flat image only
void deposit2d_device(s2g_ctx *c, const Map2DArgs &a, double *d_image, hipStream_t s,
                      s2g_stats *st)
{
    MapIn m{a, d_image, nullptr};
    deposit2d_pass(c, &m, 1, s, st);
}

// ---------------------------------------------------------------------------
// 3D deposit on device pointers (image: npix^3 x 2, plane 0 = q, 1 = weights)
// ---------------------------------------------------------------------------
void deposit3d_device(s2g_ctx *c, const Map3DArgs &a, double *d_image, hipStream_t s,
                      s2g_stats *st)
{
    const int64_t n = a.n, npix = a.npix;
    const int64_t nd = npix * npix * npix;
    if (st) {
        std::memset(st, 0, sizeof(*st));
        st->n_particles = n;
    }
    hipEvent_t *E = c->stage;
    S2G_HIP(hipEventRecord(E[0], s));
    if (n == 0) {
        S2G_HIP(hipMemsetAsync(d_image, 0, sizeof(double) * nd * 2, s));
        S2G_HIP(hipEventRecord(E[1], s));
        S2G_HIP(hipEventSynchronize(E[1]));
        return;
    }
    Rec3 *rec = (Rec3 *)c->rec.get(sizeof(Rec3) * n);
    int32_t *cls = (int32_t *)c->cls.get(sizeof(int32_t) * n);
    int64_t *hpb = (int64_t *)c->h_tab.get(2 * sizeof(int64_t));
    hpb[0] = 0;
    hpb[1] = n;
    int64_t *d_pbase = (int64_t *)c->tab.get(2 * sizeof(int64_t));
    S2G_HIP(hipMemcpyAsync(d_pbase, hpb, 2 * sizeof(int64_t), hipMemcpyHostToDevice, s));
    launch_3d_setup(a, rec, cls, s);
    const ClassLists cl = partition_classes(c, cls, n, d_pbase, 1, s);
    fill_class_stats(st, n, cl);

    // Side stream, as in the 2D pass: the explicit normalisation, then the
    // small-footprint scatter into the zeroed image, beside the pair lists
    // and the gather; the gather waits for the normalisation and the
    // finalize for the scatter, to whose sums it adds.
    hipStream_t s2 = c->side;
    S2G_HIP(hipEventRecord(c->fork, s));
    S2G_HIP(hipStreamWaitEvent(s2, c->fork, 0));
    S2G_HIP(hipEventRecord(E[2], s2));
    launch_3d_norm(a.kernel, cl.l2, cl.ne, a.npix, a.flags, rec, s2);
    S2G_HIP(hipEventRecord(E[3], s2));
    S2G_HIP(hipEventRecord(c->normed, s2));
    const int eS = s2g_ctx::kNumStageEvents - 3;
    S2G_HIP(hipMemsetAsync(d_image, 0, sizeof(double) * nd * 2, s2));
    S2G_HIP(hipEventRecord(E[eS], s2));
    Small3Args sa;
    sa.dtype = a.dtype;
    sa.kernel = a.kernel;
    sa.ns = cl.ns;
    sa.npix = npix;
    sa.list = cl.l0;
    sa.rec = rec;
    sa.binq = a.binq;
    sa.image = d_image;
    launch_3d_small(sa, s2);
    S2G_HIP(hipEventRecord(E[eS + 1], s2));
    S2G_HIP(hipEventRecord(c->smalled, s2));

    const int nb = (int)((npix + kBT - 1) / kBT);
    const int ntiles = nb * nb * nb;
    double ms_sort = 0.0, ms_gather = 0.0, ms_fin = 0.0;
    int64_t launches = 0;
    if (cl.nl > 0) {
        std::vector<TileMap> tm{TileMap{0, 0, ntiles, nb, nb, 0, 0}};
        // brick items of at most 4096 pairs (2^24 cluster set on 256^3: 16
        // items per CU of about 20k pairs, gather 69.0 ms; capped at 1024 /
        // 4096 / 8192 pairs: 60.2 / 60.9 / 63.9 ms, 1024 with a longer
        // finalize)
        const TilePlan tp = build_tile_lists(c, cl.l1, rec, true, tm, cl.lbase, cl.d_lbase,
                                             ntiles, s, &ms_sort, 64, 16, 4096);
        if (st)
            st->n_pairs = tp.npairs;
        Gather3Args g;
        g.dtype = a.dtype;
        g.kernel = a.kernel;
        g.np = 2;
        g.plane0 = 0;
        g.nq = 1;
        g.nb = nb;
        g.ntiles = ntiles;
        g.nitems = tp.nitems;
        g.npix = npix;
        g.items = tp.items;
        g.pairs = tp.pairs;
        g.rec = rec;
        g.binq = a.binq;
        g.partial = (double *)c->partial.get(sizeof(double) * (tp.nitems ? tp.nitems : 1) * 2 *
                                             kBT * kBT * kBT);
        g.tile_items = tp.tile_items;
        g.image = d_image;
        g.add = 1;
        S2G_HIP(hipStreamWaitEvent(s, c->normed, 0));
        S2G_HIP(hipEventRecord(E[10], s));
        if (tp.nitems > 0) {
            launch_3d_gather(g, s);
            ++launches;
        }
        S2G_HIP(hipEventRecord(E[11], s));
        S2G_HIP(hipStreamWaitEvent(s, c->smalled, 0));
        S2G_HIP(hipEventRecord(E[12], s));
        launch_3d_finalize(g, s);
        S2G_HIP(hipEventRecord(E[13], s));
    }
    S2G_HIP(hipStreamWaitEvent(s, c->smalled, 0));
    S2G_HIP(hipGetLastError());
    S2G_HIP(hipEventRecord(E[eS + 2], s));
    S2G_HIP(hipEventSynchronize(E[eS + 2]));
    if (st) {
        if (cl.nl > 0) {
            ms_gather = elapsed(E[10], E[11]);
            ms_fin = elapsed(E[12], E[13]);
        }
        st->ms_device = elapsed(E[0], E[eS + 2]);
        st->ms_setup = elapsed(E[0], E[1]);
        st->ms_norm = elapsed(E[2], E[3]);
        st->ms_sort = ms_sort;
        st->ms_gather = ms_gather;
        st->ms_finalize = ms_fin;
        st->ms_small = elapsed(E[eS], E[eS + 1]);
        st->n_gather_launches = launches;
    }
}

Map2DArgs make_args(int dtype, const void *pos, const void *hsml, const void *m,
                    const void *rho, const void *bin_q, int nq, const void *weights, int64_t n,
                    const s2g_params *params, int kernel, uint32_t flags)
{
    Map2DArgs a;
    a.dtype = dtype;
    a.pos = pos;
    a.hsml = hsml;
    a.m = m;
    a.rho = rho;
    a.binq = bin_q;
    a.w = weights;
    a.nq = nq;
    a.n = n;
    a.npix = params->npixels;
    a.len2pix = params->len2pix;
    a.kernel = kernel;
    a.flags = flags;
    return a;
}

// copy host inputs into the context's staging buffers
Map2DArgs stage_inputs(s2g_ctx *c, const Map2DArgs &h, hipStream_t s, s2g_stats *st)
{
    const size_t es = elem_size(h.dtype);
    const int64_t n = h.n;
    const void *src[6] = {h.pos, h.hsml, h.m, h.rho, h.binq, h.w};
    const size_t cnt[6] = {(size_t)(3 * n), (size_t)n, (size_t)n, (size_t)n,
                           (size_t)(n * h.nq), (size_t)n};
    void *dst[6];
    hipEvent_t e0 = c->ev0, e1 = c->ev1;
    S2G_HIP(hipEventRecord(e0, s));
    for (int k = 0; k < 6; ++k) {
        if (n > 0 && !src[k])
            throw Error(S2G_EINVAL, "input array pointer is NULL");
        dst[k] = c->in[k].get(es * (cnt[k] ? cnt[k] : 1));
        if (n > 0)
            S2G_HIP(hipMemcpyAsync(dst[k], src[k], es * cnt[k], hipMemcpyHostToDevice, s));
    }
    S2G_HIP(hipEventRecord(e1, s));
    S2G_HIP(hipEventSynchronize(e1));
    if (st) {
        float ms = 0.f;
        S2G_HIP(hipEventElapsedTime(&ms, e0, e1));
        st->ms_h2d = ms;
    }
    Map2DArgs d = h;
    d.pos = dst[0];
    d.hsml = dst[1];
    d.m = dst[2];
    d.rho = dst[3];
    d.binq = dst[4];
    d.w = dst[5];
    return d;
}

void d2h(s2g_ctx *c, void *dst, const void *src, size_t bytes, hipStream_t s, double *ms)
{
    S2G_HIP(hipEventRecord(c->ev0, s));
    S2G_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    S2G_HIP(hipEventRecord(c->ev1, s));
    S2G_HIP(hipEventSynchronize(c->ev1));
    if (ms) {
        float t = 0.f;
        S2G_HIP(hipEventElapsedTime(&t, c->ev0, c->ev1));
        *ms = t;
    }
}

void check_dims_nq(int dims, int nq)
{
    if (dims != 2 && dims != 3)
        throw Error(S2G_EINVAL, "dims must be 2 or 3");
    if (nq < 1 || nq > 64)
        throw Error(S2G_EINVAL, "nq must be in [1, 64]");
    if (dims == 3 && nq != 1)
        throw Error(S2G_EINVAL, "the 3D deposit maps a scalar quantity (nq == 1)");
}

void deposit_device_any(s2g_ctx *c, int dims, const Map2DArgs &a, double *d_image,
                        hipStream_t s, s2g_stats *st)
{
    if (dims == 2)
        deposit2d_device(c, a, d_image, s, st);
    else
        deposit3d_device(c, a, d_image, s, st);
}

// ---------------------------------------------------------------------------
// mass_density on device pointers
// ---------------------------------------------------------------------------
void mass_density_device(s2g_ctx *c, int dtype, const void *d_pos, const void *d_mass, int64_t n,
                         int kernel, int dim, int64_t nneigh, const double *boxsize,
                         double *d_rho, double *d_hsml, hipStream_t s, s2g_stats *st)
{
    if (dtype != S2G_F32 && dtype != S2G_F64)
        throw Error(S2G_EINVAL, "dtype must be S2G_F32 or S2G_F64");
    if (kernel < S2G_KERNEL_CUBIC || kernel > S2G_KERNEL_WENDLAND_C8)
        throw Error(S2G_EUNSUPPORTED, "unsupported kernel id " + std::to_string(kernel));
    if (dim != 2 && dim != 3)
        throw Error(S2G_EINVAL, "kernel_dim must be 2 or 3");
    if (nneigh < 1 || nneigh > md_max_neighbours())
        throw Error(S2G_EUNSUPPORTED, "Nneighbors must be in [1, " +
                                          std::to_string(md_max_neighbours()) + "]");
    if (n < nneigh + 1 || n > (int64_t)INT32_MAX)
        throw Error(S2G_EINVAL, "need at least Nneighbors + 1 particles (and fewer than 2^31)");
    if (!d_pos || !d_mass || !d_rho || !d_hsml)
        throw Error(S2G_EINVAL, "array pointer is NULL");
    MdGrid g;
    int npos = 0;
    for (int d = 0; d < 3; ++d) {
        const double b = boxsize ? boxsize[d] : 0.0;
        if (!(b >= 0.0) || !std::isfinite(b))
            throw Error(S2G_EINVAL, "boxsize must be finite and >= 0");
        g.box[d] = b;
        npos += b > 0.0;
    }
    if (npos != 0 && npos != 3)
        throw Error(S2G_EINVAL, "boxsize must be zero (non-periodic) or positive in all "
                                "three dimensions");
    g.periodic = npos == 3;
    if (st) {
        std::memset(st, 0, sizeof(*st));
        st->n_particles = n;
    }
    hipEvent_t *E = c->stage;
    S2G_HIP(hipEventRecord(E[0], s));
    const int64_t nb = md_prep_blocks(n);
    double4 *P = (double4 *)c->md_P.get(sizeof(double4) * n);
    double *bbox = (double *)c->md_bbox.get(sizeof(double) * 6 * nb);
    launch_md_prep(dtype, d_pos, d_mass, n, g, P, bbox, s);
    // key domain: the periodic box, or the bounding box of the particles
    double *h = (double *)c->h_bounds.get(sizeof(double) * 6 * nb);
    S2G_HIP(hipMemcpyAsync(h, bbox, sizeof(double) * 6 * nb, hipMemcpyDeviceToHost, s));
    S2G_HIP(hipStreamSynchronize(s));
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t b = 0; b < nb; ++b)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::fmin(lo[d], h[6 * b + d]);
            hi[d] = std::fmax(hi[d], h[6 * b + 3 + d]);
        }
    for (int d = 0; d < 3; ++d)
        if (!(lo[d] > -1e300) || !(hi[d] < 1e300))
            throw Error(S2G_EINVAL, "positions must be finite");
    if (g.periodic) {
        for (int d = 0; d < 3; ++d) {
            g.lo[d] = 0.0;
            g.ext[d] = g.box[d];
        }
    } else {
        double emax = 0.0;
        for (int d = 0; d < 3; ++d)
            emax = std::fmax(emax, hi[d] - lo[d]);
        if (!(emax > 0.0))
            emax = 1.0;
        for (int d = 0; d < 3; ++d) {
            g.lo[d] = lo[d];
            const double e = hi[d] - lo[d];
            // a little wider than the particles, so the largest maps inside
            g.ext[d] = (e > 0.0 ? e : emax) * (1.0 + 1e-9);
        }
    }
    g.min_ext = 1e300;
    for (int d = 0; d < 3; ++d) {
        g.inv_ext[d] = 1.0 / g.ext[d];
        g.min_ext = std::fmin(g.min_ext, g.ext[d]);
    }
    uint64_t *keys = (uint64_t *)c->skeys.get(sizeof(uint64_t) * 2 * n);
    int32_t *vals = (int32_t *)c->svals.get(sizeof(int32_t) * 2 * n);
    launch_md_keys(P, n, g, keys, vals, s);
    const int end_bit = 3 * md_bits();
    const size_t tb = sort_pairs64_tmp_bytes(n, end_bit);
    void *stmp = c->sort_tmp.get(tb ? tb : 1);
    sort_pairs64(stmp, tb, keys, keys + n, vals, vals + n, n, end_bit, s);
    // offset table at the level with about one particle per cell (<= 8)
    int lt = 1;
    while (lt < 8 && ((int64_t)1 << (3 * lt)) < n)
        ++lt;
    int32_t *tab = (int32_t *)c->md_tab.get(sizeof(int32_t) * (((size_t)1 << (3 * lt)) + 1));
    launch_md_table(keys + n, n, lt, tab, s);
    double4 *S = (double4 *)c->md_S.get(sizeof(double4) * n);
    launch_md_sorted(P, vals + n, n, S, s);
    S2G_HIP(hipEventRecord(E[1], s));
    const double norm = dim == 3 ? 1.0 / shape_ball_integral(kernel)
                                 : 1.0 / shape_disc_integral(kernel);
    launch_md_knn(kernel, S, keys + n, vals + n, tab, lt, n, g, (int)nneigh + 1, dim,
                  (int)nneigh, norm, d_rho, d_hsml, s);
    S2G_HIP(hipGetLastError());
    S2G_HIP(hipEventRecord(E[2], s));
    S2G_HIP(hipEventSynchronize(E[2]));
    if (st) {
        st->ms_device = elapsed(E[0], E[2]);
        st->ms_sort = elapsed(E[0], E[1]);
        st->ms_gather = elapsed(E[1], E[2]);
    }
}

// ---------------------------------------------------------------------------
// Batches: argument checks, pass groups, per-pass inputs.
// ---------------------------------------------------------------------------
void check_batch(int dtype, const s2g_map_desc *maps, int nmaps)
{
    if (nmaps < 0 || (nmaps > 0 && !maps))
        throw Error(S2G_EINVAL, "maps is NULL or nmaps < 0");
    if (dtype != S2G_F32 && dtype != S2G_F64)
        throw Error(S2G_EINVAL, "dtype must be S2G_F32 or S2G_F64");
    for (int k = 0; k < nmaps; ++k) {
        const s2g_map_desc &d = maps[k];
        const std::string at = "map " + std::to_string(k) + ": ";
        if (d.dims != 2)
            throw Error(S2G_EUNSUPPORTED, at + "batched deposits are 2D (dims == 2)");
        try {
            check_common(dtype, d.kernel, 2, d.n, &d.params);
            check_dims_nq(2, d.nq);
        } catch (const Error &e) {
            throw Error(e.code, at + e.msg);
        }
        if (!d.d_out)
            throw Error(S2G_EINVAL, at + "output pointer is NULL");
        if (d.n > 0 && (!d.d_pos || !d.d_hsml || !d.d_m || !d.d_rho || !d.d_bin_q ||
                        !d.d_weights))
            throw Error(S2G_EINVAL, at + "input array pointer is NULL");
    }
}

// one pass per (kernel, nq, accumulator) group, maps in their given order
std::vector<std::vector<int>> batch_groups(const s2g_map_desc *maps, int nmaps)
{
    std::vector<std::vector<int>> groups;
    std::vector<char> done(nmaps, 0);
    for (int k = 0; k < nmaps; ++k) {
        if (done[k])
            continue;
        std::vector<int> grp;
        for (int j = k; j < nmaps; ++j)
            if (!done[j] && maps[j].kernel == maps[k].kernel && maps[j].nq == maps[k].nq &&
                (maps[j].flags & S2G_FLAG_FP32_ACCUM) == (maps[k].flags & S2G_FLAG_FP32_ACCUM)) {
                grp.push_back(j);
                done[j] = 1;
            }
        groups.push_back(grp);
    }
    return groups;
}

// doubles of the flat images of a group's reduced maps
int64_t batch_flat_elems(const s2g_map_desc *maps, const std::vector<int> &grp)
{
    int64_t flat = 0;
    for (int j : grp)
        if (maps[j].flags & S2G_FLAG_REDUCE)
            flat += maps[j].params.npixels * maps[j].params.npixels * (maps[j].nq + 1);
    return flat;
}

// the group's pass inputs; reduced maps deposit into img (batch_flat_elems)
std::vector<MapIn> batch_inputs(int dtype, const s2g_map_desc *maps, const std::vector<int> &grp,
                                double *img)
{
    std::vector<MapIn> in;
    int64_t off = 0;
    for (int j : grp) {
        const s2g_map_desc &d = maps[j];
        MapIn m;
        m.a = make_args(dtype, d.d_pos, d.d_hsml, d.d_m, d.d_rho, d.d_bin_q, d.nq, d.d_weights,
                        d.n, &d.params, d.kernel, d.flags);
        if (d.flags & S2G_FLAG_REDUCE) {
            m.image = img + off;
            m.out = d.d_out;
            off += d.params.npixels * d.params.npixels * (d.nq + 1);
        } else {
            m.image = d.d_out;
            m.out = nullptr;
        }
        in.push_back(m);
    }
    return in;
}

}  // namespace

// ---------------------------------------------------------------------------
// Replayable batch (s2g_plan_*): the passes of s2g_deposit_batch with every
// size fixed at creation and every count kept on the device, captured once
// in a hipGraph and replayed.  Differences from the eager pass, all to keep
// the host out of the replay:
//  * the explicit-norm and gather grids are upper bounds, the counts read on
//    the device (class bounds, item totals);
//  * the pair-list chunks are ranges of each map's large list computed on the
//    device from the class bounds (nch chunks of csz entries per map, sized
//    from the creation run's large count times the headroom);
//  * pairs have a capacity (the creation run's total times the headroom);
//    k_plan_check raises a flag when either capacity is passed, the pair
//    fill writes nothing past it and the gather does not run;
//  * small-footprint buckets are per-tile lists built like the pair lists
//    (LDS counts per chunk of each map's small list, a scan, LDS cursors):
//    kernels only, no library sort and no memset node.
// ---------------------------------------------------------------------------
struct PlanPass {
    std::vector<MapIn> mi;
    int dtype = 0, kernel = 0, nq = 0, nplanes = 0, fp32acc = 0, nmaps = 0;
    int64_t ntot = 0, nb = 0, ncnt = 0, sncnt = 0, pair_cap = 0, item_cap = 0, sitem_cap = 0;
    int ntiles = 0, nstiles = 0, max_tiles = 1, smax_tiles = 1, nchunks = 0, nschunks = 0;
    int gstep = 4, sstep = 4;
    int64_t gmin = 64, gtarget = 1, gmax = 0, smin = 256, starget = 1;
    bool any_out = false;
    DevBuf tab, img, rec, cls, counts, scan_tmp, l0, l1, l2, bounds, tiles, cnt, pairs, nitems,
        items, tile_items, partial, stiles, scnt, sorted, snitems, sitems, stile, flag, sscan_tmp;
    const Map2D *d_maps = nullptr;
    const int64_t *d_pbase = nullptr;
    const TileMap *d_tm = nullptr, *d_stm = nullptr;  // large / small lists' tile grids
    const PairChunk *d_ch = nullptr, *d_sch = nullptr;
    void release()
    {
        for (DevBuf *b : {&tab, &img, &rec, &cls, &counts, &scan_tmp, &l0, &l1, &l2, &bounds,
                          &tiles, &cnt, &pairs, &nitems, &items, &tile_items, &partial, &stiles,
                          &scnt, &sorted, &snitems, &sitems, &stile, &flag, &sscan_tmp})
            b->release();
    }
};

struct s2g_plan {
    s2g_ctx *c = nullptr;
    std::vector<PlanPass> passes;
    hipStream_t cs = nullptr;   // capture stream
    hipStream_t cs2 = nullptr;  // its side branch (joined back before the capture ends)
    hipEvent_t fork = nullptr, normed = nullptr, smalled = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int64_t *h_read = nullptr;  // pinned: per pass flag + counts after a run
};

namespace {

void plan_release(s2g_plan *pl)
{
    if (pl->exec)
        (void)hipGraphExecDestroy(pl->exec);
    if (pl->graph)
        (void)hipGraphDestroy(pl->graph);
    for (hipStream_t q : {pl->cs, pl->cs2})
        if (q)
            (void)hipStreamDestroy(q);
    for (hipEvent_t e : {pl->fork, pl->normed, pl->smalled})
        if (e)
            (void)hipEventDestroy(e);
    if (pl->h_read)
        (void)hipHostFree(pl->h_read);
    for (PlanPass &pp : pl->passes)
        pp.release();
    pl->exec = nullptr;
    pl->graph = nullptr;
    pl->cs = pl->cs2 = nullptr;
    pl->fork = pl->normed = pl->smalled = nullptr;
    pl->h_read = nullptr;
}

// Everything of one pass on `s` and its side branch `s2` (fork / join by
// the plan's events), sized by the pass's capacities, with no host read-back
// (the sequence that is captured).  As in the eager pass, the explicit
// normalisation and the small-footprint scatter run on the side branch; the
// images are zeroed there by the finalize kernel with no items (no memset
// node in the graph), and the finalize adds the large particles' sums.
void plan_enqueue(s2g_plan *pl, PlanPass &pp, hipStream_t s, hipStream_t s2)
{
    Rec2 *rec = (Rec2 *)pp.rec.p;
    int32_t *cls = (int32_t *)pp.cls.p, *l0 = (int32_t *)pp.l0.p, *l1 = (int32_t *)pp.l1.p,
            *l2 = (int32_t *)pp.l2.p;
    int64_t *tmp = (int64_t *)pp.scan_tmp.p, *bounds = (int64_t *)pp.bounds.p;
    launch_2d_setup(pp.dtype, pp.d_maps, pp.nmaps, pp.ntot, pp.nq, pp.kernel, rec, cls, s);
    launch_partition(cls, pp.ntot, (int64_t *)pp.counts.p, tmp, l0, l1, l2, s);
    launch_class_bounds((const int64_t *)pp.counts.p, pp.nb, l0, l1, pp.d_pbase, pp.nmaps, bounds,
                        s);
    S2G_HIP(hipEventRecord(pl->fork, s));
    S2G_HIP(hipStreamWaitEvent(s2, pl->fork, 0));

    // side branch: normalisation, zeroed images, small-footprint lists and scatter
    launch_2d_norm(pp.kernel, l2, pp.ntot, pp.d_maps, rec, s2, bounds);
    S2G_HIP(hipEventRecord(pl->normed, s2));
    for (int plane0 = 0; plane0 < pp.nplanes; plane0 += 4) {
        GatherArgs z;
        z.np = pp.nplanes - plane0 < 4 ? pp.nplanes - plane0 : 4;
        z.plane0 = plane0;
        z.nmaps = pp.nmaps;
        z.ntiles = pp.ntiles;
        z.maps = pp.d_maps;
        z.partial = nullptr;
        z.tile_items = nullptr;  // no items: every pixel of every tile set to 0
        launch_2d_finalize(z, s2);
    }
    int64_t *stmp = (int64_t *)pp.sscan_tmp.p;
    int64_t *scnt = (int64_t *)pp.scnt.p, *snit = (int64_t *)pp.snitems.p;
    // per-tile lists over device-side chunk ranges of each map's small list
    // (bounds[5 + nmaps + m]), then items (bounds[1] = the small count)
    const int64_t *sbase = bounds + 5 + pp.nmaps;
    launch_anchor_lists(l0, rec, pp.d_sch, pp.nschunks, pp.d_stm, pp.smax_tiles, scnt, pp.sncnt,
                        stmp, nullptr, false, sbase, s2);
    launch_anchor_lists(l0, rec, pp.d_sch, pp.nschunks, pp.d_stm, pp.smax_tiles, scnt, pp.sncnt,
                        stmp, (int32_t *)pp.sorted.p, true, sbase, s2);
    launch_items(scnt, pp.d_stm, pp.nmaps, pp.nstiles, pp.smin, pp.starget, bounds + 1, snit,
                 stmp, nullptr, nullptr, false, s2);
    launch_items(scnt, pp.d_stm, pp.nmaps, pp.nstiles, pp.smin, pp.starget, bounds + 1, snit,
                 stmp, (int4 *)pp.sitems.p, (int64_t *)pp.stile.p, true, s2);
    for (int plane0 = 0; plane0 < pp.nplanes; plane0 += pp.sstep) {
        const int np = pp.nplanes - plane0 < pp.sstep ? pp.nplanes - plane0 : pp.sstep;
        launch_2d_small_tile(pp.dtype, pp.kernel, np, (const int4 *)pp.sitems.p, snit + pp.nstiles,
                             pp.sitem_cap, (const int32_t *)pp.sorted.p, rec, pp.d_maps, pp.nq,
                             plane0, s2);
    }
    S2G_HIP(hipEventRecord(pl->smalled, s2));

    // large particles: pair lists over device-side chunk ranges
    int64_t *cnt = (int64_t *)pp.cnt.p, *nit = (int64_t *)pp.nitems.p;
    const int64_t *lbase = bounds + 4, *d_npairs = cnt + pp.ncnt - 1;
    launch_pairs<Rec2>(l1, rec, pp.d_ch, pp.nchunks, pp.d_tm, pp.max_tiles, cnt, pp.ncnt, tmp,
                       nullptr, false, s, lbase);
    launch_items(cnt, pp.d_tm, pp.nmaps, pp.ntiles, pp.gmin, pp.gtarget, d_npairs, nit, tmp,
                 nullptr, nullptr, false, s, pp.gmax, pp.item_cap - pp.ntiles - 1);
    launch_plan_check(lbase, pp.nmaps, pp.d_tm, d_npairs, pp.pair_cap, (int *)pp.flag.p, s);
    launch_pairs<Rec2>(l1, rec, pp.d_ch, pp.nchunks, pp.d_tm, pp.max_tiles, cnt, pp.ncnt, tmp,
                       (int32_t *)pp.pairs.p, true, s, lbase, pp.pair_cap);
    launch_items(cnt, pp.d_tm, pp.nmaps, pp.ntiles, pp.gmin, pp.gtarget, d_npairs, nit, tmp,
                 (int4 *)pp.items.p, (int64_t *)pp.tile_items.p, true, s, pp.gmax,
                 pp.item_cap - pp.ntiles - 1);
    S2G_HIP(hipStreamWaitEvent(s, pl->normed, 0));
    for (int plane0 = 0; plane0 < pp.nplanes; plane0 += pp.gstep) {
        const int np = pp.nplanes - plane0 < pp.gstep ? pp.nplanes - plane0 : pp.gstep;
        GatherArgs g;
        g.dtype = pp.dtype;
        g.kernel = pp.kernel;
        g.fp32acc = pp.fp32acc;
        g.np = np;
        g.plane0 = plane0;
        g.nq = pp.nq;
        g.nmaps = pp.nmaps;
        g.ntiles = pp.ntiles;
        g.nitems = pp.item_cap;
        g.d_nitems = nit + pp.ntiles;
        g.d_flag = (const int *)pp.flag.p;
        g.items = (const int4 *)pp.items.p;
        g.pairs = (const int32_t *)pp.pairs.p;
        g.rec = rec;
        g.maps = pp.d_maps;
        g.partial = (double *)pp.partial.p;
        g.tile_items = (const int64_t *)pp.tile_items.p;
        g.add = 1;
        launch_2d_gather(g, s);
        if (plane0 == 0)
            S2G_HIP(hipStreamWaitEvent(s, pl->smalled, 0));
        launch_2d_finalize(g, s);
    }
    if (pp.any_out)
        launch_2d_reduce_maps(pp.d_maps, pp.nmaps, pp.ntiles, pp.nq, s);
}

// Sizes one pass from an eager run of its first stages, allocates it.
void plan_pass_build(s2g_ctx *c, PlanPass &pp, double headroom, hipStream_t s)
{
    const int nmaps = pp.nmaps;
    // tables into plan-owned memory (the eager tables live in the context)
    PassTables pt = pass_tables(c, pp.mi.data(), nmaps, s);
    const size_t btab = sizeof(Map2D) * nmaps, bpb = sizeof(int64_t) * (nmaps + 1);
    const size_t bst = sizeof(TileMap) * nmaps;
    char *d = (char *)pp.tab.get(btab + bpb + bst);
    S2G_HIP(hipMemcpyAsync(d, pt.d_maps, btab + bpb + bst, hipMemcpyDeviceToDevice, s));
    pp.d_maps = (const Map2D *)d;
    pp.d_pbase = (const int64_t *)(d + btab);
    pp.ntot = pt.ntot;
    pp.ntiles = pt.ntiles;
    pp.nstiles = pt.nstiles;
    pp.any_out = pt.any_out;
    if (pp.ntot == 0)  // the class stages need at least one particle per pass
        throw Error(S2G_EINVAL, "every pass of a plan needs particles (maps with the same "
                                "kernel, nq and accumulator have none)");
    const int64_t n = pp.ntot > 0 ? pp.ntot : 1;
    pp.nb = partition_blocks(n);
    pp.rec.get(sizeof(Rec2) * n);
    pp.cls.get(sizeof(int32_t) * n);
    pp.counts.get(sizeof(int64_t) * (3 * pp.nb + 1));
    pp.l0.get(sizeof(int32_t) * n);
    pp.l1.get(sizeof(int32_t) * n);
    pp.l2.get(sizeof(int32_t) * n);
    pp.bounds.get(sizeof(int64_t) * (2 * nmaps + 6));
    pp.flag.get(sizeof(int) * 2);
    // creation run of the class stages: the per-map large counts
    launch_2d_setup(pp.dtype, pp.d_maps, nmaps, pp.ntot, pp.nq, pp.kernel, (Rec2 *)pp.rec.p,
                    (int32_t *)pp.cls.p, s);
    pp.scan_tmp.get(sizeof(int64_t) * scan_tmp_elems(3 * pp.nb + 1));
    launch_partition((const int32_t *)pp.cls.p, pp.ntot, (int64_t *)pp.counts.p,
                     (int64_t *)pp.scan_tmp.p, (int32_t *)pp.l0.p, (int32_t *)pp.l1.p,
                     (int32_t *)pp.l2.p, s);
    launch_class_bounds((const int64_t *)pp.counts.p, pp.nb, (const int32_t *)pp.l0.p,
                        (const int32_t *)pp.l1.p, pp.d_pbase,
                        nmaps, (int64_t *)pp.bounds.p, s);
    std::vector<int64_t> hb(nmaps + 5);
    S2G_HIP(hipMemcpyAsync(hb.data(), pp.bounds.p, sizeof(int64_t) * (nmaps + 5),
                           hipMemcpyDeviceToHost, s));
    S2G_HIP(hipStreamSynchronize(s));
    // chunks: csz per map from its large count (as the eager pass), nch
    // covering headroom times that count, at least one chunk
    std::vector<TileMap> tm = pt.tm;
    std::vector<PairChunk> chunks;
    int64_t cb = 0;
    pp.max_tiles = 1;
    for (int m = 0; m < nmaps; ++m) {
        if (tm[m].ntiles > kMaxTiles)
            throw Error(S2G_EUNSUPPORTED, "grid too large for the tile lists");
        pp.max_tiles = tm[m].ntiles > pp.max_tiles ? tm[m].ntiles : pp.max_tiles;
        const int64_t nl = hb[4 + m + 1] - hb[4 + m];
        const int64_t cap = (int64_t)std::ceil((double)nl * headroom) + 64;
        int64_t csz = (cap + 32 * (int64_t)c->ncu - 1) / (32 * (int64_t)c->ncu);
        csz = csz < 64 ? 64 : csz > kPairChunk ? kPairChunk : (csz + 63) / 64 * 64;
        while ((int64_t)tm[m].ntiles * ((cap + csz - 1) / csz) > (1 << 25))
            csz *= 2;
        const int nch = (int)((cap + csz - 1) / csz);
        tm[m].cbase = cb;
        tm[m].nch = nch;
        tm[m].csz = (int32_t)csz;
        cb += (int64_t)tm[m].ntiles * nch;
        for (int k = 0; k < nch; ++k)
            chunks.push_back(PairChunk{m, k});
    }
    pp.nchunks = (int)chunks.size();
    pp.ncnt = cb + 1;
    const size_t btm = sizeof(TileMap) * nmaps, bch = sizeof(PairChunk) * chunks.size();
    char *dt = (char *)pp.tiles.get(btm + bch);
    S2G_HIP(hipMemcpy(dt, tm.data(), btm, hipMemcpyHostToDevice));
    S2G_HIP(hipMemcpy(dt + btm, chunks.data(), bch, hipMemcpyHostToDevice));
    pp.d_tm = (const TileMap *)dt;
    pp.d_ch = (const PairChunk *)(dt + btm);
    // small lists: chunks of each map's small list over its small-tile grid;
    // capacity: all of the map's particles (a true bound)
    std::vector<TileMap> stm;
    std::vector<PairChunk> schunks;
    int64_t scb = 0, stot = 0;
    pp.smax_tiles = 1;
    for (int m = 0; m < nmaps; ++m) {
        const int nts = (int)((pp.mi[m].a.npix + kST - 1) / kST);
        const int64_t ntl = (int64_t)nts * nts;
        if (ntl > kMaxTiles)
            throw Error(S2G_EUNSUPPORTED, "map " + std::to_string(m) + ": npixels above " +
                                              std::to_string(kST * 128) +
                                              " is not supported in a plan");
        pp.smax_tiles = (int)ntl > pp.smax_tiles ? (int)ntl : pp.smax_tiles;
        const int64_t cap = pp.mi[m].a.n > 0 ? pp.mi[m].a.n : 1;
        int64_t csz = (cap + 8 * (int64_t)c->ncu - 1) / (8 * (int64_t)c->ncu);
        csz = csz < kPairChunk ? kPairChunk : (csz + 63) / 64 * 64;
        while (ntl * ((cap + csz - 1) / csz) > (1 << 25))
            csz *= 2;
        const int nch = (int)((cap + csz - 1) / csz);
        stm.push_back(TileMap{scb, (int32_t)stot, (int32_t)ntl, nts, 1, nch, (int32_t)csz});
        scb += ntl * nch;
        stot += ntl;
        for (int k = 0; k < nch; ++k)
            schunks.push_back(PairChunk{m, k});
    }
    pp.nschunks = (int)schunks.size();
    pp.sncnt = scb + 1;
    const size_t bstm = sizeof(TileMap) * nmaps, bsch = sizeof(PairChunk) * schunks.size();
    char *ds = (char *)pp.stiles.get(bstm + bsch);
    S2G_HIP(hipMemcpy(ds, stm.data(), bstm, hipMemcpyHostToDevice));
    S2G_HIP(hipMemcpy(ds + bstm, schunks.data(), bsch, hipMemcpyHostToDevice));
    pp.d_stm = (const TileMap *)ds;
    pp.d_sch = (const PairChunk *)(ds + bstm);
    pp.scnt.get(sizeof(int64_t) * pp.sncnt);
    // scans: class counts, pair counts and items on the main branch; small
    // lists and items on the side branch, with scratch of their own
    int64_t nscan = 3 * pp.nb + 1;
    for (int64_t v : {pp.ncnt, (int64_t)pp.ntiles + 1})
        nscan = v > nscan ? v : nscan;
    pp.scan_tmp.get(sizeof(int64_t) * scan_tmp_elems(nscan));
    const int64_t nsscan = pp.sncnt > (int64_t)pp.nstiles + 1 ? pp.sncnt : pp.nstiles + 1;
    pp.sscan_tmp.get(sizeof(int64_t) * scan_tmp_elems(nsscan));
    pp.cnt.get(sizeof(int64_t) * pp.ncnt);
    // creation run of the pair count: the pair capacity
    launch_pairs<Rec2>((const int32_t *)pp.l1.p, (const Rec2 *)pp.rec.p, pp.d_ch, pp.nchunks,
                       pp.d_tm, pp.max_tiles, (int64_t *)pp.cnt.p, pp.ncnt,
                       (int64_t *)pp.scan_tmp.p, nullptr, false, s,
                       (const int64_t *)pp.bounds.p + 4);
    int64_t npairs = 0;
    S2G_HIP(hipMemcpyAsync(&npairs, (int64_t *)pp.cnt.p + pp.ncnt - 1, sizeof(int64_t),
                           hipMemcpyDeviceToHost, s));
    S2G_HIP(hipStreamSynchronize(s));
    pp.pair_cap = (int64_t)std::ceil((double)npairs * headroom) + 1024;
    if (pp.pair_cap >= (int64_t)INT32_MAX)
        throw Error(S2G_EUNSUPPORTED, "more than 2^31 (tile, particle) pairs");
    pp.pairs.get(sizeof(int32_t) * pp.pair_cap);
    pp.gmin = c->tune.gather_item_min;
    pp.gtarget = (int64_t)c->ncu * c->tune.gather_items_per_cu;
    pp.gmax = c->tune.gather_item_max;
    // items <= sum over tiles of (count / G + 1) <= total / G + ntiles; the
    // budget B = max(target, pair_cap / gmax) bounds total / G at B whatever
    // the replay's total (G >= total / B, item_size)
    int64_t budget = pp.gtarget;
    if (pp.gmax > 0) {
        const int64_t bg = (pp.pair_cap + pp.gmax - 1) / pp.gmax;
        budget = bg > budget ? bg : budget;
    }
    pp.item_cap = pp.ntiles + budget + 1;
    pp.nitems.get(sizeof(int64_t) * (pp.ntiles + 1));
    pp.items.get(sizeof(int4) * pp.item_cap);
    pp.tile_items.get(sizeof(int64_t) * 2 * pp.ntiles);
    pp.gstep = c->tune.gather_planes;
    pp.sstep = c->tune.small_planes;
    const int gnp = pp.nplanes < pp.gstep ? pp.nplanes : pp.gstep;
    pp.partial.get(sizeof(double) * pp.item_cap * gnp * kGT * kGT);
    pp.smin = c->tune.small_item_min;
    pp.starget = (int64_t)c->ncu * c->tune.small_items_per_cu;
    pp.sitem_cap = pp.nstiles + pp.starget + 1;
    if (pp.sitem_cap > ((int64_t)1 << 31) / 1024)
        throw Error(S2G_EUNSUPPORTED, "too many small-footprint work items");
    pp.sorted.get(sizeof(int32_t) * n);
    pp.snitems.get(sizeof(int64_t) * (pp.nstiles + 1));
    pp.sitems.get(sizeof(int4) * pp.sitem_cap);
    pp.stile.get(sizeof(int64_t) * 2 * (pp.nstiles > 0 ? pp.nstiles : 1));
}

void plan_build(s2g_ctx *c, s2g_plan *pl, int dtype, const s2g_map_desc *maps, int nmaps,
                double headroom)
{
    hipStream_t s = c->stream;
    for (const std::vector<int> &grp : batch_groups(maps, nmaps)) {
        pl->passes.emplace_back();
        PlanPass &pp = pl->passes.back();
        const int64_t flat = batch_flat_elems(maps, grp);
        double *img = flat ? (double *)pp.img.get(sizeof(double) * flat) : nullptr;
        pp.mi = batch_inputs(dtype, maps, grp, img);
        pp.nmaps = (int)pp.mi.size();
        pp.dtype = dtype;
        pp.kernel = maps[grp[0]].kernel;
        pp.nq = maps[grp[0]].nq;
        pp.nplanes = pp.nq + 1;
        pp.fp32acc = (maps[grp[0]].flags & S2G_FLAG_FP32_ACCUM) ? 1 : 0;
        plan_pass_build(c, pp, headroom, s);
    }
    S2G_HIP(hipStreamSynchronize(s));
    S2G_HIP(hipHostMalloc((void **)&pl->h_read, sizeof(int64_t) * 8 * pl->passes.size(),
                          hipHostMallocDefault));
    // the replayed sequence, captured on a stream of its own (and its side
    // branch, default priority like every stream here)
    S2G_HIP(hipStreamCreateWithFlags(&pl->cs, hipStreamNonBlocking));
    S2G_HIP(hipStreamCreateWithFlags(&pl->cs2, hipStreamNonBlocking));
    for (hipEvent_t *e : {&pl->fork, &pl->normed, &pl->smalled})
        S2G_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
    S2G_HIP(hipStreamBeginCapture(pl->cs, hipStreamCaptureModeThreadLocal));
    try {
        for (PlanPass &pp : pl->passes)
            plan_enqueue(pl, pp, pl->cs, pl->cs2);
        S2G_HIP(hipGetLastError());
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(pl->cs, &g);
        if (g)
            (void)hipGraphDestroy(g);
        throw;
    }
    S2G_HIP(hipStreamEndCapture(pl->cs, &pl->graph));
    S2G_HIP(hipGraphInstantiate(&pl->exec, pl->graph, nullptr, nullptr, 0));
}

int plan_run(s2g_plan *pl, hipStream_t s, s2g_stats *st)
{
    s2g_ctx *c = pl->c;
    hipEvent_t *E = c->stage;
    if (st)
        std::memset(st, 0, sizeof(*st));
    S2G_HIP(hipEventRecord(E[0], s));
    S2G_HIP(hipGraphLaunch(pl->exec, s));
    S2G_HIP(hipEventRecord(E[1], s));
    // per pass: the capacity flag, then the class bounds and the pair total
    for (size_t k = 0; k < pl->passes.size(); ++k) {
        const PlanPass &pp = pl->passes[k];
        int64_t *h = pl->h_read + 8 * k;
        S2G_HIP(hipMemcpyAsync(h, pp.flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
        S2G_HIP(hipMemcpyAsync(h + 1, pp.bounds.p, sizeof(int64_t) * 4, hipMemcpyDeviceToHost, s));
        S2G_HIP(hipMemcpyAsync(h + 5, (const int64_t *)pp.cnt.p + pp.ncnt - 1, sizeof(int64_t),
                               hipMemcpyDeviceToHost, s));
    }
    S2G_HIP(hipStreamSynchronize(s));
    std::string over;
    for (size_t k = 0; k < pl->passes.size(); ++k) {
        const PlanPass &pp = pl->passes[k];
        const int64_t *h = pl->h_read + 8 * k;
        const int flag = *(const int *)h;
        if (st) {
            st->n_particles += pp.ntot;
            st->n_small += h[2] - h[1];
            st->n_large += h[3] - h[2];
            st->n_large_exact += h[4] - h[3];
            st->n_pairs += h[5];
        }
        if (flag)
            over += " pass " + std::to_string(k) + ":" +
                    ((flag & 1) ? " large particles past the chunk capacity" : "") +
                    ((flag & 2) ? " " + std::to_string(h[5]) + " pairs for a capacity of " +
                                      std::to_string(pp.pair_cap)
                                : "");
    }
    if (st) {
        st->ms_device = elapsed(E[0], E[1]);
        st->n_skipped = st->n_particles - st->n_small - st->n_large;
    }
    if (!over.empty())
        throw Error(S2G_ECAPACITY, "the inputs outgrew the plan's capacities (outputs "
                                   "incomplete; create a new plan):" + over);
    return S2G_OK;
}

}  // namespace

extern "C" {

const char *s2g_version(void) { return kVersion; }

int s2g_set_tuning(s2g_ctx *c, const s2g_tuning *t)
{
    return guarded(c, [&] {
        if (!t) {
            c->tune = kDefaultTuning;
            return S2G_OK;
        }
        auto in = [](int32_t v, int32_t lo, int32_t hi) { return v >= lo && v <= hi; };
        // >= 5: a footprint of hsml < 2 px spans at most 5 pixels per axis, and
        // only the scatter has the FP64 weights such footprints need
        if (!in(t->small_side, 5, kSmallSide2D))
            throw Error(S2G_EINVAL, "small_side out of range [5, " +
                                        std::to_string(kSmallSide2D) + "]");
        if (!in(t->small_item_min, 16, 4096) || !in(t->small_items_per_cu, 1, 64))
            throw Error(S2G_EINVAL, "small_item_min / small_items_per_cu out of range "
                                    "[16, 4096] / [1, 64]");
        if (!in(t->gather_item_min, 8, 4096) || !in(t->gather_items_per_cu, 1, 256))
            throw Error(S2G_EINVAL, "gather_item_min / gather_items_per_cu out of range "
                                    "[8, 4096] / [1, 256]");
        if (!in(t->gather_planes, 1, 4) || !in(t->small_planes, 1, 4))
            throw Error(S2G_EINVAL, "gather_planes / small_planes out of range [1, 4]");
        if (t->gather_item_max != 0 &&
            !in(t->gather_item_max, t->gather_item_min, (int32_t)1 << 30))
            throw Error(S2G_EINVAL, "gather_item_max out of range (0, or gather_item_min "
                                    "to 2^30)");
        c->tune = *t;
        return S2G_OK;
    });
}

int s2g_get_tuning(s2g_ctx *c, s2g_tuning *t)
{
    return guarded(c, [&] {
        if (!t)
            throw Error(S2G_EINVAL, "tuning is NULL");
        *t = c->tune;
        return S2G_OK;
    });
}

int s2g_create(s2g_ctx **out, int device)
{
    if (!out)
        return S2G_EINVAL;
    *out = nullptr;
    s2g_ctx *c = nullptr;
    try {
        c = new s2g_ctx();
        c->device = device;
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) {
            delete c;
            return S2G_EHIP;
        }
        if (hipSetDevice(device) != hipSuccess) {
            delete c;
            return S2G_EHIP;
        }
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess)
            c->ncu = prop.multiProcessorCount;
        if (hipStreamCreateWithFlags(&c->stream, hipStreamDefault) != hipSuccess ||
            hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&c->fork, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&c->normed, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&c->smalled, hipEventDisableTiming) != hipSuccess ||
            hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
            hipHostMalloc((void **)&c->h_small, 64 * sizeof(int64_t), hipHostMallocDefault) !=
                hipSuccess) {
            s2g_destroy(c);
            return S2G_EHIP;
        }
        bool ok = true;
        for (auto &e : c->stage)
            ok = ok && hipEventCreate(&e) == hipSuccess;
        if (!ok) {
            s2g_destroy(c);
            return S2G_EHIP;
        }
        c->scratch = scratch_address(device, &c->scratch_err);
    } catch (...) {
        delete c;
        return S2G_ENOMEM;
    }
    *out = c;
    return S2G_OK;
}

void s2g_destroy(s2g_ctx *c)
{
    if (!c)
        return;
    (void)hipSetDevice(c->device);
    if (c->stream)
        (void)hipStreamSynchronize(c->stream);
    if (c->side)
        (void)hipStreamSynchronize(c->side);
    for (auto &b : c->in) b.release();
    c->rec.release();
    c->cls.release();
    c->counts.release();
    c->scan_tmp.release();
    for (auto &b : c->lists) b.release();
    c->cnt.release();
    c->pairs.release();
    c->nitems.release();
    c->items.release();
    c->tile_items.release();
    c->partial.release();
    c->tab.release();
    c->bounds.release();
    c->tiles.release();
    for (DevBuf *b : {&c->skeys, &c->svals, &c->sort_tmp, &c->soff, &c->snitems, &c->sitems,
                      &c->stile, &c->sscan_tmp, &c->md_P, &c->md_bbox, &c->md_S, &c->md_tab})
        b->release();
    c->h_tab.release();
    c->h_bounds.release();
    c->h_tiles.release();
    c->hp_rec.release();
    c->hp_nunits.release();
    c->hp_units.release();
    c->hp_part.release();
    c->hp_ring.release();
    c->hp_rtab.release();
    c->hp_list.release();
    c->tsc_off.release();
    c->tsc_gq.release();
    c->tsc_cell.release();
    c->image.release();
    c->out.release();
    if (c->h_small)
        (void)hipHostFree(c->h_small);
    if (c->ev0)
        (void)hipEventDestroy(c->ev0);
    if (c->ev1)
        (void)hipEventDestroy(c->ev1);
    for (auto &e : c->stage)
        if (e)
            (void)hipEventDestroy(e);
    for (hipEvent_t e : {c->fork, c->normed, c->smalled})
        if (e)
            (void)hipEventDestroy(e);
    if (c->side)
        (void)hipStreamDestroy(c->side);
    if (c->stream)
        (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *s2g_last_error(const s2g_ctx *c) { return c ? c->err.c_str() : "null context"; }

int s2g_deposit_device(s2g_ctx *c, int dims, int dtype, const void *d_pos, const void *d_hsml,
                       const void *d_m, const void *d_rho, const void *d_bin_q, int nq,
                       const void *d_weights, int64_t n, const s2g_params *params, int kernel,
                       int kernel_dim, uint32_t flags, double *d_image_out, void *stream,
                       s2g_stats *stats)
{
    return guarded(c, [&]() {
        check_common(dtype, kernel, kernel_dim, n, params);
        check_dims_nq(dims, nq);
        if (!d_image_out)
            throw Error(S2G_EINVAL, "output pointer is NULL");
        hipStream_t s = stream ? (hipStream_t)stream : c->stream;
        Map2DArgs a = make_args(dtype, d_pos, d_hsml, d_m, d_rho, d_bin_q, nq, d_weights, n,
                                params, kernel, flags);
        deposit_device_any(c, dims, a, d_image_out, s, stats);
        return S2G_OK;
    });
}

int s2g_deposit(s2g_ctx *c, int dims, int dtype, const void *pos, const void *hsml,
                const void *m, const void *rho, const void *bin_q, int nq, const void *weights,
                int64_t n, const s2g_params *params, int kernel, int kernel_dim, uint32_t flags,
                double *image_out, s2g_stats *stats)
{
    return guarded(c, [&]() {
        check_common(dtype, kernel, kernel_dim, n, params);
        check_dims_nq(dims, nq);
        if (!image_out)
            throw Error(S2G_EINVAL, "output pointer is NULL");
        const int64_t npix = params->npixels;
        const int64_t nd = dims == 2 ? npix * npix : npix * npix * npix;
        const int nplanes = nq + 1;
        hipStream_t s = c->stream;
        s2g_stats local;
        s2g_stats *st = stats ? stats : &local;
        Map2DArgs h = make_args(dtype, pos, hsml, m, rho, bin_q, nq, weights, n, params,
                                kernel, flags);
        double h2d_ms = 0.0;
        {
            s2g_stats tmp;
            h = stage_inputs(c, h, s, &tmp);
            h2d_ms = tmp.ms_h2d;
        }
        double *img = (double *)c->image.get(sizeof(double) * nd * nplanes);
        deposit_device_any(c, dims, h, img, s, st);
        st->ms_h2d = h2d_ms;
        d2h(c, image_out, img, sizeof(double) * nd * nplanes, s, &st->ms_d2h);
        return S2G_OK;
    });
}

int s2g_reduce_device(s2g_ctx *c, int dims, int64_t npix, int nq, uint32_t flags,
                      const double *d_image, double *d_out, void *stream)
{
    return guarded(c, [&]() {
        check_dims_nq(dims, nq);
        if (npix < 1 || npix > 65536)
            throw Error(S2G_EINVAL, "npixels out of range [1, 65536]");
        if (!d_image || !d_out)
            throw Error(S2G_EINVAL, "image pointer is NULL");
        hipStream_t s = stream ? (hipStream_t)stream : c->stream;
        if (dims == 2)
            launch_2d_reduce(d_image, npix, nq, (flags & S2G_FLAG_REDUCE) ? 1 : 0, d_out, s);
        else
            launch_3d_reduce(d_image, npix, (flags & S2G_FLAG_REDUCE) ? 1 : 0, d_out, s);
        S2G_HIP(hipGetLastError());
        return S2G_OK;
    });
}

int s2g_reduce(s2g_ctx *c, int dims, int64_t npix, int nq, uint32_t flags, const double *image,
               double *out)
{
    return guarded(c, [&]() {
        check_dims_nq(dims, nq);
        if (npix < 1 || npix > 65536)
            throw Error(S2G_EINVAL, "npixels out of range [1, 65536]");
        if (!image || !out)
            throw Error(S2G_EINVAL, "image pointer is NULL");
        const int64_t nd = dims == 2 ? npix * npix : npix * npix * npix;
        const int nin = dims == 2 ? nq + 1 : 2, nout = dims == 2 ? nq : 1;
        hipStream_t s = c->stream;
        double *img = (double *)c->image.get(sizeof(double) * nd * nin);
        double *o = (double *)c->out.get(sizeof(double) * nd * nout);
        S2G_HIP(hipMemcpyAsync(img, image, sizeof(double) * nd * nin, hipMemcpyHostToDevice, s));
        int rc = s2g_reduce_device(c, dims, npix, nq, flags, img, o, s);
        if (rc)
            return rc;
        d2h(c, out, o, sizeof(double) * nd * nout, s, nullptr);
        return S2G_OK;
    });
}

int s2g_map(s2g_ctx *c, int dims, int dtype, const void *pos, const void *hsml, const void *m,
            const void *rho, const void *bin_q, int nq, const void *weights, int64_t n,
            const s2g_params *params, int kernel, int kernel_dim, uint32_t flags,
            double *map_out, s2g_stats *stats)
{
    return guarded(c, [&]() {
        check_common(dtype, kernel, kernel_dim, n, params);
        check_dims_nq(dims, nq);
        if (!map_out)
            throw Error(S2G_EINVAL, "output pointer is NULL");
        const int64_t npix = params->npixels;
        const int64_t nd = dims == 2 ? npix * npix : npix * npix * npix;
        const int nout = dims == 2 ? nq : 1;
        hipStream_t s = c->stream;
        s2g_stats local;
        s2g_stats *st = stats ? stats : &local;
        Map2DArgs h = make_args(dtype, pos, hsml, m, rho, bin_q, nq, weights, n, params,
                                kernel, flags);
        double h2d_ms = 0.0;
        {
            s2g_stats tmp;
            h = stage_inputs(c, h, s, &tmp);
            h2d_ms = tmp.ms_h2d;
        }
        double *img = (double *)c->image.get(sizeof(double) * nd * (nq + 1));
        deposit_device_any(c, dims, h, img, s, st);
        st->ms_h2d = h2d_ms;
        double *o = (double *)c->out.get(sizeof(double) * nd * nout);
        int rc = s2g_reduce_device(c, dims, npix, nq, flags, img, o, s);
        if (rc)
            return rc;
        d2h(c, map_out, o, sizeof(double) * nd * nout, s, &st->ms_d2h);
        return S2G_OK;
    });
}

// all-sky deposit of a.nmaps maps of one particle set on device pointers:
// map k into maps[k] / wmaps[k] (every pixel written).  The disc queries, the
// normalisation sums and the per-ring lists are computed once for all maps.
static void healpix_device(s2g_ctx *c, const HpArgs &a, double *const *maps,
                           double *const *wmaps, hipStream_t s, s2g_stats *st)
{
    const int64_t n = a.n, nside = a.nside;
    const int64_t npx = 12 * nside * nside;
    hipEvent_t *E = c->stage;
    S2G_HIP(hipEventRecord(E[0], s));
    // the ring-major deposit writes every pixel; the atomic one adds
    const bool rings = n > 0 && hp_ring_deposit_fits(nside);
    if (!rings)
        for (int k = 0; k < a.nmaps; ++k) {
            S2G_HIP(hipMemsetAsync(maps[k], 0, sizeof(double) * npx, s));
            S2G_HIP(hipMemsetAsync(wmaps[k], 0, sizeof(double) * npx, s));
        }
    // per-ring half-angle terms, once per call
    double4 *rtab = (double4 *)c->hp_rtab.get(sizeof(double4) * 4 * nside);
    launch_hp_ring_table(nside, rtab, s);
    auto group = [&](int m0) {
        HpGroup g;
        g.rtab = rtab;
        g.nm = a.nmaps - m0 < 2 ? a.nmaps - m0 : 2;
        for (int k = 0; k < 2; ++k) {
            const int mk = m0 + (k < g.nm ? k : 0);
            g.binq[k] = a.binq[mk];
            g.w[k] = a.w[mk];
            g.map[k] = maps[mk];
            g.wmap[k] = wmaps[mk];
        }
        return g;
    };
    if (n > 0) {
        HpRec *rec = (HpRec *)c->hp_rec.get(sizeof(HpRec) * n);
        int64_t *nun = (int64_t *)c->hp_nunits.get(sizeof(int64_t) * (n + 1));
        int64_t *tmp = (int64_t *)c->scan_tmp.get(sizeof(int64_t) * scan_tmp_elems(n + 1));
        launch_hp_setup(a, rec, nun, s);
        exclusive_scan_i64(nun, n + 1, tmp, s);
        S2G_HIP(hipEventRecord(E[1], s));
        int64_t nu;
        read_i64(c, s, {nun + n}, &nu);
        if (nu >= (int64_t)INT32_MAX)
            throw Error(S2G_EUNSUPPORTED, "too many all-sky work units");
        int2 *units = (int2 *)c->hp_units.get(sizeof(int2) * (nu ? nu : 1));
        double4 *part = (double4 *)c->hp_part.get(sizeof(double4) * (nu ? nu : 1));
        launch_hp_units(nun, n, rec, units, s);
        // the per-ring lists (side stream) do not read what the normalisation
        // writes (the units' partial sums, the records' scale): they are
        // built beside it, and the deposit waits for both
        hipStream_t s2 = c->side;
        if (rings) {
            S2G_HIP(hipEventRecord(c->fork, s));
            S2G_HIP(hipStreamWaitEvent(s2, c->fork, 0));
        }
        S2G_HIP(hipEventRecord(E[2], s));
        launch_hp_norm(a.kernel, units, nu, rec, nside, rtab, part, s);
        launch_hp_finish(nun, n, part, rec, s);
        S2G_HIP(hipEventRecord(E[3], s));
        if (rings) {
            // per-ring particle lists: sorted entries (units + extra centre
            // rings), per-ring counts, scan, ordered fill
            const int64_t nslots = 4 * nside + 1;
            const int64_t ne = nu + n;
            if (ne >= (int64_t)INT32_MAX)
                throw Error(S2G_EUNSUPPORTED, "too many all-sky work units");
            uint32_t *keys = (uint32_t *)c->skeys.get(sizeof(uint32_t) * 2 * ne);
            int32_t *vals = (int32_t *)c->svals.get(sizeof(int32_t) * 2 * ne);
            launch_hp_entry_keys(units, nu, rec, n, nside, keys, vals, s2);
            const int kbits = hp_entry_key_bits(nside);
            const size_t tb = sort_pairs_tmp_bytes(ne, kbits);
            void *sortmp = c->sort_tmp.get(tb ? tb : 1);
            sort_pairs(sortmp, tb, keys, keys + ne, vals, vals + ne, ne, kbits, s2);
            int64_t *rb = (int64_t *)c->hp_ring.get(sizeof(int64_t) * 2 * nslots);
            int64_t *eoff = rb, *off = rb + nslots;
            int64_t *stmp =
                (int64_t *)c->sscan_tmp.get(sizeof(int64_t) * scan_tmp_elems(nslots));
            launch_hp_entry_offsets(keys + ne, ne, nside, eoff, s2);
            launch_hp_ring_count(keys + ne, eoff, nside, off, s2);
            exclusive_scan_i64(off, nslots, stmp, s2);
            int64_t npairs;
            read_i64(c, s2, {off + nslots - 1}, &npairs);
            if (npairs >= (int64_t)INT32_MAX)
                throw Error(S2G_EUNSUPPORTED, "too many all-sky (particle, ring) pairs");
            int32_t *list = (int32_t *)c->hp_list.get(sizeof(int32_t) * (npairs ? npairs : 1));
            launch_hp_ring_fill(keys + ne, vals + ne, eoff, off, nside, list, s2);
            S2G_HIP(hipEventRecord(c->smalled, s2));
            S2G_HIP(hipStreamWaitEvent(s, c->smalled, 0));
            for (int m0 = 0; m0 < a.nmaps; m0 += 2)
                launch_hp_ring_deposit(a, group(m0), list, off, rec, s);
        } else {
            for (int m0 = 0; m0 < a.nmaps; m0 += 2)
                launch_hp_deposit(a, group(m0), units, nu, rec, s);
        }
        S2G_HIP(hipGetLastError());
        if (st)
            st->n_pairs = nu;
    } else {
        // nothing to deposit: every map is zero (written by the memset above
        // unless the ring path was chosen, which needs particles)
    }
    S2G_HIP(hipEventRecord(E[4], s));
    S2G_HIP(hipEventSynchronize(E[4]));
    if (st) {
        st->ms_device = elapsed(E[0], E[4]);
        if (n > 0) {
            st->ms_setup = elapsed(E[0], E[1]);
            st->ms_norm = elapsed(E[2], E[3]);
            st->ms_small = elapsed(E[3], E[4]);
        }
    }
}

static HpArgs hp_args(int dtype, int kernel, const void *pos, const void *hsml, const void *m,
                      const void *rho, int nmaps, const void *const *bin_q,
                      const void *const *w, int64_t n, int64_t nside, uint32_t flags)
{
    HpArgs a;
    a.dtype = dtype;
    a.kernel = kernel;
    a.pos = pos;
    a.hsml = hsml;
    a.m = m;
    a.rho = rho;
    a.nmaps = nmaps;
    for (int k = 0; k < kHpMaxMaps; ++k) {
        a.binq[k] = k < nmaps ? bin_q[k] : nullptr;
        a.w[k] = k < nmaps ? w[k] : nullptr;
    }
    a.n = n;
    a.nside = nside;
    a.flags = flags;
    return a;
}

static void check_healpix(int dtype, int kernel, int kernel_dim, int64_t n, int64_t nside,
                          int nmaps)
{
    s2g_params dummy{1, 1.0};
    check_common(dtype, kernel, kernel_dim, n, &dummy);
    if (nside < 1 || nside > 8192)
        throw Error(S2G_EINVAL, "nside out of range [1, 8192]");
    if (nmaps < 1 || nmaps > kHpMaxMaps)
        throw Error(S2G_EINVAL, "nmaps out of range [1, " + std::to_string(kHpMaxMaps) + "]");
}

int s2g_healpix_deposit_maps_device(s2g_ctx *c, int dtype, const void *d_pos,
                                    const void *d_hsml, const void *d_m, const void *d_rho,
                                    int nmaps, const void *const *d_bin_q,
                                    const void *const *d_weights, int64_t n, int64_t nside,
                                    int kernel, int kernel_dim, uint32_t flags,
                                    double *const *d_map_out, double *const *d_weight_out,
                                    void *stream, s2g_stats *stats)
{
    return guarded(c, [&]() {
        check_healpix(dtype, kernel, kernel_dim, n, nside, nmaps);
        if (!d_bin_q || !d_weights || !d_map_out || !d_weight_out)
            throw Error(S2G_EINVAL, "per-map pointer array is NULL");
        for (int k = 0; k < nmaps; ++k) {
            if (!d_map_out[k] || !d_weight_out[k])
                throw Error(S2G_EINVAL, "map " + std::to_string(k) + ": output pointer is NULL");
            if (n > 0 && (!d_bin_q[k] || !d_weights[k]))
                throw Error(S2G_EINVAL, "map " + std::to_string(k) + ": input pointer is NULL");
        }
        if (n > 0 && (!d_pos || !d_hsml || !d_m || !d_rho))
            throw Error(S2G_EINVAL, "input array pointer is NULL");
        hipStream_t s = stream ? (hipStream_t)stream : c->stream;
        if (stats) {
            std::memset(stats, 0, sizeof(*stats));
            stats->n_particles = n;
        }
        healpix_device(c, hp_args(dtype, kernel, d_pos, d_hsml, d_m, d_rho, nmaps, d_bin_q,
                                  d_weights, n, nside, flags),
                       d_map_out, d_weight_out, s, stats);
        return S2G_OK;
    });
}

int s2g_healpix_deposit_device(s2g_ctx *c, int dtype, const void *d_pos, const void *d_hsml,
                               const void *d_m, const void *d_rho, const void *d_bin_q,
                               const void *d_weights, int64_t n, int64_t nside, int kernel,
                               int kernel_dim, uint32_t flags, double *d_map_out,
                               double *d_weight_out, void *stream, s2g_stats *stats)
{
    return s2g_healpix_deposit_maps_device(c, dtype, d_pos, d_hsml, d_m, d_rho, 1, &d_bin_q,
                                           &d_weights, n, nside, kernel, kernel_dim, flags,
                                           &d_map_out, &d_weight_out, stream, stats);
}

int s2g_healpix_deposit_maps(s2g_ctx *c, int dtype, const void *pos, const void *hsml,
                             const void *m, const void *rho, int nmaps,
                             const void *const *bin_q, const void *const *weights, int64_t n,
                             int64_t nside, int kernel, int kernel_dim, uint32_t flags,
                             double *const *map_out, double *const *weight_out,
                             s2g_stats *stats)
{
    return guarded(c, [&]() {
        check_healpix(dtype, kernel, kernel_dim, n, nside, nmaps);
        if (!bin_q || !weights || !map_out || !weight_out)
            throw Error(S2G_EINVAL, "per-map pointer array is NULL");
        for (int k = 0; k < nmaps; ++k)
            if (!map_out[k] || !weight_out[k] || (n > 0 && (!bin_q[k] || !weights[k])))
                throw Error(S2G_EINVAL, "map " + std::to_string(k) + ": pointer is NULL");
        if (n > 0 && (!pos || !hsml || !m || !rho))
            throw Error(S2G_EINVAL, "input array pointer is NULL");
        hipStream_t s = c->stream;
        s2g_stats local;
        s2g_stats *st = stats ? stats : &local;
        std::memset(st, 0, sizeof(*st));
        st->n_particles = n;
        const size_t es = elem_size(dtype);
        const size_t cnt = (size_t)(n > 0 ? n : 1);
        hipEvent_t e0 = c->ev0, e1 = c->ev1;
        S2G_HIP(hipEventRecord(e0, s));
        // staging: pos, hsml, m, rho in in[0..3]; per map bin_q and weights in
        // one buffer (in[4])
        const void *src[4] = {pos, hsml, m, rho};
        void *dst[4];
        for (int k = 0; k < 4; ++k) {
            const size_t b = es * cnt * (k == 0 ? 3 : 1);
            dst[k] = c->in[k].get(b);
            if (n > 0)
                S2G_HIP(hipMemcpyAsync(dst[k], src[k], b, hipMemcpyHostToDevice, s));
        }
        char *qw = (char *)c->in[4].get(es * cnt * 2 * nmaps);
        const void *dq[kHpMaxMaps], *dw[kHpMaxMaps];
        for (int k = 0; k < nmaps; ++k) {
            dq[k] = qw + es * cnt * (2 * k);
            dw[k] = qw + es * cnt * (2 * k + 1);
            if (n > 0) {
                S2G_HIP(hipMemcpyAsync((void *)dq[k], bin_q[k], es * n, hipMemcpyHostToDevice, s));
                S2G_HIP(hipMemcpyAsync((void *)dw[k], weights[k], es * n, hipMemcpyHostToDevice,
                                       s));
            }
        }
        S2G_HIP(hipEventRecord(e1, s));
        S2G_HIP(hipEventSynchronize(e1));
        const double h2d = elapsed(e0, e1);
        const int64_t npx = 12 * nside * nside;
        double *out = (double *)c->image.get(sizeof(double) * 2 * npx * nmaps);
        double *dm[kHpMaxMaps], *dwm[kHpMaxMaps];
        for (int k = 0; k < nmaps; ++k) {
            dm[k] = out + 2 * npx * k;
            dwm[k] = out + 2 * npx * k + npx;
        }
        healpix_device(c, hp_args(dtype, kernel, dst[0], dst[1], dst[2], dst[3], nmaps, dq, dw, n,
                                  nside, flags),
                       dm, dwm, s, st);
        st->n_particles = n;
        st->ms_h2d = h2d;
        double d2h_ms = 0.0;
        for (int k = 0; k < nmaps; ++k) {
            double t1 = 0.0, t2 = 0.0;
            d2h(c, map_out[k], dm[k], sizeof(double) * npx, s, &t1);
            d2h(c, weight_out[k], dwm[k], sizeof(double) * npx, s, &t2);
            d2h_ms += t1 + t2;
        }
        st->ms_d2h = d2h_ms;
        return S2G_OK;
    });
}

int s2g_healpix_deposit(s2g_ctx *c, int dtype, const void *pos, const void *hsml,
                        const void *m, const void *rho, const void *bin_q,
                        const void *weights, int64_t n, int64_t nside, int kernel,
                        int kernel_dim, uint32_t flags, double *map_out, double *weight_out,
                        s2g_stats *stats)
{
    return s2g_healpix_deposit_maps(c, dtype, pos, hsml, m, rho, 1, &bin_q, &weights, n, nside,
                                    kernel, kernel_dim, flags, &map_out, &weight_out, stats);
}

// Work space of one TSC deposit of n particles on an npix^3 grid.
struct TscWork {
    uint32_t *keys;   // 2n: keys, sorted keys
    int32_t *vals;    // 2n: particles, sorted particles
    int64_t *off;     // nbricks + 1
    double4 *gq;      // n
    uint64_t *cell;   // n
    void *sort_tmp;
    size_t sort_bytes;
};

// Bytes of a TscWork laid out in one buffer (each part 256-byte aligned).
static size_t tsc_work_bytes(int64_t n, int64_t npix, size_t sort_bytes)
{
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t m = (size_t)(n > 0 ? n : 1);
    return al(8 * m) + al(8 * m) + al(sizeof(int64_t) * (size_t)(tsc_nbricks(npix) + 1)) +
           al(sizeof(double4) * m) + al(8 * m) + al(sort_bytes ? sort_bytes : 1);
}

static TscWork tsc_work_in(void *base, int64_t n, int64_t npix, size_t sort_bytes)
{
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t m = (size_t)(n > 0 ? n : 1);
    char *p = (char *)base;
    TscWork w;
    w.keys = (uint32_t *)p;
    p += al(8 * m);
    w.vals = (int32_t *)p;
    p += al(8 * m);
    w.off = (int64_t *)p;
    p += al(sizeof(int64_t) * (size_t)(tsc_nbricks(npix) + 1));
    w.gq = (double4 *)p;
    p += al(sizeof(double4) * m);
    w.cell = (uint64_t *)p;
    p += al(8 * m);
    w.sort_tmp = p;
    w.sort_bytes = sort_bytes;
    return w;
}

// limits of s2g_tsc3d_device_noalloc (its work space fits kScratchBytes)
constexpr int64_t kNoallocMaxParticles = (int64_t)1 << 20;
constexpr int64_t kNoallocMaxNpix = 512;

static void check_tsc(int dtype, int64_t n, const s2g_params *params)
{
    check_common(dtype, S2G_KERNEL_CUBIC, 3, n, params);
    if (params->npixels > 2048)
        throw Error(S2G_EINVAL, "npixels out of range [1, 2048] for the TSC grid");
}

// The whole deposit on `s` (enqueues only): sort by brick, offsets, prep,
// gather.  Writes every cell of d_grid.
static void tsc_run(int dtype, const void *d_pos, const void *d_q, int64_t n, int64_t npix,
             double len2pix, int periodic, const TscWork &w, double *d_grid, hipStream_t s)
{
    if (n > 0) {
        launch_tsc_keys(dtype, d_pos, n, npix, len2pix, periodic, w.keys, w.vals, s);
        sort_pairs(w.sort_tmp, w.sort_bytes, w.keys, w.keys + n, w.vals, w.vals + n, n,
                   tsc_key_bits(npix), s);
    }
    // n == 0: every offset is 0 (no key is read)
    launch_tsc_offsets(w.keys + n, n, npix, w.off, s);
    launch_tsc_prep(dtype, d_pos, d_q, w.vals + n, w.off, n, npix, len2pix, periodic, w.gq,
                    w.cell, s);
    launch_tsc_gather(w.gq, w.cell, w.off, npix, periodic, d_grid, s);
}

// the same on the context's buffers (grown as needed)
static void tsc_ctx(s2g_ctx *c, int dtype, const void *d_pos, const void *d_q, int64_t n,
             const s2g_params *params, uint32_t flags, double *d_grid, hipStream_t s)
{
    const int64_t npix = params->npixels;
    const size_t m = (size_t)(n > 0 ? n : 1);
    TscWork w;
    w.keys = (uint32_t *)c->skeys.get(8 * m);
    w.vals = (int32_t *)c->svals.get(8 * m);
    w.off = (int64_t *)c->tsc_off.get(sizeof(int64_t) * (size_t)(tsc_nbricks(npix) + 1));
    w.gq = (double4 *)c->tsc_gq.get(sizeof(double4) * m);
    w.cell = (uint64_t *)c->tsc_cell.get(8 * m);
    w.sort_bytes = n > 0 ? sort_pairs_tmp_bytes(n, tsc_key_bits(npix)) : 0;
    w.sort_tmp = c->sort_tmp.get(w.sort_bytes ? w.sort_bytes : 1);
    tsc_run(dtype, d_pos, d_q, n, npix, params->len2pix, (flags & S2G_FLAG_PERIODIC) ? 1 : 0, w,
            d_grid, s);
}

int s2g_tsc3d(s2g_ctx *c, int dtype, const void *pos, const void *bin_q, int64_t n,
              const s2g_params *params, uint32_t flags, double *grid_out, s2g_stats *stats)
{
    return guarded(c, [&]() {
        check_tsc(dtype, n, params);
        if (!grid_out)
            throw Error(S2G_EINVAL, "output pointer is NULL");
        hipStream_t s = c->stream;
        s2g_stats local;
        s2g_stats *st = stats ? stats : &local;
        std::memset(st, 0, sizeof(*st));
        st->n_particles = n;
        const size_t es = elem_size(dtype);
        void *dpos = c->in[0].get(es * (size_t)(3 * n > 0 ? 3 * n : 1));
        void *dq = c->in[4].get(es * (size_t)(n > 0 ? n : 1));
        hipEvent_t *E = c->stage;
        S2G_HIP(hipEventRecord(E[0], s));
        if (n > 0) {
            if (!pos || !bin_q)
                throw Error(S2G_EINVAL, "input array pointer is NULL");
            S2G_HIP(hipMemcpyAsync(dpos, pos, es * 3 * n, hipMemcpyHostToDevice, s));
            S2G_HIP(hipMemcpyAsync(dq, bin_q, es * n, hipMemcpyHostToDevice, s));
        }
        const int64_t npix = params->npixels;
        const int64_t nd = npix * npix * npix;
        double *grid = (double *)c->image.get(sizeof(double) * 2 * nd);
        S2G_HIP(hipEventRecord(E[1], s));
        tsc_ctx(c, dtype, dpos, dq, n, params, flags, grid, s);
        S2G_HIP(hipGetLastError());
        S2G_HIP(hipEventRecord(E[2], s));
        S2G_HIP(hipEventSynchronize(E[2]));
        st->ms_h2d = elapsed(E[0], E[1]);
        st->ms_device = elapsed(E[1], E[2]);
        st->n_small = n;
        d2h(c, grid_out, grid, sizeof(double) * 2 * nd, s, &st->ms_d2h);
        return S2G_OK;
    });
}

int s2g_tsc3d_device(s2g_ctx *c, int dtype, const void *d_pos, const void *d_bin_q, int64_t n,
                     const s2g_params *params, uint32_t flags, double *d_grid_out, void *stream)
{
    return guarded(c, [&]() {
        check_tsc(dtype, n, params);
        if (!d_grid_out)
            throw Error(S2G_EINVAL, "output pointer is NULL");
        if (n > 0 && (!d_pos || !d_bin_q))
            throw Error(S2G_EINVAL, "input array pointer is NULL");
        hipStream_t s = stream ? (hipStream_t)stream : c->stream;
        tsc_ctx(c, dtype, d_pos, d_bin_q, n, params, flags, d_grid_out, s);
        S2G_HIP(hipGetLastError());
        // the particle order lives in the context's buffers: finish before
        // another call (possibly on another stream) reuses them
        S2G_HIP(hipStreamSynchronize(s));
        return S2G_OK;
    });
}

int s2g_tsc3d_device_noalloc(s2g_ctx *c, int dtype, const void *d_pos, const void *d_bin_q,
                             int64_t n, const s2g_params *params, uint32_t flags,
                             double *d_grid_out, void *stream)
{
    return guarded(c, [&]() {
        check_tsc(dtype, n, params);
        if (!d_grid_out)
            throw Error(S2G_EINVAL, "output pointer is NULL");
        if (n > 0 && (!d_pos || !d_bin_q))
            throw Error(S2G_EINVAL, "input array pointer is NULL");
        if (n > kNoallocMaxParticles)
            throw Error(S2G_EUNSUPPORTED,
                        "s2g_tsc3d_device_noalloc takes at most 2^20 particles (its fixed work "
                        "space); got " + std::to_string(n));
        if (params->npixels > kNoallocMaxNpix)
            throw Error(S2G_EUNSUPPORTED,
                        "s2g_tsc3d_device_noalloc takes npixels <= 512 (its fixed work space); "
                        "got " + std::to_string(params->npixels));
        if (!c->scratch)
            throw Error(S2G_EHIP, "fixed work space unavailable: " + c->scratch_err);
        const int64_t npix = params->npixels;
        const size_t sb = n > 0 ? sort_pairs_tmp_bytes(n, tsc_key_bits(npix)) : 0;
        const size_t need = tsc_work_bytes(n, npix, sb);
        if (need > kScratchBytes)
            throw Error(S2G_EUNSUPPORTED, "work space of " + std::to_string(need) +
                                              " bytes exceeds the fixed " +
                                              std::to_string(kScratchBytes));
        hipStream_t s = stream ? (hipStream_t)stream : c->stream;
        const TscWork w = tsc_work_in(c->scratch, n, npix, sb);
        tsc_run(dtype, d_pos, d_bin_q, n, npix, params->len2pix,
                (flags & S2G_FLAG_PERIODIC) ? 1 : 0, w, d_grid_out, s);
        S2G_HIP(hipGetLastError());
        return S2G_OK;
    });
}

int s2g_deposit_batch(s2g_ctx *c, int dtype, const s2g_map_desc *maps, int nmaps, void *stream,
                      s2g_stats *stats)
{
    return guarded(c, [&]() {
        check_batch(dtype, maps, nmaps);
        hipStream_t s = stream ? (hipStream_t)stream : c->stream;
        s2g_stats total;
        std::memset(&total, 0, sizeof(total));
        for (const std::vector<int> &grp : batch_groups(maps, nmaps)) {
            // reduced maps deposit into context-owned flat images first
            double *img = nullptr;
            const int64_t flat = batch_flat_elems(maps, grp);
            if (flat)
                img = (double *)c->image.get(sizeof(double) * flat);
            std::vector<MapIn> in = batch_inputs(dtype, maps, grp, img);
            s2g_stats st;
            deposit2d_pass(c, in.data(), (int)in.size(), s, &st);
            total.ms_device += st.ms_device;
            total.ms_setup += st.ms_setup;
            total.ms_norm += st.ms_norm;
            total.ms_sort += st.ms_sort;
            total.ms_gather += st.ms_gather;
            total.ms_finalize += st.ms_finalize;
            total.ms_small += st.ms_small;
            total.n_particles += st.n_particles;
            total.n_skipped += st.n_skipped;
            total.n_small += st.n_small;
            total.n_large += st.n_large;
            total.n_large_exact += st.n_large_exact;
            total.n_pairs += st.n_pairs;
            total.n_gather_launches += st.n_gather_launches;
        }
        if (stats)
            *stats = total;
        return S2G_OK;
    });
}

int s2g_plan_create(s2g_ctx *c, int dtype, const s2g_map_desc *maps, int nmaps, double headroom,
                    s2g_plan **out)
{
    return guarded(c, [&]() {
        if (!out)
            throw Error(S2G_EINVAL, "out is NULL");
        *out = nullptr;
        check_batch(dtype, maps, nmaps);
        if (nmaps < 1)
            throw Error(S2G_EINVAL, "a plan needs at least one map");
        if (!(headroom >= 1.0 && headroom <= 64.0))
            throw Error(S2G_EINVAL, "headroom out of range [1, 64]");
        std::unique_ptr<s2g_plan> pl(new s2g_plan());
        pl->c = c;
        try {
            plan_build(c, pl.get(), dtype, maps, nmaps, headroom);
        } catch (...) {
            plan_release(pl.get());
            throw;
        }
        *out = pl.release();
        return S2G_OK;
    });
}

int s2g_plan_run(s2g_plan *pl, void *stream, s2g_stats *stats)
{
    if (!pl)
        return S2G_EINVAL;
    return guarded(pl->c, [&]() {
        return plan_run(pl, stream ? (hipStream_t)stream : pl->c->stream, stats);
    });
}

void s2g_plan_destroy(s2g_plan *pl)
{
    if (!pl)
        return;
    int dev = -1;
    if (hipGetDevice(&dev) == hipSuccess && dev != pl->c->device)
        (void)hipSetDevice(pl->c->device);
    plan_release(pl);
    delete pl;
}

int s2g_footprints(s2g_ctx *c, int dims, int dtype, const void *pos, const void *hsml,
                   const void *m, const void *rho, const void *bin_q, int nq,
                   const void *weights, int64_t n, const s2g_params *params, int kernel,
                   uint32_t flags, int32_t *box, int32_t *cls_out, int32_t *fb_out,
                   int64_t *tile_off, int32_t *pairs_out, int64_t pair_cap, int64_t *small_off,
                   int32_t *small_order, int64_t *counts)
{
    return guarded(c, [&]() {
        check_common(dtype, kernel, dims == 3 ? 3 : 2, n, params);
        check_dims_nq(dims, nq);
        if (!box || !cls_out || !fb_out || !tile_off || !small_off || !small_order || !counts)
            throw Error(S2G_EINVAL, "output pointer is NULL");
        if (pair_cap > 0 && !pairs_out)
            throw Error(S2G_EINVAL, "pairs is NULL with pair_cap > 0");
        hipStream_t s = c->stream;
        const int64_t npix = params->npixels;
        Map2DArgs h = make_args(dtype, pos, hsml, m, rho, bin_q, nq, weights, n, params, kernel,
                                flags);
        {
            s2g_stats tmp;
            h = stage_inputs(c, h, s, &tmp);
        }
        ClassLists cl;
        std::vector<TileMap> tm;
        int ntiles = 0, nstiles = 1;
        const void *rec = nullptr;
        const int32_t *sorted = nullptr;
        int64_t *soff = nullptr;
        if (dims == 2) {
            MapIn mi{h, nullptr, nullptr};
            PassTables pt = pass_tables(c, &mi, 1, s);
            tm = pt.tm;
            ntiles = pt.ntiles;
            nstiles = pt.nstiles;
            if (n > 0) {
                Rec2 *r2 = (Rec2 *)c->rec.get(sizeof(Rec2) * n);
                int32_t *cl2 = (int32_t *)c->cls.get(sizeof(int32_t) * n);
                launch_2d_setup(dtype, pt.d_maps, 1, n, nq, kernel, r2, cl2, s);
                cl = partition_classes(c, cl2, n, pt.d_pbase, 1, s);
                rec = r2;
                if (cl.ns > 0)
                    sorted = bucket_small(c, cl, r2, pt.d_maps, nstiles, &soff, s);
            }
        } else {
            const int nb = (int)((npix + kBT - 1) / kBT);
            ntiles = nb * nb * nb;
            tm.push_back(TileMap{0, 0, ntiles, nb, nb, 0, 0});
            if (n > 0) {
                Rec3 *r3 = (Rec3 *)c->rec.get(sizeof(Rec3) * n);
                int32_t *cl3 = (int32_t *)c->cls.get(sizeof(int32_t) * n);
                int64_t *hpb = (int64_t *)c->h_tab.get(2 * sizeof(int64_t));
                hpb[0] = 0;
                hpb[1] = n;
                int64_t *d_pbase = (int64_t *)c->tab.get(2 * sizeof(int64_t));
                S2G_HIP(hipMemcpyAsync(d_pbase, hpb, 2 * sizeof(int64_t), hipMemcpyHostToDevice,
                                       s));
                launch_3d_setup(h, r3, cl3, s);
                cl = partition_classes(c, cl3, n, d_pbase, 1, s);
                rec = r3;
                sorted = cl.l0;  // the 3D small path keeps particle order
            }
        }
        // large particles: per-tile pair lists
        int64_t npairs = 0;
        std::vector<int64_t> cnt;
        if (cl.nl > 0) {
            double ms = 0.0;
            const TilePlan tp = build_tile_lists(c, cl.l1, rec, dims == 3, tm, cl.lbase,
                                                 cl.d_lbase, ntiles, s, &ms);
            npairs = tp.npairs;
            const int64_t ncnt = (int64_t)ntiles * tm[0].nch + 1;
            cnt.resize(ncnt);
            d2h(c, cnt.data(), c->cnt.p, sizeof(int64_t) * ncnt, s, nullptr);
            if (npairs > 0 && npairs <= pair_cap)
                d2h(c, pairs_out, tp.pairs, sizeof(int32_t) * npairs, s, nullptr);
        }
        for (int t = 0; t < ntiles; ++t)
            tile_off[t] = cl.nl > 0 ? cnt[(size_t)t * tm[0].nch] : 0;
        tile_off[ntiles] = npairs;
        // small particles: buckets (2D) or particle order (3D)
        if (cl.ns > 0) {
            d2h(c, small_order, sorted, sizeof(int32_t) * cl.ns, s, nullptr);
            if (dims == 2) {
                d2h(c, small_off, soff, sizeof(int64_t) * (nstiles + 1), s, nullptr);
            } else {
                small_off[0] = 0;
                small_off[1] = cl.ns;
            }
        } else {
            for (int t = 0; t <= nstiles; ++t)
                small_off[t] = 0;
        }
        // per-particle records
        if (n > 0) {
            const int nb = dims == 2 ? 4 : 6;
            if (dims == 2) {
                std::vector<Rec2> r((size_t)n);
                d2h(c, r.data(), rec, sizeof(Rec2) * n, s, nullptr);
                for (int64_t p = 0; p < n; ++p) {
                    const int32_t b[4] = {r[p].i0, r[p].i1, r[p].j0, r[p].j1};
                    std::memcpy(box + nb * p, b, sizeof(b));
                    cls_out[p] = r[p].cls;
                    fb_out[p] = r[p].fb;
                }
            } else {
                std::vector<Rec3> r((size_t)n);
                d2h(c, r.data(), rec, sizeof(Rec3) * n, s, nullptr);
                for (int64_t p = 0; p < n; ++p) {
                    const int32_t b[6] = {r[p].i0, r[p].i1, r[p].j0, r[p].j1, r[p].k0, r[p].k1};
                    std::memcpy(box + nb * p, b, sizeof(b));
                    cls_out[p] = r[p].cls;
                    fb_out[p] = r[p].fb;
                }
            }
        }
        counts[0] = cl.ns;
        counts[1] = cl.nl;
        counts[2] = cl.ne;
        counts[3] = npairs;
        counts[4] = ntiles;
        counts[5] = dims == 2 ? nstiles : 1;
        return S2G_OK;
    });
}

int s2g_accumulate_device(s2g_ctx *c, const double *d_src, double *d_dst, int64_t count,
                          void *stream)
{
    return guarded(c, [&]() {
        if (count < 0)
            throw Error(S2G_EINVAL, "count < 0");
        if (count > 0 && (!d_src || !d_dst))
            throw Error(S2G_EINVAL, "array pointer is NULL");
        hipStream_t s = stream ? (hipStream_t)stream : c->stream;
        launch_accumulate(d_src, d_dst, count, s);
        S2G_HIP(hipGetLastError());
        return S2G_OK;
    });
}

int s2g_healpix_reduce_device(s2g_ctx *c, double *d_map, const double *d_weights, int64_t npix,
                              void *stream)
{
    return guarded(c, [&]() {
        if (npix < 0)
            throw Error(S2G_EINVAL, "npix < 0");
        if (npix > 0 && (!d_map || !d_weights))
            throw Error(S2G_EINVAL, "map pointer is NULL");
        hipStream_t s = stream ? (hipStream_t)stream : c->stream;
        launch_hp_reduce(d_map, d_weights, npix, s);
        S2G_HIP(hipGetLastError());
        return S2G_OK;
    });
}

int s2g_mass_density_device(s2g_ctx *c, int dtype, const void *d_pos, const void *d_mass,
                            int64_t n, int kernel, int kernel_dim, int64_t nneighbors,
                            const double *boxsize, double *d_rho_out, double *d_hsml_out,
                            void *stream, s2g_stats *stats)
{
    return guarded(c, [&]() {
        hipStream_t s = stream ? (hipStream_t)stream : c->stream;
        mass_density_device(c, dtype, d_pos, d_mass, n, kernel, kernel_dim, nneighbors, boxsize,
                            d_rho_out, d_hsml_out, s, stats);
        return S2G_OK;
    });
}

int s2g_mass_density(s2g_ctx *c, int dtype, const void *pos, const void *mass, int64_t n,
                     int kernel, int kernel_dim, int64_t nneighbors, const double *boxsize,
                     double *rho_out, double *hsml_out, s2g_stats *stats)
{
    return guarded(c, [&]() {
        if (!pos || !mass || !rho_out || !hsml_out)
            throw Error(S2G_EINVAL, "array pointer is NULL");
        if (n < 0 || n > (int64_t)INT32_MAX)
            throw Error(S2G_EINVAL, "particle count out of range [0, 2^31)");
        const size_t es = elem_size(dtype);
        hipStream_t s = c->stream;
        void *dp = c->in[0].get(es * 3 * (n ? n : 1));
        void *dm = c->in[1].get(es * (n ? n : 1));
        double *dr = (double *)c->out.get(sizeof(double) * 2 * (n ? n : 1));
        double h2d = 0.0;
        S2G_HIP(hipEventRecord(c->ev0, s));
        S2G_HIP(hipMemcpyAsync(dp, pos, es * 3 * n, hipMemcpyHostToDevice, s));
        S2G_HIP(hipMemcpyAsync(dm, mass, es * n, hipMemcpyHostToDevice, s));
        S2G_HIP(hipEventRecord(c->ev1, s));
        S2G_HIP(hipEventSynchronize(c->ev1));
        h2d = elapsed(c->ev0, c->ev1);
        mass_density_device(c, dtype, dp, dm, n, kernel, kernel_dim, nneighbors, boxsize, dr,
                            dr + n, s, stats);
        double t1 = 0.0, t2 = 0.0;
        d2h(c, rho_out, dr, sizeof(double) * n, s, &t1);
        d2h(c, hsml_out, dr + n, sizeof(double) * n, s, &t2);
        if (stats) {
            stats->ms_h2d = h2d;
            stats->ms_d2h = t1 + t2;
        }
        return S2G_OK;
    });
}

}  // extern "C"
