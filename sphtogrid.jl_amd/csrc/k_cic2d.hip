// k_cic2d.hip -- the 2D Smac pixel-overlap deposit (cic_interpolation/cic_2D.jl)
// on gfx950.
//
// Pipeline of one pass (all on one stream).  A pass deposits one map or a
// batch of maps; the maps' particles and tiles are concatenated and every
// stage reads the per-map table (Map2D) where it needs map data.
//   k2d_setup        per particle, FP64: pixel coordinates, integer footprint,
//                    zero-quantity skip, fallback decision, class
//   k2d_partition    stable (particle-order) lists of small / large / exact
//   k2d_norm         explicit FP64 normalisation sum of large clipped footprints
//                    (closed form minus the border-cut part when smaller)
//   k_pair_count     counting sort of (tile, particle) pairs: per-chunk counts
//   (scan)           tile-major exclusive scan -> deterministic list offsets
//   k_pair_fill      one wave per chunk fills the per-tile lists in list order
//   k_item_*         split every tile list into work items of <= G pairs
//   k2d_gather       per-tile gather, FP64 register accumulation per pixel
//                    (k2d_gather_wide for three or four planes)
//   k2d_finalize     fixed-order sum of a tile's partials -> image (stored, or added
//                    to the small scatter's sums when that ran first)
//   k2d_small_*      small footprints: sorted by tile, LDS accumulation per item,
//                    one atomic add per touched pixel into the image
//   k2d_reduce_maps  reduce_image_2D: divide by the weight plane + transpose
#include "k_cic2d.h"

namespace s2g {

// ---------------------------------------------------------------------------
// setup: get_quantities_2D (cic_2D.jl:80-91), get_xyz (cic_shared.jl:85-100),
// pix_index_min_max (cic_shared.jl:46-52) and the fallback test of
// calculate_weights (cic_2D.jl:49-69), all in FP64 without FMA contraction.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k2d_setup(const Map2D *__restrict__ maps, int nmaps,
                                                 int64_t ntot, int nq, int kernel,
                                                 Rec2 *__restrict__ rec,
                                                 int32_t *__restrict__ cls)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= ntot)
        return;
    const int mi = find_entry(maps, nmaps, p, [](const Map2D &e) { return e.pbase; });
    const Map2D &md = maps[mi];
    const int64_t pl = p - md.pbase;
    const T *pos = (const T *)md.pos, *hsml = (const T *)md.hsml, *m = (const T *)md.m;
    const T *rho = (const T *)md.rho, *binq = (const T *)md.binq, *w = (const T *)md.w;
    const int64_t npix = md.npix;
    const double len2pix = md.len2pix;
    const uint32_t flags = md.flags;
    Rec2 r;
    r.x = r.y = r.h = r.s = 0.0;
    r.i0 = r.j0 = 0;
    r.i1 = r.j1 = -1;
    r.cls = kClsSkip;
    r.fb = 0;
    r.map = mi;

    // cic_2D.jl:155-168: skip if every quantity is zero and !calc_mean
    bool allzero = true;
    for (int q = 0; q < nq; ++q)
        if (ld(binq, pl * nq + q) != 0.0)
            allzero = false;
    if (allzero && !(flags & S2G_FLAG_CALC_MEAN)) {
        rec[p] = r;
        cls[p] = r.cls;
        return;
    }

    const double h = mul_rn(ld(hsml, pl), len2pix);
    const double area = mul_rn(2.0 * h, 2.0 * h);
    const double l3 = mul_rn(mul_rn(len2pix, len2pix), len2pix);
    const double rh = mul_rn(ld(rho, pl), 1.0 / l3);
    const double dz = (ld(m, pl) / rh) / area;
    const double wl = ld(w, pl);
    const double half = 0.5 * (double)npix;
    const double x = add_rn(mul_rn(ld(pos, 3 * pl + 0), len2pix), half);
    const double y = add_rn(mul_rn(ld(pos, 3 * pl + 1), len2pix), half);

    const double lim = (double)npix + 1.0;
    double xa = floor(sub_rn(x, h)), xb = floor(add_rn(x, h));
    double ya = floor(sub_rn(y, h)), yb = floor(add_rn(y, h));
    if (!(xa == xa) || !(xb == xb) || !(ya == ya) || !(yb == yb) || !(h > 0.0)) {
        rec[p] = r;  // NaN input or h <= 0: nothing is deposited
        cls[p] = r.cls;
        return;
    }
    xa = fmin(fmax(xa, -1.0), lim);
    xb = fmin(fmax(xb, -1.0), lim);
    ya = fmin(fmax(ya, -1.0), lim);
    yb = fmin(fmax(yb, -1.0), lim);
    const int64_t a0 = (int64_t)xa, a1 = (int64_t)xb, b0 = (int64_t)ya, b1 = (int64_t)yb;
    const int64_t i0 = a0 > 0 ? a0 : 0, i1 = a1 < npix - 1 ? a1 : npix - 1;
    const int64_t j0 = b0 > 0 ? b0 : 0, j1 = b1 < npix - 1 ? b1 : npix - 1;
    if (i0 > i1 || j0 > j1) {
        rec[p] = r;
        cls[p] = r.cls;
        return;
    }

    // Fallback iff no footprint pixel centre has u < 1 (u from the
    // reference's FP64 expression).  u is monotone in |x - i - 1/2| and
    // |y - j - 1/2|, so the minimum sits at the nearest column and row; three
    // candidates per axis cover ties and clamping.
    const double h_inv = 1.0 / h;
    double bx = 1e300, by = 1e300;
    {
        const double fx = floor(x), fy = floor(y);
        for (int d = -1; d <= 1; ++d) {
            double ci = fmin(fmax(fx + d, (double)i0), (double)i1);
            double cj = fmin(fmax(fy + d, (double)j0), (double)j1);
            bx = fmin(bx, fabs(coff1(x, ci)));
            by = fmin(by, fabs(coff1(y, cj)));
        }
    }
    const double umin = mul_rn(sqrt(add_rn(mul_rn(bx, bx), mul_rn(by, by))), h_inv);
    r.fb = !(umin < 1.0);

    r.x = x;
    r.y = y;
    r.h = h;
    r.i0 = (int32_t)i0;
    r.i1 = (int32_t)i1;
    r.j0 = (int32_t)j0;
    r.j1 = (int32_t)j1;
    const double num = mul_rn(mul_rn(area, wl), dz);
    const bool small = (i1 - i0 + 1) <= md.small_side && (j1 - j0 + 1) <= md.small_side;
    if (small) {
        r.cls = kClsSmall;
        r.s = num;
    } else {
        const bool clipped = a0 < 0 || a1 > npix - 1 || b0 < 0 || b1 > npix - 1;
        if (!(flags & S2G_FLAG_EXACT_NORM) && !clipped && !r.fb &&
            h >= closed_form_min_h(kernel)) {
            r.cls = kClsLargeAn;
            r.s = num / (mul_rn(h, h) * shape_disc_integral(kernel));
        } else {
            r.cls = kClsLargeEx;
            r.s = num;
        }
    }
    rec[p] = r;
    cls[p] = r.cls;
}

// ---------------------------------------------------------------------------
// Pixel weights shared by the normalisation sum and the tile gather: both
// form s = (x offset^2 + y offset^2) / h^2 with the same FP32 operations and
// evaluate ax * shape(sqrt(s)) with shape_row (s2g_internal.h), so a
// particle's deposited weights and its normalisation sum come from the same
// per-pixel values.
//
// Explicit normalisation for large footprints that are clipped by the image
// border, below the closed-form threshold or with S2G_FLAG_EXACT_NORM: one
// wave per particle, deterministic FP64 sum (fixed per-lane order + fixed
// butterfly).
// ---------------------------------------------------------------------------
constexpr int kNormThreads = 256;

// This lane's share of sum over rows [ia, ib] x columns [ja, jb] (any
// integers: pixels outside the image follow the same formulas) of
// shape_row(u, ax^(1/n)) * ay in FP64.  Lanes take the columns; rows are
// taken 64 at a time, lane l computing the terms of row ic + l once and the
// row loop broadcasting them with v_readlane.  With `disc`, each 64-column
// chunk runs only over the rows that can reach the kernel disc from it
// (|x - i - 1/2| < sqrt(h^2 - dy^2) + 1, dy the chunk's nearest column-centre
// distance): the rows left out have u >= 1, so the sum is unchanged.
template <int K>
__device__ __forceinline__ double norm_rect(const Rec2 &r, double hinv, bool disc, int64_t ia,
                                            int64_t ib, int64_t ja, int64_t jb, int ln)
{
    double acc = 0.0;
    for (int64_t jc = ja; jc <= jb; jc += kWave) {
        int64_t ra = ia, rb = ib;
        if (disc) {
            const float xc = (float)r.x, yc = (float)r.y, hf = (float)r.h;
            const int64_t jh = jc + kWave - 1 < jb ? jc + kWave - 1 : jb;
            const float dy = fmaxf(0.0f, fmaxf((float)jc + 0.5f - yc, yc - ((float)jh + 0.5f)));
            if (!(dy < hf + 1.0f))
                continue;
            const float hx = sqrtf(fmaxf(hf * hf - dy * dy, 0.0f)) + 1.0f;
            const int64_t da = (int64_t)floorf(xc - 0.5f - hx), db = (int64_t)ceilf(xc - 0.5f + hx);
            ra = da > ra ? da : ra;
            rb = db < rb ? db : rb;
        }
        const int64_t j = jc + ln;
        const double fj = (double)j;
        const float ys = (float)(coff1(r.y, fj) * hinv);
        const float yd2 = fmul_rn(ys, ys);  // rounded, as the gather's column table
        const float ay = j <= jb ? fmaxf((float)overlap1(r.y, r.h, fj), 0.0f) : 0.0f;
        double col = 0.0;
        for (int64_t ic = ra; ic <= rb; ic += kWave) {
            const double fi = (double)(ic + ln);
            const float xs = (float)(coff1(r.x, fi) * hinv);
            const float rx = xs * xs;
            const float a = overlap_root<K>((float)overlap1(r.x, r.h, fi));
            const int nrow = rb - ic + 1 < kWave ? (int)(rb - ic + 1) : kWave;
            for (int rr = 0; rr < nrow; ++rr) {
                const float rxr = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rx), rr));
                const float ar = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(a), rr));
                col += (double)shape_row<K>(__builtin_amdgcn_sqrtf(rxr + yd2), ar);
            }
        }
        acc += col * (double)ay;
    }
    return acc;
}

// Sum over the footprint in the image.  When the closed form h^2 I_k holds for
// the unclipped footprint (h >= the kernel's threshold, no fallback, no
// S2G_FLAG_EXACT_NORM), the particle's centre is in the image and the part cut
// off by the border is the smaller one, the sum is h^2 I_k minus the sum over
// that part (the border rows, then the border columns of the kept rows);
// otherwise the footprint in the image is summed directly.
template <int K>
__device__ __forceinline__ void k2d_norm_one(int64_t gw, const int32_t *__restrict__ list,
                                             const Map2D *__restrict__ maps,
                                             Rec2 *__restrict__ rec)
{
    const int ln = threadIdx.x % kWave;
    const int p = list[gw];
    const Rec2 r = rec[p];
    const double hinv = r.fb ? 0.0 : 1.0 / r.h;
    const int64_t i0 = r.i0, i1 = r.i1, j0 = r.j0, j1 = r.j1;
    bool cut = false;
    double A0 = 0.0, A1 = 0.0, B0 = 0.0, B1 = 0.0;
    if (!r.fb && r.h >= closed_form_min_h(K) && !(maps[r.map].flags & S2G_FLAG_EXACT_NORM)) {
        const double np = (double)maps[r.map].npix;
        A0 = floor(sub_rn(r.x, r.h));
        A1 = floor(add_rn(r.x, r.h));
        B0 = floor(sub_rn(r.y, r.h));
        B1 = floor(add_rn(r.y, r.h));
        const double kept = (double)(i1 - i0 + 1) * (double)(j1 - j0 + 1);
        const double full = (A1 - A0 + 1.0) * (B1 - B0 + 1.0);
        cut = r.x >= 0.0 && r.x < np && r.y >= 0.0 && r.y < np && full - kept < kept;
    }
    double acc;
    if (cut) {
        const int64_t a0 = (int64_t)A0, a1 = (int64_t)A1, b0 = (int64_t)B0, b1 = (int64_t)B1;
        double out = 0.0;
        if (a0 < i0)
            out += norm_rect<K>(r, hinv, true, a0, i0 - 1, b0, b1, ln);
        if (a1 > i1)
            out += norm_rect<K>(r, hinv, true, i1 + 1, a1, b0, b1, ln);
        if (b0 < j0)
            out += norm_rect<K>(r, hinv, true, i0, i1, b0, j0 - 1, ln);
        if (b1 > j1)
            out += norm_rect<K>(r, hinv, true, i0, i1, j1 + 1, b1, ln);
        out = wave_sum(out);
        acc = mul_rn(r.h, r.h) * shape_disc_integral(K) - out;
    } else {
        acc = wave_sum(norm_rect<K>(r, hinv, !r.fb, i0, i1, j0, j1, ln));
    }
    if (ln == 0)
        rec[p].s = acc != 0.0 ? r.s / acc : 0.0;
}

template <int K>
// d_bounds (null: ne as given): the class bounds on the device, ne =
// d_bounds[3] - d_bounds[2]
__global__ __launch_bounds__(kNormThreads) void k2d_norm(const int32_t *__restrict__ list,
                                                         int64_t ne,
                                                         const int64_t *__restrict__ d_bounds,
                                                         const Map2D *__restrict__ maps,
                                                         Rec2 *__restrict__ rec)
{
    if (d_bounds)
        ne = d_bounds[3] - d_bounds[2];
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x / kWave);
    for (int64_t gw = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; gw < ne;
         gw += nw)
        k2d_norm_one<K>(gw, list, maps, rec);
}


// ---------------------------------------------------------------------------
// Tile gather.  Block = 4 waves = one 64x64 tile.  Wave w owns the 16-column
// strip [16w, 16w+16) of the tile; lane l works on column 16w + (l & 15) and
// rows 4k + (l >> 4), k = 0..15, so each lane accumulates 16 pixels in FP64
// registers.  Per batch of particles the row terms (x offset^2, x overlap) of
// all 64 tile rows are tabulated once in LDS; per particle each lane computes
// its column terms, and waves skip particles that miss their strip or row
// groups (wave-uniform scalar tests).  Per pixel: one sqrt, the kernel
// polynomial with the row overlap folded into its (1 - u) factor (shape_row),
// and NP FP64 FMAs (the plane multiplier already carries the particle's scale
// and the lane's column overlap).
// ---------------------------------------------------------------------------
constexpr int kGBatch = 32;
constexpr int kGStep = 4;  // row groups per skip test (measured best of 1, 2, 4, 8, 16)

// A: the accumulator type, double, or float with S2G_FLAG_FP32_ACCUM (an
// item's pixel sums in FP32, written to its FP64 partials)
// d_nitems (null in a call sized on the host): the item count on the device,
// the grid being its upper bound; d_flag != 0 (a replay whose pair lists
// outgrew their capacity): no item runs
template <typename T, int K, int NP, typename A>
__device__ __forceinline__ void gather_tile(const int4 *__restrict__ items,
                                            const int32_t *__restrict__ pairs,
                                            const Rec2 *__restrict__ rec,
                                            const Map2D *__restrict__ maps, int nq, int plane0,
                                            double *__restrict__ partial,
                                            const int64_t *__restrict__ d_nitems,
                                            const int *__restrict__ d_flag)
{
    if (d_nitems && ((int64_t)blockIdx.x >= *d_nitems || *d_flag))
        return;  // the whole block
    __shared__ float2 s_row[kGBatch][kGT];
    __shared__ float2 s_col[kGBatch][kGT];
    __shared__ double s_x[kGBatch], s_y[kGBatch], s_h[kGBatch];
    __shared__ float s_hinv[kGBatch];
    __shared__ double s_hinvd[kGBatch];
    __shared__ int s_i0[kGBatch], s_i1[kGBatch], s_j0[kGBatch], s_j1[kGBatch];
    __shared__ double s_mult[NP][kGBatch];
    __shared__ int s_k[kGBatch][4];  // row-group range per (particle, strip), -1: none

    const int4 it = items[blockIdx.x];
    const int ntj = maps[it.w].ntj;
    const int tile = it.x - maps[it.w].tbase;
    const T *binq = (const T *)maps[it.w].binq;
    const int64_t pbase = maps[it.w].pbase;
    const int ti0 = (tile / ntj) * kGT, tj0 = (tile % ntj) * kGT;
    const int wv = threadIdx.x / kWave, ln = threadIdx.x % kWave;
    const int rg = ln >> 4;

    A acc[kGRows][NP];
#pragma unroll
    for (int k = 0; k < kGRows; ++k)
#pragma unroll
        for (int q = 0; q < NP; ++q)
            acc[k][q] = A(0);

    for (int base = 0; base < it.z; base += kGBatch) {
        const int cnt = it.z - base < kGBatch ? it.z - base : kGBatch;
        if (threadIdx.x < cnt) {
            const int p = pairs[it.y + base + threadIdx.x];
            const Rec2 r = rec[p];
            s_x[threadIdx.x] = r.x;
            s_y[threadIdx.x] = r.y;
            s_h[threadIdx.x] = r.h;
            const double hi = r.fb ? 0.0 : 1.0 / r.h;
            s_hinv[threadIdx.x] = (float)hi;
            s_hinvd[threadIdx.x] = hi;
            s_i0[threadIdx.x] = r.i0;
            s_i1[threadIdx.x] = r.i1;
            s_j0[threadIdx.x] = r.j0;
            s_j1[threadIdx.x] = r.j1;
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                const int P = plane0 + q;
                const double mult =
                    P == nq ? 1.0 : ld(binq, (int64_t)(p - pbase) * nq + (P < nq ? P : 0));
                s_mult[q][threadIdx.x] = mult * r.s;
            }
        }
        __syncthreads();
        // row and column tables of the batch for the 64 tile rows and
        // columns: ((x offset / h)^2, x overlap^(1/n)) and
        // ((y offset / h)^2, y overlap); offset / h is formed in FP64 and
        // rounded once, so no rounding of 1/h scales a whole footprint
        for (int e = threadIdx.x; e < cnt * kGT; e += 256) {
            const int b = e / kGT, rr = e % kGT;
            const double fi = (double)(ti0 + rr), fj = (double)(tj0 + rr);
            const float xs = (float)(coff1(s_x[b], fi) * s_hinvd[b]);
            const float a = overlap_root<K>((float)overlap1(s_x[b], s_h[b], fi));
            s_row[b][rr] = make_float2(xs * xs, a);
            const float ys = (float)(coff1(s_y[b], fj) * s_hinvd[b]);
            const float ay = fmaxf((float)overlap1(s_y[b], s_h[b], fj), 0.0f);
            s_col[b][rr] = make_float2(ys * ys, ay);
        }
        // row groups each wave's strip needs for each particle: the footprint
        // box, narrowed to the rows that can reach the kernel disc from the
        // strip's 16 columns, |x - i - 1/2| < sqrt(h^2 - dy^2) + 1/4 with dy
        // the smallest distance of a column centre from y.  The quarter-pixel
        // margin keeps every skipped pixel at u >= 1 up to FP32 rounding
        // (W = 0, or below 1e-20 relative).
        if (threadIdx.x < cnt * 4) {
            const int b = threadIdx.x >> 2, w = threadIdx.x & 3;
            const int c0 = tj0 + 16 * w;
            int packed = -1;
            if (!(c0 + 15 < s_j0[b] || c0 > s_j1[b])) {
                const int i0 = s_i0[b] - ti0, i1 = s_i1[b] - ti0;
                int klo = i0 > 0 ? (i0 >> 2) : 0;
                int khi = (i1 >> 2) < kGRows - 1 ? (i1 >> 2) : kGRows - 1;
                if (s_hinv[b] != 0.0f) {
                    const float xc = (float)s_x[b], yc = (float)s_y[b], hf = (float)s_h[b];
                    const float dy = fmaxf(0.0f, fmaxf((float)c0 + 0.5f - yc,
                                                       yc - ((float)c0 + 15.5f)));
                    if (dy < hf + 0.25f) {
                        const float hx = sqrtf(fmaxf(hf * hf - dy * dy, 0.0f)) + 0.25f;
                        const int ra = (int)floorf(xc - 0.5f - hx) - ti0;
                        const int rb = (int)ceilf(xc - 0.5f + hx) - ti0;
                        klo = (ra >> 2) > klo ? (ra >> 2) : klo;
                        khi = (rb >> 2) < khi ? (rb >> 2) : khi;
                    } else {
                        khi = -1;
                    }
                }
                if (klo <= khi)
                    packed = klo | (khi << 8);
            }
            s_k[b][w] = packed;
        }
        __syncthreads();
        for (int b = 0; b < cnt; ++b) {
            const int kk = __builtin_amdgcn_readfirstlane(s_k[b][wv]);
            if (kk < 0)
                continue;
            const int klo = kk & 0xff, khi = kk >> 8;
            const float2 cv = s_col[b][16 * wv + (ln & 15)];
            const float yd2 = cv.x;
            // the lane's column overlap is common to its 16 pixels: fold it
            // into the FP64 plane multipliers
            const double ayd = (double)cv.y;
            A mult[NP];
#pragma unroll
            for (int q = 0; q < NP; ++q)
                mult[q] = (A)(s_mult[q][b] * ayd);
            // kGStep row groups per wave-uniform test, so their LDS reads and
            // dependency chains overlap.  A group outside [klo, khi] adds
            // exact zeros (a = 0 outside the footprint, u >= 1 outside the
            // row range), so the image is the same as with one group per test
            // (evaluating only the groups in the range, with the chunks at
            // the range's ends taken one group at a time, was slower: 272 ->
            // 314 ms on the cluster set).
#pragma unroll
            for (int k0 = 0; k0 < kGRows; k0 += kGStep) {
                if (k0 + kGStep - 1 < klo || k0 > khi)
                    continue;
                float wk[kGStep];
#pragma unroll
                for (int d = 0; d < kGStep; ++d) {
                    const float2 rvk = s_row[b][4 * (k0 + d) + rg];
                    wk[d] = shape_row<K>(__builtin_amdgcn_sqrtf(rvk.x + yd2), rvk.y);
                }
#pragma unroll
                for (int d = 0; d < kGStep; ++d) {
                    const A f = (A)wk[d];
#pragma unroll
                    for (int q = 0; q < NP; ++q)
                        acc[k0 + d][q] = fma(f, mult[q], acc[k0 + d][q]);
                }
            }
        }
        __syncthreads();
    }
    double *dst = partial + (int64_t)blockIdx.x * NP * kGT * kGT;
#pragma unroll
    for (int q = 0; q < NP; ++q)
#pragma unroll
        for (int k = 0; k < kGRows; ++k)
            dst[(q * kGT + 4 * k + rg) * kGT + 16 * wv + (ln & 15)] = (double)acc[k][q];
}

// up to two planes: 128 VGPRs, four waves per SIMD (the accumulators need
// 64); three or four planes: 168 VGPRs, three waves per SIMD (four planes
// spill a few registers, which costs less than the lost wave: the batched
// 64-map set's gather 85 -> 73 ms).  Tried and dropped: eight waves per
// block with half the tile's rows per wave, so that four planes fit 128
// VGPRs at four waves per SIMD: the batched set's gather 70.6 -> 81.5 ms
// (each wave repeats the per-particle set-up for half the pixels).
template <typename T, int K, int NP>
__global__ __launch_bounds__(256, 4) void k2d_gather(const int4 *__restrict__ items,
                                                     const int32_t *__restrict__ pairs,
                                                     const Rec2 *__restrict__ rec,
                                                     const Map2D *__restrict__ maps, int nq,
                                                     int plane0, double *__restrict__ partial,
                                                     const int64_t *__restrict__ d_nitems,
                                                     const int *__restrict__ d_flag)
{
    gather_tile<T, K, NP, double>(items, pairs, rec, maps, nq, plane0, partial, d_nitems,
                                  d_flag);
}
template <typename T, int K, int NP>
__global__ __launch_bounds__(256, 3) void k2d_gather_wide(const int4 *__restrict__ items,
                                                       const int32_t *__restrict__ pairs,
                                                       const Rec2 *__restrict__ rec,
                                                       const Map2D *__restrict__ maps, int nq,
                                                       int plane0, double *__restrict__ partial,
                                                     const int64_t *__restrict__ d_nitems,
                                                     const int *__restrict__ d_flag)
{
    gather_tile<T, K, NP, double>(items, pairs, rec, maps, nq, plane0, partial, d_nitems,
                                  d_flag);
}
// FP32 accumulators: at most 64 VGPRs for four planes, four waves per SIMD
template <typename T, int K, int NP>
__global__ __launch_bounds__(256, 4) void k2d_gather_f32acc(const int4 *__restrict__ items,
                                                         const int32_t *__restrict__ pairs,
                                                         const Rec2 *__restrict__ rec,
                                                         const Map2D *__restrict__ maps, int nq,
                                                         int plane0, double *__restrict__ partial,
                                                         const int64_t *__restrict__ d_nitems,
                                                         const int *__restrict__ d_flag)
{
    gather_tile<T, K, NP, float>(items, pairs, rec, maps, nq, plane0, partial, d_nitems,
                                 d_flag);
}

// fixed-order sum of a tile's item partials -> image planes (plain stores).
// Every pixel of every tile of the pass is written (zero where no large
// particle reaches it), which initialises the images for the small scatter.
constexpr int kFinSplit = 16;

template <int NP>
// add: the image already holds the small scatter's sums (zeroed before it)
__global__ __launch_bounds__(256) void k2d_finalize(const double *__restrict__ partial,
                                                    const int64_t *__restrict__ tile_items,
                                                    const Map2D *__restrict__ maps, int nmaps,
                                                    int plane0, int add)
{
    // kFinSplit blocks per tile, 4096 / kFinSplit pixels each, so that a pass
    // of few tiles (a small map) still spreads the sums over the CUs
    const int t = blockIdx.x / kFinSplit;
    const int px0 = (blockIdx.x % kFinSplit) * (kGT * kGT / kFinSplit);
    const int mi = find_entry(maps, nmaps, t, [](const Map2D &e) { return e.tbase; });
    const int ntj = maps[mi].ntj;
    const int64_t npix = maps[mi].npix;
    double *image = maps[mi].image;
    const int tl = t - maps[mi].tbase;
    const int ti0 = (tl / ntj) * kGT, tj0 = (tl % ntj) * kGT;
    const int64_t b = tile_items ? tile_items[2 * t] : 0;
    const int64_t k = tile_items ? tile_items[2 * t + 1] : 0;
    const int64_t nd = npix * npix;
    for (int px = px0 + threadIdx.x; px < px0 + kGT * kGT / kFinSplit; px += 256) {
        const int i = ti0 + px / kGT, j = tj0 + px % kGT;
        if (i >= npix || j >= npix)
            continue;
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            double s = 0.0;
#pragma unroll 4
            for (int64_t it = 0; it < k; ++it)
                s += partial[((b + it) * NP + q) * kGT * kGT + px];
            double *dst = image + (int64_t)(plane0 + q) * nd + (int64_t)i * npix + j;
            *dst = add ? *dst + s : s;
        }
    }
}

// ---------------------------------------------------------------------------
// Small footprints (box side <= 34).  Particles are sorted by the 32 x 32
// tile that holds their footprint's first pixel; a block takes a work item of
// one tile, each wave one particle at a time: the FP64 normalisation sum over
// the footprint (wave-reduced), then the FP64 pixel weights added into the
// block's LDS window of the tile + 34 pixels (ds_add_f64).  The window is
// added to the image once per item (FP64 atomics, only non-zero cells).
// ---------------------------------------------------------------------------
// threads per block: 8 waves; 16 for three or four planes, whose window
// (104 or 139 KB of LDS) leaves room for one block per CU, so that the CU
// still runs four waves per SIMD
constexpr int small_threads(int np) { return np > 2 ? 1024 : 512; }
// particles whose records a block loads into LDS at a time
constexpr int kSmallBatch = 64;
// pixels per lane of the largest small footprint (34 x 34 over 64 lanes)
constexpr int kPerLane = (kSmallSide2D * kSmallSide2D + kWave - 1) / kWave;

__global__ void k2d_small_keys(const int32_t *__restrict__ list, int64_t ns,
                               const Rec2 *__restrict__ rec, const Map2D *__restrict__ maps,
                               uint32_t *__restrict__ keys, int32_t *__restrict__ vals)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ns)
        return;
    const int p = list[e];
    const Rec2 r = rec[p];
    const Map2D &md = maps[r.map];
    keys[e] = (uint32_t)(md.sbase + (r.i0 / kST) * md.nts + r.j0 / kST);
    vals[e] = p;
}

__global__ void k_tile_offsets(const uint32_t *__restrict__ keys, int64_t ns, int ntiles,
                               int64_t *__restrict__ off)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > ntiles)
        return;
    int64_t lo = 0, hi = ns;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (keys[mid] < (uint32_t)t)
            lo = mid + 1;
        else
            hi = mid;
    }
    off[t] = lo;
}

// One small particle on one wave: row / column terms in the wave's LDS
// tables (((x offset / h)^2, x overlap^(1/n)) and ((y offset / h)^2,
// y overlap), as the tile gather forms them), the pixel weights
// ay * shape_row(u, a) in FP32 kept in registers (at most KMAX per lane), the
// normalisation sum in FP64 reduced across the wave, then the weights times
// the particle's FP64 plane multipliers added into the block's LDS window
// (ds_add_f64).
// FP64 weights, computed per pixel without tables (two passes: the sum,
// then the deposit), for the footprints whose FP32 weights cannot resolve the
// reference's: h < 2 px, where every pixel centre inside the kernel may lie
// near its edge (1 - u of the order of the FP32 rounding of u), and the
// fallback to area weights.  Such footprints have at most 5 x 5 pixels.
template <int K, int NP>
__device__ __forceinline__ void small_particle64(const Rec2 &r, int e, int ln, int oi, int oj,
                                                 double (*s_mult)[kSmallBatch],
                                                 double (*acc)[kSW * kSW])
{
    const int jw = r.j1 - r.j0 + 1;
    const int fp = (r.i1 - r.i0 + 1) * jw;
    const double hinv = r.fb ? 0.0 : 1.0 / r.h;
    auto weight = [&](int f, int &cell) {
        const int di = f / jw, dj = f - di * jw;
        const double fi = (double)(r.i0 + di), fj = (double)(r.j0 + dj);
        const double xd = coff1(r.x, fi), yd = coff1(r.y, fj);
        const double u = sqrt(xd * xd + yd * yd) * hinv;
        cell = (r.i0 + di - oi) * kSW + (r.j0 + dj - oj);
        return shape_of<K, double>(u) * (overlap1(r.x, r.h, fi) * overlap1(r.y, r.h, fj));
    };
    double sum = 0.0;
    for (int f = ln; f < fp; f += kWave) {
        int c;
        sum += weight(f, c);
    }
    sum = wave_sum(sum);
    if (!(sum != 0.0))
        return;
    const double inv = 1.0 / sum;
    double mult[NP];
#pragma unroll
    for (int q = 0; q < NP; ++q)
        mult[q] = s_mult[q][e] * inv;
    for (int f = ln; f < fp; f += kWave) {
        int c;
        const double w = weight(f, c);
        if (w != 0.0) {
#pragma unroll
            for (int q = 0; q < NP; ++q)
                unsafeAtomicAdd(&acc[q][c], w * mult[q]);
        }
    }
}

// FP32 weights for h >= 2 px: the pixel holding the particle's centre has
// u <= 0.36, so the largest weights are well conditioned and the FP32
// rounding of u only moves weights far below them.  Returns false (nothing
// deposited) if the weight sum is below 1e-30, which leaves the particle to
// the FP64 path.
template <int K, int NP, int KMAX>
__device__ __forceinline__ bool small_particle(const Rec2 &r, int e, int ln, int oi, int oj,
                                               float2 *rt_tab, float2 *ct_tab,
                                               double (*s_mult)[kSmallBatch],
                                               double (*acc)[kSW * kSW])
{
    const int iw = r.i1 - r.i0 + 1, jw = r.j1 - r.j0 + 1;
    const int fp = iw * jw;
    const double hinv = r.fb ? 0.0 : 1.0 / r.h;
    if (ln < iw) {
        const double fi = (double)(r.i0 + ln);
        const float xs = (float)(coff1(r.x, fi) * hinv);
        rt_tab[ln] = make_float2(xs * xs, overlap_root<K>((float)overlap1(r.x, r.h, fi)));
    }
    if (ln < jw) {
        const double fj = (double)(r.j0 + ln);
        const float ys = (float)(coff1(r.y, fj) * hinv);
        ct_tab[ln] = make_float2(ys * ys, fmaxf((float)overlap1(r.y, r.h, fj), 0.0f));
    }
    const float invjw = 1.0f / (float)jw;
    const int cbase = (r.i0 - oi) * kSW + (r.j0 - oj);
    float wk[KMAX];
    int cell[KMAX];
    double sum = 0.0;
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        wk[k] = 0.0f;
        cell[k] = 0;
        const int f = ln + k * kWave;
        if (f < fp) {
            const int di = (int)(((float)f + 0.5f) * invjw);
            const int dj = f - di * jw;
            const float2 rt = rt_tab[di], ct = ct_tab[dj];
            wk[k] = shape_row<K>(__builtin_amdgcn_sqrtf(rt.x + ct.x), rt.y) * ct.y;
            cell[k] = cbase + di * kSW + dj;
            sum += (double)wk[k];
        }
    }
    sum = wave_sum(sum);
    if (!(sum >= 1e-30))
        return false;
    const double inv = 1.0 / sum;
    double mult[NP];
#pragma unroll
    for (int q = 0; q < NP; ++q)
        mult[q] = s_mult[q][e] * inv;
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        if (wk[k] != 0.0f) {
            const double w = (double)wk[k];
#pragma unroll
            for (int q = 0; q < NP; ++q)
                unsafeAtomicAdd(&acc[q][cell[k]], w * mult[q]);
        }
    }
    return true;
}

// Tried and dropped: FP32 windows (ds_add_f32, half the LDS, two 512-thread
// blocks per CU) with S2G_FLAG_FP32_ACCUM: the batched set's four-plane
// scatter went from 30 to 140 ms.
template <typename T, int K, int NP>
__global__ __launch_bounds__(small_threads(NP)) void k2d_small_tile(
    const int4 *__restrict__ items, const int64_t *__restrict__ nitems,
    const int32_t *__restrict__ plist, const Rec2 *__restrict__ rec,
    const Map2D *__restrict__ maps, int nq, int plane0)
{
    constexpr int kSmallThreads = small_threads(NP);
    __shared__ double acc[NP][kSW * kSW];
    // per-wave row / column terms of the current particle: (offset^2, overlap)
    __shared__ float2 s_rt[kSmallThreads / kWave][kSmallSide2D];
    __shared__ float2 s_ct[kSmallThreads / kWave][kSmallSide2D];
    // records and plane multipliers of a batch of particles, loaded together
    __shared__ Rec2 s_rec[kSmallBatch];
    __shared__ double s_mult[NP][kSmallBatch];
    if ((int64_t)blockIdx.x >= *nitems)
        return;
    const int4 it = items[blockIdx.x];
    const int nts = maps[it.w].nts;
    const int tl = it.x - maps[it.w].sbase;
    const int oi = (tl / nts) * kST, oj = (tl % nts) * kST;
    const int64_t npix = maps[it.w].npix, pbase = maps[it.w].pbase;
    const T *binq = (const T *)maps[it.w].binq;
    double *image = maps[it.w].image;
    for (int c = threadIdx.x; c < NP * kSW * kSW; c += kSmallThreads)
        (&acc[0][0])[c] = 0.0;
    const int wv = threadIdx.x / kWave, ln = threadIdx.x % kWave;
    for (int base = 0; base < it.z; base += kSmallBatch) {
        const int cnt = it.z - base < kSmallBatch ? it.z - base : kSmallBatch;
        __syncthreads();
        if (threadIdx.x < cnt) {
            const int p = plist[it.y + base + threadIdx.x];
            const Rec2 r = rec[p];
            s_rec[threadIdx.x] = r;
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                const int P = plane0 + q;
                s_mult[q][threadIdx.x] =
                    P == nq ? r.s : r.s * ld(binq, (int64_t)(p - pbase) * nq + (P < nq ? P : 0));
            }
        }
        __syncthreads();
        for (int e = wv; e < cnt; e += kSmallThreads / kWave) {
            const Rec2 r = s_rec[e];
            const int fp = (r.i1 - r.i0 + 1) * (r.j1 - r.j0 + 1);
            bool done = false;
            if (!r.fb && r.h >= 2.0) {
                // most footprints need at most 4 pixels per lane: a short
                // unrolled variant for them, the full one for the rest
                if (__builtin_amdgcn_readfirstlane(fp) <= 4 * kWave)
                    done = small_particle<K, NP, 4>(r, e, ln, oi, oj, s_rt[wv], s_ct[wv],
                                                    s_mult, acc);
                else
                    done = small_particle<K, NP, kPerLane>(r, e, ln, oi, oj, s_rt[wv],
                                                           s_ct[wv], s_mult, acc);
            }
            if (!done)
                small_particle64<K, NP>(r, e, ln, oi, oj, s_mult, acc);
        }
    }
    __syncthreads();
    const int64_t nd = npix * npix;
    for (int c = threadIdx.x; c < kSW * kSW; c += kSmallThreads) {
        const int i = oi + c / kSW, j = oj + c % kSW;
        if (i >= npix || j >= npix)
            continue;
        const int64_t idx = (int64_t)i * npix + j;
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            const double v = acc[q][c];
            if (v != 0.0)
                unsafeAtomicAdd(&image[(int64_t)(plane0 + q) * nd + idx], v);
        }
    }
}

// ---------------------------------------------------------------------------
// reduce_image_2D (reduce_image.jl:8-31): out[q][y][x] = img[q][x*N+y] / w.
// 64x64 tiles transposed through LDS.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void reduce_tile(const double *__restrict__ image, int64_t npix,
                                            int nq, int reduce, double *__restrict__ out,
                                            int ti, int tj, double (*tile)[65])
{
    const int64_t nd = npix * npix;
    const double *w = image + (int64_t)nq * nd;
    const int tx = threadIdx.x % 64, ty = threadIdx.x / 64;
    // weights of the 16 output elements this thread writes, transposed via LDS
    for (int r = ty; r < 64; r += 4) {
        int i = ti + r, j = tj + tx;
        tile[r][tx] = (i < npix && j < npix) ? w[(int64_t)i * npix + j] : 0.0;
    }
    __syncthreads();
    double wt[16];
#pragma unroll
    for (int k = 0; k < 16; ++k)
        wt[k] = tile[tx][ty + 4 * k];
    for (int q = 0; q < nq; ++q) {
        __syncthreads();
        for (int r = ty; r < 64; r += 4) {
            int i = ti + r, j = tj + tx;
            tile[r][tx] = (i < npix && j < npix) ? image[q * nd + (int64_t)i * npix + j] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int r = ty + 4 * k;
            const int j = tj + r, i = ti + tx;  // output row j (y), column i (x)
            if (i < npix && j < npix) {
                double v = tile[tx][r];
                if (reduce && wt[k] > 0.0)
                    v /= wt[k];
                out[q * nd + (int64_t)j * npix + i] = v;
            }
        }
    }
}

__global__ __launch_bounds__(256) void k2d_reduce(const double *__restrict__ image,
                                                  int64_t npix, int nq, int reduce,
                                                  double *__restrict__ out)
{
    __shared__ double tile[64][65];
    reduce_tile(image, npix, nq, reduce, out, blockIdx.y * 64, blockIdx.x * 64, tile);
}

// one block per gather tile of the pass (gather and reduce tiles are both
// 64 x 64); maps without an output are skipped
__global__ __launch_bounds__(256) void k2d_reduce_maps(const Map2D *__restrict__ maps,
                                                       int nmaps, int nq)
{
    __shared__ double tile[64][65];
    const int t = blockIdx.x;
    const int mi = find_entry(maps, nmaps, t, [](const Map2D &e) { return e.tbase; });
    const Map2D &md = maps[mi];
    if (!md.out)
        return;
    const int tl = t - md.tbase;
    reduce_tile(md.image, md.npix, nq, (md.flags & S2G_FLAG_REDUCE) ? 1 : 0, md.out,
                (tl / md.ntj) * kGT, (tl % md.ntj) * kGT, tile);
}

// ---------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------
#define S2G_K2D_SWITCH(kernel, CALL)                                                    \
    switch (kernel) {                                                                  \
    case S2G_KERNEL_CUBIC: { constexpr int KK = S2G_KERNEL_CUBIC; CALL; } break;       \
    case S2G_KERNEL_QUINTIC: { constexpr int KK = S2G_KERNEL_QUINTIC; CALL; } break;   \
    case S2G_KERNEL_WENDLAND_C2: { constexpr int KK = S2G_KERNEL_WENDLAND_C2; CALL; } break; \
    case S2G_KERNEL_WENDLAND_C4: { constexpr int KK = S2G_KERNEL_WENDLAND_C4; CALL; } break; \
    case S2G_KERNEL_WENDLAND_C6: { constexpr int KK = S2G_KERNEL_WENDLAND_C6; CALL; } break; \
    default: { constexpr int KK = S2G_KERNEL_WENDLAND_C8; CALL; } break;               \
    }

void launch_2d_setup(int dtype, const Map2D *maps, int nmaps, int64_t ntot, int nq, int kernel,
                     Rec2 *rec, int32_t *cls, hipStream_t s)
{
    if (ntot == 0)
        return;
    const dim3 g((unsigned)((ntot + 255) / 256)), b(256);
    if (dtype == S2G_F32)
        hipLaunchKernelGGL(k2d_setup<float>, g, b, 0, s, maps, nmaps, ntot, nq, kernel, rec, cls);
    else
        hipLaunchKernelGGL(k2d_setup<double>, g, b, 0, s, maps, nmaps, ntot, nq, kernel, rec,
                           cls);
}

void launch_2d_norm(int kernel, const int32_t *list, int64_t ne, const Map2D *maps, Rec2 *rec,
                    hipStream_t s, const int64_t *d_bounds)
{
    if (ne == 0)
        return;
    const dim3 g((unsigned)wave_blocks(ne, kNormThreads)), b(kNormThreads);
    S2G_K2D_SWITCH(kernel,
                   hipLaunchKernelGGL(k2d_norm<KK>, g, b, 0, s, list, ne, d_bounds, maps, rec));
}

template <typename T, int K, int NP>
static void gather_knp(const GatherArgs &g, hipStream_t s)
{
    if (g.fp32acc)
        hipLaunchKernelGGL((k2d_gather_f32acc<T, K, NP>), dim3((unsigned)g.nitems), dim3(256), 0,
                           s, g.items, g.pairs, g.rec, g.maps, g.nq, g.plane0, g.partial,
                           g.d_nitems, g.d_flag);
    else if constexpr (NP <= 2)
        hipLaunchKernelGGL((k2d_gather<T, K, NP>), dim3((unsigned)g.nitems), dim3(256), 0, s,
                           g.items, g.pairs, g.rec, g.maps, g.nq, g.plane0, g.partial,
                           g.d_nitems, g.d_flag);
    else
        hipLaunchKernelGGL((k2d_gather_wide<T, K, NP>), dim3((unsigned)g.nitems), dim3(256), 0,
                           s, g.items, g.pairs, g.rec, g.maps, g.nq, g.plane0,
                           g.partial, g.d_nitems, g.d_flag);
}

template <typename T, int K>
static void gather_k(const GatherArgs &g, hipStream_t s)
{
    switch (g.np) {
    case 1: gather_knp<T, K, 1>(g, s); break;
    case 2: gather_knp<T, K, 2>(g, s); break;
    case 3: gather_knp<T, K, 3>(g, s); break;
    default: gather_knp<T, K, 4>(g, s); break;
    }
}

void launch_2d_gather(const GatherArgs &g, hipStream_t s)
{
    if (g.dtype == S2G_F32) {
        S2G_K2D_SWITCH(g.kernel, (gather_k<float, KK>(g, s)));
    } else {
        S2G_K2D_SWITCH(g.kernel, (gather_k<double, KK>(g, s)));
    }
}

void launch_2d_finalize(const GatherArgs &g, hipStream_t s)
{
    const dim3 grid((unsigned)(g.ntiles * kFinSplit)), block(256);
    switch (g.np) {
    case 1: hipLaunchKernelGGL(k2d_finalize<1>, grid, block, 0, s, g.partial, g.tile_items, g.maps, g.nmaps, g.plane0, g.add); break;
    case 2: hipLaunchKernelGGL(k2d_finalize<2>, grid, block, 0, s, g.partial, g.tile_items, g.maps, g.nmaps, g.plane0, g.add); break;
    case 3: hipLaunchKernelGGL(k2d_finalize<3>, grid, block, 0, s, g.partial, g.tile_items, g.maps, g.nmaps, g.plane0, g.add); break;
    default: hipLaunchKernelGGL(k2d_finalize<4>, grid, block, 0, s, g.partial, g.tile_items, g.maps, g.nmaps, g.plane0, g.add); break;
    }
}

void launch_2d_small_keys(const int32_t *list, int64_t ns, const Rec2 *rec, const Map2D *maps,
                          uint32_t *keys, int32_t *vals, hipStream_t s)
{
    if (ns == 0)
        return;
    hipLaunchKernelGGL(k2d_small_keys, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, s, list,
                       ns, rec, maps, keys, vals);
}

void launch_tile_offsets(const uint32_t *keys, int64_t ns, int ntiles, int64_t *off,
                         hipStream_t s)
{
    hipLaunchKernelGGL(k_tile_offsets, dim3((unsigned)((ntiles + 1 + 255) / 256)), dim3(256), 0,
                       s, keys, ns, ntiles, off);
}

template <typename T, int K>
static void small_tile_k(int np, dim3 g, const int4 *items, const int64_t *nitems,
                         const int32_t *plist, const Rec2 *rec, const Map2D *maps, int nq,
                         int plane0, hipStream_t s)
{
    const dim3 b(small_threads(np));
    switch (np) {
    case 1: hipLaunchKernelGGL((k2d_small_tile<T, K, 1>), g, b, 0, s, items, nitems, plist, rec, maps, nq, plane0); break;
    case 2: hipLaunchKernelGGL((k2d_small_tile<T, K, 2>), g, b, 0, s, items, nitems, plist, rec, maps, nq, plane0); break;
    case 3: hipLaunchKernelGGL((k2d_small_tile<T, K, 3>), g, b, 0, s, items, nitems, plist, rec, maps, nq, plane0); break;
    default: hipLaunchKernelGGL((k2d_small_tile<T, K, 4>), g, b, 0, s, items, nitems, plist, rec, maps, nq, plane0); break;
    }
}

void launch_2d_small_tile(int dtype, int kernel, int np, const int4 *items,
                          const int64_t *nitems, int64_t max_items, const int32_t *plist,
                          const Rec2 *rec, const Map2D *maps, int nq, int plane0, hipStream_t s)
{
    if (max_items == 0)
        return;
    const dim3 g((unsigned)max_items);
    if (dtype == S2G_F32) {
        S2G_K2D_SWITCH(kernel, (small_tile_k<float, KK>(np, g, items, nitems, plist, rec, maps,
                                                         nq, plane0, s)));
    } else {
        S2G_K2D_SWITCH(kernel, (small_tile_k<double, KK>(np, g, items, nitems, plist, rec, maps,
                                                          nq, plane0, s)));
    }
}

void launch_2d_reduce_maps(const Map2D *maps, int nmaps, int ntiles, int nq, hipStream_t s)
{
    if (ntiles == 0)
        return;
    hipLaunchKernelGGL(k2d_reduce_maps, dim3((unsigned)ntiles), dim3(256), 0, s, maps, nmaps, nq);
}

void launch_2d_reduce(const double *image, int64_t npix, int nq, int reduce, double *out,
                      hipStream_t s)
{
    const unsigned nt = (unsigned)((npix + 63) / 64);
    hipLaunchKernelGGL(k2d_reduce, dim3(nt, nt), dim3(256), 0, s, image, npix, nq, reduce, out);
}

}  // namespace s2g
