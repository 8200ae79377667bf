// k_cic3d.hip -- the 3D deposit (cic_interpolation/cic_3D.jl) on gfx950.
//
// Same structure as the 2D pipeline: FP64 per-particle setup and class,
// stable class lists, explicit normalisation of clipped / moderate footprints,
// a deterministic counting sort of (brick, particle) pairs, a brick gather
// with FP64 register accumulation, and a scatter for small footprints.
// Bricks are 16^3 voxels; a block of 4 waves owns one brick.
#include "k_cic3d.h"

namespace s2g {

// ---------------------------------------------------------------------------
// setup: get_quantities_3D (cic_3D.jl:87-97), get_xyz (cic_shared.jl:85-100),
// pix_index_min_max x3 (cic_3D.jl:153-155) and the fallback test of
// calculate_weights (cic_3D.jl:55-75), FP64 without contraction.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k3d_setup(const T *__restrict__ pos,
                                                 const T *__restrict__ hsml,
                                                 const T *__restrict__ m,
                                                 const T *__restrict__ rho,
                                                 const T *__restrict__ binq,
                                                 const T *__restrict__ w, int64_t n,
                                                 int64_t npix, double len2pix, int kernel,
                                                 uint32_t flags, Rec3 *__restrict__ rec,
                                                 int32_t *__restrict__ cls)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n)
        return;
    Rec3 r;
    r.x = r.y = r.z = r.h = r.s = 0.0;
    r.i0 = r.j0 = r.k0 = 0;
    r.i1 = r.j1 = r.k1 = -1;
    r.cls = kClsSkip;
    r.fb = 0;
    // cic_3D.jl:140-144
    if (ld(binq, p) == 0.0 && !(flags & S2G_FLAG_CALC_MEAN)) {
        rec[p] = r;
        cls[p] = r.cls;
        return;
    }
    const double h = mul_rn(ld(hsml, p), len2pix);
    const double l3 = mul_rn(mul_rn(len2pix, len2pix), len2pix);
    const double rh = ld(rho, p) / l3;
    const double vol = ld(m, p) / rh;
    const double wl = ld(w, p);
    const double half = 0.5 * (double)npix;
    double c[3];
    for (int d = 0; d < 3; ++d)
        c[d] = add_rn(mul_rn(ld(pos, 3 * p + d), len2pix), half);

    const double lim = (double)npix + 1.0;
    int64_t lo[3], hi[3], a[3], b[3];
    bool bad = !(h > 0.0);
    for (int d = 0; d < 3; ++d) {
        double fa = floor(sub_rn(c[d], h)), fb = floor(add_rn(c[d], h));
        if (!(fa == fa) || !(fb == fb))
            bad = true;
        fa = fmin(fmax(fa, -1.0), lim);
        fb = fmin(fmax(fb, -1.0), lim);
        a[d] = (int64_t)fa;
        b[d] = (int64_t)fb;
        lo[d] = a[d] > 0 ? a[d] : 0;
        hi[d] = b[d] < npix - 1 ? b[d] : npix - 1;
        if (lo[d] > hi[d])
            bad = true;
    }
    if (bad) {
        rec[p] = r;
        cls[p] = r.cls;
        return;
    }
    // fallback iff no voxel centre of the footprint has u < 1 (FP64, nearest
    // centre per axis; three candidates cover ties and clamping)
    const double h_inv = 1.0 / h;
    double best[3];
    for (int d = 0; d < 3; ++d) {
        best[d] = 1e300;
        const double fc = floor(c[d]);
        for (int k = -1; k <= 1; ++k) {
            const double ci = fmin(fmax(fc + k, (double)lo[d]), (double)hi[d]);
            best[d] = fmin(best[d], fabs(coff1(c[d], ci)));
        }
    }
    const double r2 = add_rn(add_rn(mul_rn(best[0], best[0]), mul_rn(best[1], best[1])),
                             mul_rn(best[2], best[2]));
    const double umin = mul_rn(sqrt(r2), h_inv);
    r.fb = !(umin < 1.0);
    r.x = c[0];
    r.y = c[1];
    r.z = c[2];
    r.h = h;
    r.i0 = (int32_t)lo[0];
    r.i1 = (int32_t)hi[0];
    r.j0 = (int32_t)lo[1];
    r.j1 = (int32_t)hi[1];
    r.k0 = (int32_t)lo[2];
    r.k1 = (int32_t)hi[2];
    // volume_norm numerator (cic_3D.jl:168-169): vol * w * len2pix
    const double num = mul_rn(mul_rn(vol, wl), len2pix);
    bool small = true, clipped = false;
    for (int d = 0; d < 3; ++d) {
        small = small && (hi[d] - lo[d] + 1) <= kSmallSide3D;
        clipped = clipped || a[d] < 0 || b[d] > npix - 1;
    }
    if (small) {
        r.cls = kClsSmall;
        r.s = num;
    } else if (!(flags & S2G_FLAG_EXACT_NORM) && !clipped && !r.fb &&
               h >= closed_form_min_h3(kernel)) {
        r.cls = kClsLargeAn;
        r.s = num / (mul_rn(mul_rn(h, h), h) * shape_ball_integral(kernel));
    } else {
        r.cls = kClsLargeEx;
        r.s = num;
    }
    rec[p] = r;
    cls[p] = r.cls;
}

// voxel weight factor shared by the normalisation and the gather:
//   f = shape(sqrt(r2xy + zd2) * hinv) * (ax*ay) * az
template <int K>
__device__ __forceinline__ float vox_factor(float r2xy, float zd2, float hinv, float axy,
                                            float az)
{
    const float u = __builtin_amdgcn_sqrtf(r2xy + zd2) * hinv;
    return (shape_of<K, float>(u) * axy) * az;
}

// ---------------------------------------------------------------------------
// Explicit normalisation: one wave per particle.  The (j, k) plane of the
// footprint is flattened over the lanes (64 at a time, per-lane y/z terms
// computed once per chunk); x rows are taken 64 at a time with the row terms
// broadcast by v_readlane.  Fixed per-lane order + butterfly: reproducible.
// ---------------------------------------------------------------------------
template <int K>
__device__ __forceinline__ void k3d_norm_one(int64_t gw, const int32_t *__restrict__ list, int64_t ne, Rec3 *__restrict__ rec)
{
    const int ln = threadIdx.x % kWave;
    const int p = list[gw];
    const Rec3 r = rec[p];
    const float hinv = r.fb ? 0.0f : (float)(1.0 / r.h);
    const int jw = r.j1 - r.j0 + 1, kw = r.k1 - r.k0 + 1;
    const int njk = jw * kw;
    double acc = 0.0;
    for (int fc = 0; fc < njk; fc += kWave) {
        const int f = fc + ln;
        const int dj = f / kw, dk = f - dj * kw;
        const double fj = (double)(r.j0 + dj), fk = (double)(r.k0 + dk);
        const float yd = (float)coff1(r.y, fj), zd = (float)coff1(r.z, fk);
        const float ay = fmaxf((float)overlap1(r.y, r.h, fj), 0.0f);
        const float az = fmaxf((float)overlap1(r.z, r.h, fk), 0.0f);
        const bool on = f < njk;
        const float yzd2 = yd * yd + zd * zd;
        const float ayz = on ? ay * az : 0.0f;
        for (int ic = r.i0; ic <= r.i1; ic += kWave) {
            const double fi = (double)(ic + ln);
            const float xd = (float)coff1(r.x, fi);
            const float xd2 = xd * xd;
            const float ax = fmaxf((float)overlap1(r.x, r.h, fi), 0.0f);
            const int nrow = r.i1 - ic + 1 < kWave ? r.i1 - ic + 1 : kWave;
            for (int rr = 0; rr < nrow; ++rr) {
                const float xd2r =
                    __int_as_float(__builtin_amdgcn_readlane(__float_as_int(xd2), rr));
                const float axr =
                    __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ax), rr));
                const float u = __builtin_amdgcn_sqrtf(xd2r + yzd2) * hinv;
                acc += (double)((shape_of<K, float>(u) * axr) * ayz);
            }
        }
    }
    acc = wave_sum(acc);
    if (ln == 0)
        rec[p].s = acc != 0.0 ? r.s / acc : 0.0;
}
template <int K>
__global__ __launch_bounds__(256) void k3d_norm(const int32_t *__restrict__ list, int64_t ne, Rec3 *__restrict__ rec)
{
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x / kWave);
    for (int64_t gw = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; gw < ne;
         gw += nw)
        k3d_norm_one<K>(gw, list, ne, rec);
}


// ---------------------------------------------------------------------------
// Brick gather.  Block = 4 waves = one 16^3 brick (4096 voxels); voxel
// (ii, jj, kk) of the brick, "row" r = ii*16 + jj.  Wave w owns rows
// [64w, 64w+64) (ii in [4w, 4w+4)); lane l works on kk = l & 15 and rows
// 64w + 4m + (l >> 4), m = 0..15.  Per batch of particles the row terms
// (xd^2 + yd^2, ax*ay) of all 256 rows are tabulated in LDS; per particle
// each lane computes its z terms, and the z overlap is folded into the FP64
// plane multipliers.  Per voxel the work equals the 2D gather's per pixel.
// ---------------------------------------------------------------------------
constexpr int kGBatch3 = 16;

template <typename T, int K, int NP>
__global__ __launch_bounds__(256) void k3d_gather(const int4 *__restrict__ items,
                                                  const int32_t *__restrict__ pairs,
                                                  const int32_t *__restrict__ large,
                                                  const Rec3 *__restrict__ rec,
                                                  const T *__restrict__ binq, int nq,
                                                  int plane0, int nb,
                                                  double *__restrict__ partial)
{
    __shared__ float2 s_row[kGBatch3][kBT * kBT];
    __shared__ double s_x[kGBatch3], s_y[kGBatch3], s_z[kGBatch3], s_h[kGBatch3];
    __shared__ float s_hinv[kGBatch3];
    __shared__ int s_i0[kGBatch3], s_i1[kGBatch3], s_j0[kGBatch3], s_j1[kGBatch3];
    __shared__ double s_mult[NP][kGBatch3];

    const int4 it = items[blockIdx.x];
    const int t = it.x;
    const int bi0 = (t / (nb * nb)) * kBT, bj0 = ((t / nb) % nb) * kBT, bk0 = (t % nb) * kBT;
    const int wv = threadIdx.x / kWave, ln = threadIdx.x % kWave;
    const int rg = ln >> 4;
    const int kk = ln & 15;
    const double fk = (double)(bk0 + kk);

    double acc[16][NP];
#pragma unroll
    for (int m = 0; m < 16; ++m)
#pragma unroll
        for (int q = 0; q < NP; ++q)
            acc[m][q] = 0.0;

    for (int base = 0; base < it.z; base += kGBatch3) {
        const int cnt = it.z - base < kGBatch3 ? it.z - base : kGBatch3;
        if (threadIdx.x < cnt) {
            const int l = pairs[it.y + base + threadIdx.x];
            const int p = large[l];
            const Rec3 r = rec[p];
            s_x[threadIdx.x] = r.x;
            s_y[threadIdx.x] = r.y;
            s_z[threadIdx.x] = r.z;
            s_h[threadIdx.x] = r.h;
            s_hinv[threadIdx.x] = r.fb ? 0.0f : (float)(1.0 / r.h);
            s_i0[threadIdx.x] = r.i0;
            s_i1[threadIdx.x] = r.i1;
            s_j0[threadIdx.x] = r.j0;
            s_j1[threadIdx.x] = r.j1;
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                const int P = plane0 + q;
                const double mult =
                    P == nq ? 1.0 : ld(binq, (int64_t)p * nq + (P < nq ? P : 0));
                s_mult[q][threadIdx.x] = mult * r.s;
            }
        }
        __syncthreads();
        for (int e = threadIdx.x; e < cnt * kBT * kBT; e += 256) {
            const int b = e / (kBT * kBT), rr = e % (kBT * kBT);
            const double fi = (double)(bi0 + (rr >> 4)), fj = (double)(bj0 + (rr & 15));
            const float xd = (float)coff1(s_x[b], fi), yd = (float)coff1(s_y[b], fj);
            const float ax = fmaxf((float)overlap1(s_x[b], s_h[b], fi), 0.0f);
            const float ay = fmaxf((float)overlap1(s_y[b], s_h[b], fj), 0.0f);
            s_row[b][rr] = make_float2(xd * xd + yd * yd, ax * ay);
        }
        __syncthreads();
        for (int b = 0; b < cnt; ++b) {
            const int i0 = __builtin_amdgcn_readfirstlane(s_i0[b]) - bi0;
            const int i1 = __builtin_amdgcn_readfirstlane(s_i1[b]) - bi0;
            if (4 * wv + 3 < i0 || 4 * wv > i1)
                continue;
            const int j0 = __builtin_amdgcn_readfirstlane(s_j0[b]) - bj0;
            const int j1 = __builtin_amdgcn_readfirstlane(s_j1[b]) - bj0;
            const double z = s_z[b], h = s_h[b];
            const float hinv = s_hinv[b];
            const float zd = (float)coff1(z, fk);
            const float zd2 = zd * zd;
            const double azd = (double)fmaxf((float)overlap1(z, h, fk), 0.0f);
            double mult[NP];
#pragma unroll
            for (int q = 0; q < NP; ++q)
                mult[q] = s_mult[q][b] * azd;
#pragma unroll
            for (int m = 0; m < 16; ++m) {
                // row 64w + 4m + rg: ii = 4w + m/4, jj = 4(m%4) + rg
                const int ii = 4 * wv + (m >> 2);
                const int jg = 4 * (m & 3);
                if (ii < i0 || ii > i1 || jg + 3 < j0 || jg > j1)
                    continue;
                const float2 rv = s_row[b][64 * wv + 4 * m + rg];
                const float u = __builtin_amdgcn_sqrtf(rv.x + zd2) * hinv;
                const double f = (double)(shape_of<K, float>(u) * rv.y);
#pragma unroll
                for (int q = 0; q < NP; ++q)
                    acc[m][q] = fma(f, mult[q], acc[m][q]);
            }
        }
        __syncthreads();
    }
    double *dst = partial + (int64_t)blockIdx.x * NP * kBT * kBT * kBT;
#pragma unroll
    for (int q = 0; q < NP; ++q)
#pragma unroll
        for (int m = 0; m < 16; ++m)
            dst[(q * kBT * kBT + 64 * wv + 4 * m + rg) * kBT + kk] = acc[m][q];
}

// fixed-order sum of a brick's item partials -> image planes (plain stores)
template <int NP>
__global__ __launch_bounds__(256) void k3d_finalize(const double *__restrict__ partial,
                                                    const int64_t *__restrict__ tile_items,
                                                    int nb, int64_t npix, int plane0,
                                                    double *__restrict__ image)
{
    const int t = blockIdx.x;
    const int bi0 = (t / (nb * nb)) * kBT, bj0 = ((t / nb) % nb) * kBT, bk0 = (t % nb) * kBT;
    const int64_t b = tile_items[2 * t], k = tile_items[2 * t + 1];
    const int64_t nd = npix * npix * npix;
    for (int v = threadIdx.x; v < kBT * kBT * kBT; v += 256) {
        const int i = bi0 + (v >> 8), j = bj0 + ((v >> 4) & 15), kz = bk0 + (v & 15);
        if (i >= npix || j >= npix || kz >= npix)
            continue;
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            double s = 0.0;
            for (int64_t itm = 0; itm < k; ++itm)
                s += partial[((b + itm) * NP + q) * (kBT * kBT * kBT) + v];
            image[(int64_t)(plane0 + q) * nd + ((int64_t)i * npix + j) * npix + kz] = s;
        }
    }
}

// ---------------------------------------------------------------------------
// Small footprints (box <= 8^3): one wave per particle, FP64 weights,
// wave-reduced normalisation, FP64 atomics (cic_3D.jl:158-188).
// ---------------------------------------------------------------------------
template <typename T, int K>
__device__ __forceinline__ void k3d_small_atomic_one(int64_t gw, const int32_t *__restrict__ list, int64_t ns, const Rec3 *__restrict__ rec,
        const T *__restrict__ binq, int64_t npix, double *__restrict__ image)
{
    const int ln = threadIdx.x % kWave;
    const int p = list[gw];
    const Rec3 r = rec[p];
    const int iw = r.i1 - r.i0 + 1, jw = r.j1 - r.j0 + 1, kw = r.k1 - r.k0 + 1;
    const int jk = jw * kw, fp = iw * jk;
    const double hinv = r.fb ? 0.0 : 1.0 / r.h;
    const float invjk = 1.0f / (float)jk, invkw = 1.0f / (float)kw;
    double sum = 0.0;
    for (int f = ln; f < fp; f += kWave) {
        const int di = (int)(((float)f + 0.5f) * invjk);
        const int rem = f - di * jk;
        const int dj = (int)(((float)rem + 0.5f) * invkw);
        const int dk = rem - dj * kw;
        const double fi = (double)(r.i0 + di), fj = (double)(r.j0 + dj), fk = (double)(r.k0 + dk);
        const double xd = coff1(r.x, fi), yd = coff1(r.y, fj), zd = coff1(r.z, fk);
        const double V = overlap1(r.x, r.h, fi) * overlap1(r.y, r.h, fj) * overlap1(r.z, r.h, fk);
        const double u = sqrt(xd * xd + yd * yd + zd * zd) * hinv;
        sum += shape_of<K, double>(u) * V;
    }
    sum = wave_sum(sum);
    if (!(sum != 0.0))
        return;
    const double sc = r.s / sum;
    const double bq = ld(binq, p);
    const int64_t nd = npix * npix * npix;
    for (int f = ln; f < fp; f += kWave) {
        const int di = (int)(((float)f + 0.5f) * invjk);
        const int rem = f - di * jk;
        const int dj = (int)(((float)rem + 0.5f) * invkw);
        const int dk = rem - dj * kw;
        const double fi = (double)(r.i0 + di), fj = (double)(r.j0 + dj), fk = (double)(r.k0 + dk);
        const double xd = coff1(r.x, fi), yd = coff1(r.y, fj), zd = coff1(r.z, fk);
        const double V = overlap1(r.x, r.h, fi) * overlap1(r.y, r.h, fj) * overlap1(r.z, r.h, fk);
        const double u = sqrt(xd * xd + yd * yd + zd * zd) * hinv;
        const double pw = shape_of<K, double>(u) * V * sc;
        if (pw != 0.0) {
            const int64_t idx = ((int64_t)(r.i0 + di) * npix + (r.j0 + dj)) * npix + (r.k0 + dk);
            unsafeAtomicAdd(&image[nd + idx], pw);
            unsafeAtomicAdd(&image[idx], bq * pw);
        }
    }
}
template <typename T, int K>
__global__ __launch_bounds__(256) void k3d_small_atomic(const int32_t *__restrict__ list, int64_t ns, const Rec3 *__restrict__ rec,
        const T *__restrict__ binq, int64_t npix, double *__restrict__ image)
{
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x / kWave);
    for (int64_t gw = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; gw < ns;
         gw += nw)
        k3d_small_atomic_one<T, K>(gw, list, ns, rec, binq, npix, image);
}


// reduce_image_3D (reduce_image.jl:39-55): out = q / w where q > 0; without
// reduce_image the weights are 1 (cic_interpolation.jl:230-232)
__global__ void k3d_reduce(const double *__restrict__ image, int64_t nd, int reduce,
                           double *__restrict__ out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nd)
        return;
    double v = image[t];
    if (reduce && v > 0.0)
        v /= image[nd + t];
    out[t] = v;
}

// ---------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------
#define S2G_KERNEL_SWITCH(kernel, CALL)                                         \
    switch (kernel) {                                                           \
    case S2G_KERNEL_CUBIC: { constexpr int KK = S2G_KERNEL_CUBIC; CALL; } break; \
    case S2G_KERNEL_QUINTIC: { constexpr int KK = S2G_KERNEL_QUINTIC; CALL; } break; \
    case S2G_KERNEL_WENDLAND_C2: { constexpr int KK = S2G_KERNEL_WENDLAND_C2; CALL; } break; \
    case S2G_KERNEL_WENDLAND_C4: { constexpr int KK = S2G_KERNEL_WENDLAND_C4; CALL; } break; \
    case S2G_KERNEL_WENDLAND_C6: { constexpr int KK = S2G_KERNEL_WENDLAND_C6; CALL; } break; \
    default: { constexpr int KK = S2G_KERNEL_WENDLAND_C8; CALL; } break;       \
    }

void launch_3d_setup(const Map3DArgs &a, Rec3 *rec, int32_t *cls, hipStream_t s)
{
    if (a.n == 0)
        return;
    const dim3 g((unsigned)((a.n + 255) / 256)), blk(256);
    if (a.dtype == S2G_F32)
        hipLaunchKernelGGL(k3d_setup<float>, g, blk, 0, s, (const float *)a.pos,
                           (const float *)a.hsml, (const float *)a.m, (const float *)a.rho,
                           (const float *)a.binq, (const float *)a.w, a.n, a.npix, a.len2pix,
                           a.kernel, a.flags, rec, cls);
    else
        hipLaunchKernelGGL(k3d_setup<double>, g, blk, 0, s, (const double *)a.pos,
                           (const double *)a.hsml, (const double *)a.m, (const double *)a.rho,
                           (const double *)a.binq, (const double *)a.w, a.n, a.npix,
                           a.len2pix, a.kernel, a.flags, rec, cls);
}

void launch_3d_norm(int kernel, const int32_t *list, int64_t ne, Rec3 *rec, hipStream_t s)
{
    if (ne == 0)
        return;
    const dim3 g((unsigned)wave_blocks(ne, 256)), blk(256);
    S2G_KERNEL_SWITCH(kernel, hipLaunchKernelGGL(k3d_norm<KK>, g, blk, 0, s, list, ne, rec));
}

template <typename T, int K>
static void gather3_tk(const Gather3Args &g, hipStream_t s)
{
    const dim3 grid((unsigned)g.nitems), blk(256);
    if (g.np == 1)
        hipLaunchKernelGGL((k3d_gather<T, K, 1>), grid, blk, 0, s, g.items, g.pairs, g.large,
                           g.rec, (const T *)g.binq, g.nq, g.plane0, g.nb, g.partial);
    else
        hipLaunchKernelGGL((k3d_gather<T, K, 2>), grid, blk, 0, s, g.items, g.pairs, g.large,
                           g.rec, (const T *)g.binq, g.nq, g.plane0, g.nb, g.partial);
}

void launch_3d_gather(const Gather3Args &g, hipStream_t s)
{
    if (g.dtype == S2G_F32) {
        S2G_KERNEL_SWITCH(g.kernel, (gather3_tk<float, KK>(g, s)));
    } else {
        S2G_KERNEL_SWITCH(g.kernel, (gather3_tk<double, KK>(g, s)));
    }
}

void launch_3d_finalize(const Gather3Args &g, hipStream_t s)
{
    const dim3 grid((unsigned)g.ntiles), blk(256);
    if (g.np == 1)
        hipLaunchKernelGGL(k3d_finalize<1>, grid, blk, 0, s, g.partial, g.tile_items, g.nb,
                           g.npix, g.plane0, g.image);
    else
        hipLaunchKernelGGL(k3d_finalize<2>, grid, blk, 0, s, g.partial, g.tile_items, g.nb,
                           g.npix, g.plane0, g.image);
}

template <typename T>
static void small3_t(const Small3Args &a, hipStream_t s)
{
    const dim3 g((unsigned)wave_blocks(a.ns, 256)), blk(256);
    S2G_KERNEL_SWITCH(a.kernel, hipLaunchKernelGGL((k3d_small_atomic<T, KK>), g, blk, 0, s,
                                                   a.list, a.ns, a.rec, (const T *)a.binq,
                                                   a.npix, a.image));
}

void launch_3d_small(const Small3Args &a, hipStream_t s)
{
    if (a.ns == 0)
        return;
    if (a.dtype == S2G_F32)
        small3_t<float>(a, s);
    else
        small3_t<double>(a, s);
}

void launch_3d_reduce(const double *image, int64_t npix, int reduce, double *out, hipStream_t s)
{
    const int64_t nd = npix * npix * npix;
    hipLaunchKernelGGL(k3d_reduce, dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, s, image, nd,
                       reduce, out);
}

}  // namespace s2g
