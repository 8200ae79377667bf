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

// ---------------------------------------------------------------------------
// Explicit normalisation: one wave per particle.  The voxel weights are
// formed as the brick gather forms them (s = (x offset / h)^2 +
// ((y offset / h)^2 + (z offset / h)^2) in the same FP32 operations,
// shape_row with the x overlap's root times the y overlap in FP32, the z
// overlap as an FP64 factor), so the sum and the deposited weights agree.
// ---------------------------------------------------------------------------

constexpr int kNX3 = 128;  // x terms tabulated per pass (per wave, in LDS)

struct NRow3 {
    int off;  // first position of the row in the wave's (j, k) list
    int kl;   // first k of the row, relative to the box
};

__device__ __forceinline__ int norm_incl_scan(int v, int ln)
{
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int t = __shfl_up(v, o, kWave);
        v += ln >= o ? t : 0;
    }
    return v;
}

// max over the 64 lanes of a fully active wave of values >= 0 (as integers:
// the bit patterns of non-negative floats order as the floats do), returned
// wave-uniform: DPP within each 16-lane row, then the four rows
template <int kCtrl>
__device__ __forceinline__ int dpp_max(int v)
{
    const int o = __builtin_amdgcn_update_dpp(0, v, kCtrl, 0xF, 0xF, true);
    return v > o ? v : o;
}
__device__ __forceinline__ float wave_max_nonneg(float x)
{
    int v = __float_as_int(x);
    v = dpp_max<0xB1>(v);   // quad_perm [1, 0, 3, 2]
    v = dpp_max<0x4E>(v);   // quad_perm [2, 3, 0, 1]
    v = dpp_max<0x141>(v);  // row_half_mirror
    v = dpp_max<0x140>(v);  // row_mirror
    const int r0 = __builtin_amdgcn_readlane(v, 0), r1 = __builtin_amdgcn_readlane(v, 16);
    const int r2 = __builtin_amdgcn_readlane(v, 32), r3 = __builtin_amdgcn_readlane(v, 48);
    const int m01 = r0 > r1 ? r0 : r1, m23 = r2 > r3 ? r2 : r3;
    return __int_as_float(m01 > m23 ? m01 : m23);
}

// This lane's share of the sum over the box [ia, ib] x [ja, jb] x [ka, kb]
// (any integers; voxels outside the grid follow the same formulas).  Only
// voxels within h + 1 voxel of the centre are visited (the others have
// u >= 1 in FP32 too, and weight 0); a fallback footprint keeps its whole
// box.  Per 64 rows j the (j, k) pairs inside that ball's section are listed
// (per-row k ranges, wave prefix sum) and taken 64 at a time, a pair per
// lane; each 64 pairs run over the x range of their union, with the x terms
// read from the wave's LDS table.  Fixed per-lane order: reproducible.
template <int K>
__device__ __forceinline__ double norm_box3(const Rec3 &r, double hinv, int ia, int ib, int ja,
                                            int jb, int ka, int kb, int ln, float2 *xt,
                                            NRow3 *rows)
{
    const float xc = (float)r.x, yc = (float)r.y, zc = (float)r.z;
    const float hr = (float)r.h + 1.0f;
    const bool all = r.fb != 0;
    double acc = 0.0;
    for (int xa = ia; xa <= ib; xa += kNX3) {
        const int xb = ib < xa + kNX3 - 1 ? ib : xa + kNX3 - 1;
        __builtin_amdgcn_wave_barrier();
        for (int e = ln; e <= xb - xa; e += kWave) {
            const double fi = (double)(xa + e);
            const float xs = (float)(coff1(r.x, fi) * hinv);
            xt[e] = make_float2(xs * xs,
                                overlap_root<K>(fmaxf((float)overlap1(r.x, r.h, fi), 0.0f)));
        }
        for (int jc = ja; jc <= jb; jc += kWave) {
            const int j = jc + ln;
            int len = 0, kl = 0;
            if (j <= jb) {
                int lo = ka, hi = kb;
                if (!all) {
                    const float dy = fabsf((float)j + 0.5f - yc);
                    if (dy < hr) {
                        const float wz = sqrtf(hr * hr - dy * dy);
                        const int a = (int)floorf(zc - 0.5f - wz);
                        const int c = (int)ceilf(zc - 0.5f + wz);
                        lo = a > lo ? a : lo;
                        hi = c < hi ? c : hi;
                    } else {
                        hi = lo - 1;
                    }
                }
                if (hi >= lo) {
                    len = hi - lo + 1;
                    kl = lo - ka;
                }
            }
            const int incl = norm_incl_scan(len, ln);
            const int total = __builtin_amdgcn_readlane(incl, kWave - 1);
            if (total == 0)
                continue;
            __builtin_amdgcn_wave_barrier();
            rows[ln] = NRow3{incl - len, kl};
            __builtin_amdgcn_wave_barrier();
            int jr = -1, next = 0;
            NRow3 cur = {0, 0};
            for (int q0 = 0; q0 < total; q0 += kWave) {
                const int q = q0 + ln;
                float yz2 = 0.0f, wx = 0.0f, ay = 0.0f;
                double azd = 0.0;
                if (q < total) {
                    while (q >= next) {
                        ++jr;
                        cur = rows[jr];
                        next = jr + 1 < kWave ? rows[jr + 1].off : total;
                    }
                    const int jj = jc + jr, kk = ka + cur.kl + (q - cur.off);
                    const double fj = (double)jj, fk = (double)kk;
                    const float ys = (float)(coff1(r.y, fj) * hinv);
                    const float zs = (float)(coff1(r.z, fk) * hinv);
                    // squares rounded one by one, as the gather's tables hold
                    // them: the sum is then the gather's to the bit
                    yz2 = fmul_rn(ys, ys) + fmul_rn(zs, zs);
                    ay = fmaxf((float)overlap1(r.y, r.h, fj), 0.0f);
                    azd = (double)fmaxf((float)overlap1(r.z, r.h, fk), 0.0f);
                    const float dy = (float)jj + 0.5f - yc, dz = (float)kk + 0.5f - zc;
                    wx = sqrtf(fmaxf(hr * hr - dy * dy - dz * dz, 0.0f));
                }
                // the x range of the 64 pairs: the union of their sections
                int ra = xa, rb = xb;
                if (!all) {
                    const float wm = wave_max_nonneg(wx);
                    const int a = (int)floorf(xc - 0.5f - wm);
                    const int c = (int)ceilf(xc - 0.5f + wm);
                    ra = a > ra ? a : ra;
                    rb = c < rb ? c : rb;
                }
                double col = 0.0;
#pragma unroll 4
                for (int i = ra; i <= rb; ++i) {
                    const float2 t = xt[i - xa];
                    col += (double)(shape_row<K>(__builtin_amdgcn_sqrtf(t.x + yz2), t.y) * ay);
                }
                acc += col * azd;
            }
        }
    }
    return acc;
}

// Sum over the footprint in the grid.  When the closed form h^3 I_k holds for
// the unclipped footprint (h >= the kernel's threshold, no fallback, no
// S2G_FLAG_EXACT_NORM), the particle's centre is in the grid and the part cut
// off by the border is the smaller one, the sum is h^3 I_k minus the sum over
// that part (x slabs, then y slabs of the kept x range, then z slabs of the
// kept x and y ranges); otherwise the footprint in the grid is summed.  The
// boxes go through one loop (one copy of the box sum in the kernel's code).
template <int K>
__device__ __forceinline__ void k3d_norm_one(int64_t gw, const int32_t *__restrict__ list,
                                             int64_t npix, uint32_t flags,
                                             Rec3 *__restrict__ rec, float2 *xt, NRow3 *rows)
{
    const int ln = threadIdx.x % kWave;
    const int p = list[gw];
    const Rec3 r = rec[p];
    const double hinv = r.fb ? 0.0 : 1.0 / r.h;
    const int i0 = r.i0, i1 = r.i1, j0 = r.j0, j1 = r.j1, k0 = r.k0, k1 = r.k1;
    bool cut = false;
    double A0 = 0.0, A1 = 0.0, B0 = 0.0, B1 = 0.0, C0 = 0.0, C1 = 0.0;
    if (!r.fb && r.h >= closed_form_min_h3(K) && !(flags & S2G_FLAG_EXACT_NORM)) {
        const double np = (double)npix;
        A0 = floor(sub_rn(r.x, r.h));
        A1 = floor(add_rn(r.x, r.h));
        B0 = floor(sub_rn(r.y, r.h));
        B1 = floor(add_rn(r.y, r.h));
        C0 = floor(sub_rn(r.z, r.h));
        C1 = floor(add_rn(r.z, r.h));
        const double kept =
            (double)(i1 - i0 + 1) * (double)(j1 - j0 + 1) * (double)(k1 - k0 + 1);
        const double full = (A1 - A0 + 1.0) * (B1 - B0 + 1.0) * (C1 - C0 + 1.0);
        cut = r.x >= 0.0 && r.x < np && r.y >= 0.0 && r.y < np && r.z >= 0.0 && r.z < np &&
              full - kept < kept;
    }
    // a cut footprint's full box is at most twice the grid's: int holds it
    const int a0 = (int)A0, a1 = (int)A1, b0 = (int)B0, b1 = (int)B1;
    const int c0 = (int)C0, c1 = (int)C1;
    double out = 0.0;
    for (int sl = cut ? 0 : 6; sl < (cut ? 6 : 7); ++sl) {
        int ia = i0, ib = i1, ja = j0, jb = j1, ka = k0, kb = k1;
        if (sl < 2) {
            ja = b0;
            jb = b1;
            ka = c0;
            kb = c1;
            ia = sl == 0 ? a0 : i1 + 1;
            ib = sl == 0 ? i0 - 1 : a1;
        } else if (sl < 4) {
            ka = c0;
            kb = c1;
            ja = sl == 2 ? b0 : j1 + 1;
            jb = sl == 2 ? j0 - 1 : b1;
        } else if (sl < 6) {
            ka = sl == 4 ? c0 : k1 + 1;
            kb = sl == 4 ? k0 - 1 : c1;
        }
        if (ia > ib || ja > jb || ka > kb)
            continue;
        out += norm_box3<K>(r, hinv, ia, ib, ja, jb, ka, kb, ln, xt, rows);
    }
    out = wave_sum(out);
    const double acc = cut ? mul_rn(mul_rn(r.h, r.h), r.h) * shape_ball_integral(K) - out : out;
    if (ln == 0)
        rec[p].s = acc != 0.0 ? r.s / acc : 0.0;
}
template <int K>
__global__ __launch_bounds__(256) void k3d_norm(const int32_t *__restrict__ list, int64_t ne,
                                                int64_t npix, uint32_t flags,
                                                Rec3 *__restrict__ rec)
{
    __shared__ float2 s_xt[256 / kWave][kNX3];
    __shared__ NRow3 s_rows[256 / kWave][kWave];
    const int wv = threadIdx.x / kWave;
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x / kWave);
    for (int64_t gw = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; gw < ne;
         gw += nw)
        k3d_norm_one<K>(gw, list, npix, flags, rec, s_xt[wv], s_rows[wv]);
}


// ---------------------------------------------------------------------------
// Brick gather.  Block = 4 waves = one 16^3 brick (4096 voxels), cut into
// 4^3 sub-bricks.  Wave w owns the z slab kk in [4w, 4w + 4): its 16
// sub-bricks m = 4 bi + bj; lane l works on voxel (4 bi + (l >> 4),
// 4 bj + ((l >> 2) & 3), 4w + (l & 3)) of sub-brick m.  Per batch of
// particles the block tabulates per axis the 16 offset terms (offset / h)^2
// and overlaps, and one bit per sub-brick that can reach the kernel sphere
// (voxel-centre box within h + 1 voxel of the centre, inside the footprint
// box); a wave evaluates only its sub-bricks with that bit set.  The x
// overlap enters shape_row as ax^(1/n), the y overlap multiplies the FP32
// weight and the z overlap (fixed per lane) folds into the FP64 plane
// multipliers: one instruction per voxel more than the 2D gather's per
// pixel, and a per-particle set-up of one FP64 multiplier per plane.
// ---------------------------------------------------------------------------
constexpr int kGBatch3 = 32;
// sub-bricks per uniform test (1, 2 or 4: one bi per test; 4 was slower on
// the 2^24 cluster set at 256^3, 64.3 -> 69.6 ms, more sub-bricks outside
// the footprint evaluated)
constexpr int kG3 = 2;

template <typename T, int K, int NP>
__global__ __launch_bounds__(256, 4) void k3d_gather(const int4 *__restrict__ items,
                                                     const int32_t *__restrict__ pairs,
                                                     const Rec3 *__restrict__ rec,
                                                     const T *__restrict__ binq, int nq,
                                                     int plane0, int nb,
                                                     double *__restrict__ partial)
{
    __shared__ float2 s_tab[3][kGBatch3][kBT];
    __shared__ uint32_t s_hit[kGBatch3][2];
    __shared__ double s_c[kGBatch3][3], s_h[kGBatch3];
    __shared__ double s_hinvd[kGBatch3];
    __shared__ float s_dd[kGBatch3][3][4];  // per axis: squared sub-brick distance
    __shared__ float s_rr[kGBatch3];        // (h + 1)^2
    __shared__ double s_mult[NP][kGBatch3];

    const int4 it = items[blockIdx.x];
    const int t = it.x;
    int b0[3];
    b0[0] = (t / (nb * nb)) * kBT;
    b0[1] = ((t / nb) % nb) * kBT;
    b0[2] = (t % nb) * kBT;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), ln = threadIdx.x % kWave;
    const int di = ln >> 4, dj = (ln >> 2) & 3, dk = ln & 3;

    double acc[16][NP];
#pragma unroll
    for (int m = 0; m < 16; ++m)
#pragma unroll
        for (int q = 0; q < NP; ++q)
            acc[m][q] = 0.0;

    for (int base = 0; base < it.z; base += kGBatch3) {
        const int cnt = it.z - base < kGBatch3 ? it.z - base : kGBatch3;
        if (threadIdx.x < cnt) {
            const int p = pairs[it.y + base + threadIdx.x];
            const Rec3 r = rec[p];
            s_c[threadIdx.x][0] = r.x;
            s_c[threadIdx.x][1] = r.y;
            s_c[threadIdx.x][2] = r.z;
            s_h[threadIdx.x] = r.h;
            s_hinvd[threadIdx.x] = r.fb ? 0.0 : 1.0 / r.h;
            // sub-brick test terms: per axis and sub-brick the squared
            // distance from the centre to its voxel centres, infinite
            // outside the footprint box (0 inside it for a fallback
            // footprint, which keeps its whole box); one voxel of margin on
            // h, so FP32 rounding cannot drop a voxel with u < 1
            const int lo[3] = {r.i0 - b0[0], r.j0 - b0[1], r.k0 - b0[2]};
            const int hi[3] = {r.i1 - b0[0], r.j1 - b0[1], r.k1 - b0[2]};
            const double cc[3] = {r.x, r.y, r.z};
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                const float c = (float)(cc[ax] - (double)b0[ax]);
#pragma unroll
                for (int sb = 0; sb < 4; ++sb) {
                    const float dd =
                        fmaxf(fmaxf(4.0f * sb + 0.5f - c, c - (4.0f * sb + 3.5f)), 0.0f);
                    const bool inbox = 4 * sb <= hi[ax] && 4 * sb + 3 >= lo[ax];
                    s_dd[threadIdx.x][ax][sb] = !inbox ? __builtin_inff() : r.fb ? 0.0f : dd * dd;
                }
            }
            const float rr = (float)r.h + 1.0f;
            s_rr[threadIdx.x] = rr * rr;
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                const int P = plane0 + q;
                const double mult =
                    P == nq ? 1.0 : ld(binq, (int64_t)p * nq + (P < nq ? P : 0));
                s_mult[q][threadIdx.x] = mult * r.s;
            }
        }
        __syncthreads();
        // per axis the 16 terms ((offset / h)^2, overlap); offset / h in FP64,
        // rounded once; the x overlap as its n-th root.  One axis per wave.
        for (int e = threadIdx.x; e < 3 * kGBatch3 * kBT; e += 256) {
            const int ax = e / (kGBatch3 * kBT), b = (e / kBT) % kGBatch3, v = e % kBT;
            if (b >= cnt)
                continue;
            const double f = (double)(b0[ax] + v);
            const double c = s_c[b][ax];
            const float os = (float)(coff1(c, f) * s_hinvd[b]);
            const float ov = fmaxf((float)overlap1(c, s_h[b], f), 0.0f);
            s_tab[ax][b][v] = make_float2(os * os, ax == 0 ? overlap_root<K>(ov) : ov);
        }
        // sub-brick bits: lane = 16 bk + 4 bi + bj
        const int sbx = (ln >> 2) & 3, sby = ln & 3, sbz = ln >> 4;
        for (int b = wv; b < cnt; b += 4) {
            const float d2 = s_dd[b][0][sbx] + s_dd[b][1][sby] + s_dd[b][2][sbz];
            const uint64_t bits = __ballot(d2 < s_rr[b]);
            if (ln == 0) {
                s_hit[b][0] = (uint32_t)bits;
                s_hit[b][1] = (uint32_t)(bits >> 32);
            }
        }
        __syncthreads();
        for (int b = 0; b < cnt; ++b) {
            const uint32_t word = __builtin_amdgcn_readfirstlane(s_hit[b][wv >> 1]);
            const uint32_t wm = (word >> (16 * (wv & 1))) & 0xffffu;
            if (wm == 0)
                continue;
            const float2 zt = s_tab[2][b][4 * wv + dk];
            const double az = (double)zt.y;
            // the z overlap is common to the lane's voxels: fold it into the
            // FP64 plane multipliers; the y overlap of each bj multiplies the
            // FP32 weight (for the bj pairs a test of this particle will
            // evaluate)
            double mult[NP];
#pragma unroll
            for (int q = 0; q < NP; ++q)
                mult[q] = s_mult[q][b] * az;
            float yz[4], ay[4];
#pragma unroll
            for (int bj = 0; bj < 4; ++bj) {
                if (!(wm & (((1u << kG3) - 1u) * 0x1111u << (bj / kG3 * kG3))))
                    continue;
                const float2 yt = s_tab[1][b][4 * bj + dj];
                yz[bj] = yt.x + zt.x;
                ay[bj] = yt.y;
            }
#pragma unroll
            for (int m = 0; m < 16; m += kG3) {
                if (!((wm >> m) & ((1u << kG3) - 1u)))
                    continue;
                // the kG3 sub-bricks of one test share bi
                const float2 xt = s_tab[0][b][4 * (m >> 2) + di];
                float wk[kG3];
#pragma unroll
                for (int g = 0; g < kG3; ++g) {
                    const int bj = (m + g) & 3;
                    wk[g] = shape_row<K>(__builtin_amdgcn_sqrtf(xt.x + yz[bj]), xt.y) * ay[bj];
                }
#pragma unroll
                for (int g = 0; g < kG3; ++g) {
                    const double f = (double)wk[g];
#pragma unroll
                    for (int q = 0; q < NP; ++q)
                        acc[m + g][q] = fma(f, mult[q], acc[m + g][q]);
                }
            }
        }
        __syncthreads();
    }
    double *dst = partial + (int64_t)blockIdx.x * NP * kBT * kBT * kBT;
#pragma unroll
    for (int q = 0; q < NP; ++q)
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const int i = 4 * (m >> 2) + di, j = 4 * (m & 3) + dj, k = 4 * wv + dk;
            dst[((q * kBT + i) * kBT + j) * kBT + k] = acc[m][q];
        }
}

// fixed-order sum of a brick's item partials -> image planes (plain stores)
constexpr int kFinSplit3 = 16;

template <int NP>
// add: the image already holds the small scatter's sums (zeroed before it)
__global__ __launch_bounds__(256) void k3d_finalize(const double *__restrict__ partial,
                                                    const int64_t *__restrict__ tile_items,
                                                    int nb, int64_t npix, int plane0,
                                                    double *__restrict__ image, int add)
{
    // kFinSplit3 blocks per brick, so that a grid of few bricks still
    // spreads the sums over the CUs
    const int t = blockIdx.x / kFinSplit3;
    const int v0 = (blockIdx.x % kFinSplit3) * (kBT * kBT * kBT / kFinSplit3);
    const int bi0 = (t / (nb * nb)) * kBT, bj0 = ((t / nb) % nb) * kBT, bk0 = (t % nb) * kBT;
    const int64_t b = tile_items[2 * t], k = tile_items[2 * t + 1];
    const int64_t nd = npix * npix * npix;
    for (int v = v0 + threadIdx.x; v < v0 + kBT * kBT * kBT / kFinSplit3; v += 256) {
        const int i = bi0 + (v >> 8), j = bj0 + ((v >> 4) & 15), kz = bk0 + (v & 15);
        if (i >= npix || j >= npix || kz >= npix)
            continue;
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            double s = 0.0;
#pragma unroll 4
            for (int64_t itm = 0; itm < k; ++itm)
                s += partial[((b + itm) * NP + q) * (kBT * kBT * kBT) + v];
            double *dst = image + (int64_t)(plane0 + q) * nd + ((int64_t)i * npix + j) * npix + kz;
            *dst = add ? *dst + s : s;
        }
    }
}

// ---------------------------------------------------------------------------
// Small footprints (box <= 8^3)
// ---------------------------------------------------------------------------

// Flat small scatter: a wave takes 64 consecutive particles of the list and
// lays their footprint voxels end to end (exclusive prefix sum of the box
// sizes), 64 voxels per step, so a particle of 1 - 8 voxels no longer holds
// a whole wave.  In each 64-voxel window the particles that start there are
// flagged in LDS; a ballot of the flags and a lane-prefix count (mbcnt) give
// every lane its particle.  Pass 0 sums wk V per particle into LDS (one wave
// sum when the window holds one particle, LDS FP64 atomics otherwise); pass 1
// recomputes the same products and deposits wk V * vol w len2pix / sum with
// FP64 atomics (cic_3D.jl:158-188).
struct Small3Wave {
    double x[kWave], y[kWave], z[kWave], h[kWave], hinv[kWave], s[kWave], bq[kWave];
    double sum[kWave];
    int off[kWave], i0[kWave], j0[kWave], k0[kWave], jk[kWave], kw[kWave], flag[kWave];
    float invjk[kWave], invkw[kWave];
};

template <typename T, int K>
__global__ __launch_bounds__(256) void k3d_small_flat(const int32_t *__restrict__ list, int64_t ns,
                                                      const Rec3 *__restrict__ rec,
                                                      const T *__restrict__ binq, int64_t npix,
                                                      double *__restrict__ image)
{
    __shared__ Small3Wave s_w[256 / kWave];
    const int ln = threadIdx.x % kWave;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    Small3Wave &W = s_w[wv];
    const int64_t nd = npix * npix * npix;
    const int64_t ngroups = (ns + kWave - 1) / kWave;
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x / kWave);
    for (int64_t g = (int64_t)blockIdx.x * (blockDim.x / kWave) + wv; g < ngroups; g += nw) {
        const int64_t first = g * kWave;
        const int cnt = ns - first < kWave ? (int)(ns - first) : kWave;
        int fp = 0;
        __builtin_amdgcn_wave_barrier();
        if (ln < cnt) {
            const int p = list[first + ln];
            const Rec3 r = rec[p];
            const int iw = r.i1 - r.i0 + 1, jw = r.j1 - r.j0 + 1, kw = r.k1 - r.k0 + 1;
            fp = iw * jw * kw;
            W.x[ln] = r.x;
            W.y[ln] = r.y;
            W.z[ln] = r.z;
            W.h[ln] = r.h;
            W.hinv[ln] = r.fb ? 0.0 : 1.0 / r.h;
            W.s[ln] = r.s;
            W.bq[ln] = (double)ld(binq, p);
            W.sum[ln] = 0.0;
            W.i0[ln] = r.i0;
            W.j0[ln] = r.j0;
            W.k0[ln] = r.k0;
            W.jk[ln] = jw * kw;
            W.kw[ln] = kw;
            W.invjk[ln] = 1.0f / (float)(jw * kw);
            W.invkw[ln] = 1.0f / (float)kw;
        }
        const int incl = norm_incl_scan(fp, ln);
        W.off[ln] = incl - fp;
        const int total = __builtin_amdgcn_readlane(incl, kWave - 1);
        __builtin_amdgcn_wave_barrier();
        for (int pass = 0; pass < 2; ++pass) {
            int bprev = -1;  // the particle holding voxel e0 - 1
            for (int e0 = 0; e0 < total; e0 += kWave) {
                W.flag[ln] = 0;
                __builtin_amdgcn_wave_barrier();
                const int bj = bprev + 1 + ln;
                if (bj < cnt) {
                    const int pj = W.off[bj] - e0;  // >= 0: particle bj starts at or after e0
                    if (pj < kWave)
                        W.flag[pj] = 1;
                }
                __builtin_amdgcn_wave_barrier();
                const uint64_t mask = __ballot(W.flag[ln] != 0);
                const int below = (int)__builtin_amdgcn_mbcnt_hi(
                    (uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
                const int b = bprev + below + (int)((mask >> ln) & 1u);
                const int e = e0 + ln;
                double wvol = 0.0;
                if (e < total) {
                    const int f = e - W.off[b];
                    const int jk = W.jk[b], kw = W.kw[b];
                    const int di = (int)(((float)f + 0.5f) * W.invjk[b]);
                    const int rem = f - di * jk;
                    const int dj = (int)(((float)rem + 0.5f) * W.invkw[b]);
                    const int dk = rem - dj * kw;
                    const double x = W.x[b], y = W.y[b], z = W.z[b], h = W.h[b];
                    const double fi = (double)(W.i0[b] + di), fj = (double)(W.j0[b] + dj),
                                 fk = (double)(W.k0[b] + dk);
                    const double xd = coff1(x, fi), yd = coff1(y, fj), zd = coff1(z, fk);
                    const double V = overlap1(x, h, fi) * overlap1(y, h, fj) * overlap1(z, h, fk);
                    const double u = sqrt(xd * xd + yd * yd + zd * zd) * W.hinv[b];
                    wvol = shape_of<K, double>(u) * V;
                    if (pass == 1) {
                        const double pw = wvol * W.s[b];
                        if (pw != 0.0) {
                            const int64_t idx =
                                ((int64_t)(W.i0[b] + di) * npix + (W.j0[b] + dj)) * npix +
                                (W.k0[b] + dk);
                            unsafeAtomicAdd(&image[nd + idx], pw);
                            unsafeAtomicAdd(&image[idx], W.bq[b] * pw);
                        }
                    }
                }
                if (pass == 0) {
                    if ((mask & ~1ull) == 0) {
                        // one particle in the window
                        const double t = wave_sum(wvol);
                        if (ln == 0)
                            W.sum[b] += t;
                    } else if (wvol != 0.0) {
                        unsafeAtomicAdd(&W.sum[b], wvol);
                    }
                }
                bprev += __popcll(mask);
                __builtin_amdgcn_wave_barrier();
            }
            if (pass == 0) {
                // the scale of pass 1: vol w len2pix / sum (0 when nothing
                // in the box has weight)
                if (ln < cnt) {
                    const double sm = W.sum[ln];
                    W.s[ln] = sm != 0.0 ? W.s[ln] / sm : 0.0;
                }
                __builtin_amdgcn_wave_barrier();
            }
        }
    }
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

void launch_3d_norm(int kernel, const int32_t *list, int64_t ne, int64_t npix, uint32_t flags,
                    Rec3 *rec, hipStream_t s)
{
    if (ne == 0)
        return;
    const dim3 g((unsigned)wave_blocks(ne, 256)), blk(256);
    S2G_KERNEL_SWITCH(kernel, hipLaunchKernelGGL(k3d_norm<KK>, g, blk, 0, s, list, ne, npix,
                                                 flags, rec));
}

template <typename T, int K>
static void gather3_tk(const Gather3Args &g, hipStream_t s)
{
    const dim3 grid((unsigned)g.nitems), blk(256);
    if (g.np == 1)
        hipLaunchKernelGGL((k3d_gather<T, K, 1>), grid, blk, 0, s, g.items, g.pairs,
                           g.rec, (const T *)g.binq, g.nq, g.plane0, g.nb, g.partial);
    else
        hipLaunchKernelGGL((k3d_gather<T, K, 2>), grid, blk, 0, s, g.items, g.pairs,
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
    const dim3 grid((unsigned)(g.ntiles * kFinSplit3)), blk(256);
    if (g.np == 1)
        hipLaunchKernelGGL(k3d_finalize<1>, grid, blk, 0, s, g.partial, g.tile_items, g.nb,
                           g.npix, g.plane0, g.image, g.add);
    else
        hipLaunchKernelGGL(k3d_finalize<2>, grid, blk, 0, s, g.partial, g.tile_items, g.nb,
                           g.npix, g.plane0, g.image, g.add);
}

template <typename T>
static void small3_t(const Small3Args &a, hipStream_t s)
{
    const dim3 g((unsigned)wave_blocks((a.ns + kWave - 1) / kWave, 256)), blk(256);
    S2G_KERNEL_SWITCH(a.kernel, hipLaunchKernelGGL((k3d_small_flat<T, KK>), g, blk, 0, s,
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
