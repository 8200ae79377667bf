// k_healpix.hip -- the all-sky deposit of healpix_map
// (healpix_interpolation/main.jl:146-213) on gfx950.
//
//   k_hp_setup      per particle, FP64: skips (main.jl:160-179), projected
//                   hsml (:182), area and depth (:56-63, 193), direction, the
//                   RING-scheme disc query's ring range and polar caps, the
//                   pixel containing the centre (constributing_pixels.jl:7-22)
//   k_hp_units      (particle, group of rings) work units from a scan
//   k_hp_norm       per unit: partial sums of W*A and A (FP64, fixed order)
//   k_hp_finish     per particle: ordered sum of its units -> scale, fallback
//   k_hp_entry_*    (Nside <= 1024) work units keyed by first and last ring,
//   k_hp_ring_count radix-sorted; per ring the particles that touch it,
//   k_hp_ring_fill  counted, scanned and written in sorted order (no atomics)
//   k_hp_ring_deposit  one block per ring: W*A*scale of every particle on the
//                   ring into both map rows held in LDS, then plain stores
//   k_hp_deposit    (Nside > 1024) per unit, FP64 global atomics
//
// Pixel distances use the haversine form: the ring's colatitude difference
// and the pixel's azimuth difference are formed in FP64, the remaining
// trigonometry in FP32.  This gives the great-circle distance to FP32
// relative accuracy even at pixel-scale angles.
#include "k_healpix.h"

namespace s2g {

constexpr double kPi = 3.14159265358979323846;
constexpr int kHpRingsPerUnit = 64;  // one ring per lane in the interval phase

// ---------------------------------------------------------------------------
// RING-scheme geometry (Gorski et al. 2005), FP64
// ---------------------------------------------------------------------------
__device__ __forceinline__ void hp_ring_info(int64_t nside, int64_t ring, int64_t &start,
                                             int64_t &nr, double &shift)
{
    const int64_t npix = 12 * nside * nside, ncap = 2 * nside * (nside - 1);
    if (ring < nside) {
        shift = 0.5;
        nr = 4 * ring;
        start = 2 * ring * (ring - 1);
    } else if (ring < 3 * nside) {
        shift = ((ring - nside) & 1) == 0 ? 0.5 : 0.0;
        nr = 4 * nside;
        start = ncap + (ring - nside) * 4 * nside;
    } else {
        const int64_t r = 4 * nside - ring;
        shift = 0.5;
        nr = 4 * r;
        start = npix - 2 * r * (r + 1);
    }
}

__device__ __forceinline__ double hp_ring_z(int64_t nside, int64_t ring)
{
    const double fact2 = 4.0 / (double)(12 * nside * nside);
    if (ring < nside)
        return 1.0 - (double)(ring * ring) * fact2;
    if (ring <= 3 * nside)
        return (double)(2 * nside - ring) * (2.0 / (3.0 * (double)nside));
    const int64_t r = 4 * nside - ring;
    return -1.0 + (double)(r * r) * fact2;
}

// z, 1 - z and 1 + z of a ring, the small one of the last two formed
// without cancellation (polar rings)
__device__ __forceinline__ void hp_ring_zz(int64_t nside, int64_t ring, double &z, double &omz,
                                           double &opz)
{
    const double fact2 = 4.0 / (double)(12 * nside * nside);
    if (ring < nside) {
        omz = (double)(ring * ring) * fact2;
        z = 1.0 - omz;
        opz = 2.0 - omz;
    } else if (ring <= 3 * nside) {
        z = (double)(2 * nside - ring) * (2.0 / (3.0 * (double)nside));
        omz = 1.0 - z;
        opz = 1.0 + z;
    } else {
        const int64_t r = 4 * nside - ring;
        opz = (double)(r * r) * fact2;
        z = -1.0 + opz;
        omz = 2.0 - opz;
    }
}

__device__ __forceinline__ int64_t hp_ring_above(int64_t nside, double z)
{
    const double az = fabs(z);
    if (az <= 2.0 / 3.0)
        return (int64_t)((double)nside * (2.0 - 1.5 * z));
    const int64_t ir = (int64_t)((double)nside * sqrt(3.0 * (1.0 - az)));
    return (z > 0) ? ir : 4 * nside - ir - 1;
}

__device__ __forceinline__ int64_t hp_ang2pix(int64_t nside, double z, double phi)
{
    const int64_t npix = 12 * nside * nside, ncap = 2 * nside * (nside - 1);
    const double za = fabs(z);
    double tt = fmod(phi, 2.0 * kPi);
    if (tt < 0.0)
        tt += 2.0 * kPi;
    tt = tt / (0.5 * kPi);
    if (za <= 2.0 / 3.0) {
        const double t1 = (double)nside * (0.5 + tt);
        const double t2 = (double)nside * z * 0.75;
        const int64_t jp = (int64_t)(t1 - t2), jm = (int64_t)(t1 + t2);
        const int64_t ir = nside + 1 + jp - jm;
        const int64_t kshift = 1 - (ir & 1);
        int64_t ip = (jp + jm - nside + kshift + 1) / 2;
        ip = ((ip % (4 * nside)) + 4 * nside) % (4 * nside);
        return ncap + (ir - 1) * 4 * nside + ip;
    }
    const double tp = tt - (double)(int64_t)tt;
    const double tmp = (double)nside * sqrt(3.0 * (1.0 - za));
    const int64_t jp = (int64_t)(tp * tmp), jm = (int64_t)((1.0 - tp) * tmp);
    const int64_t ir = jp + jm + 1;
    int64_t ip = (int64_t)(tt * (double)ir);
    ip = ((ip % (4 * ir)) + 4 * ir) % (4 * ir);
    return (z > 0) ? 2 * ir * (ir - 1) + ip : npix - 2 * ir * (ir + 1) + ip;
}

// ring index (1-based) and ring-local index of a RING pixel
__device__ __forceinline__ void hp_pix2ring(int64_t nside, int64_t pix, int64_t &ring,
                                            int64_t &k)
{
    const int64_t npix = 12 * nside * nside, ncap = 2 * nside * (nside - 1);
    if (pix < ncap) {
        int64_t r = (int64_t)((1.0 + sqrt(1.0 + 2.0 * (double)pix)) * 0.5);
        while (2 * r * (r - 1) > pix) --r;
        while (2 * (r + 1) * r <= pix) ++r;
        ring = r;
        k = pix - 2 * r * (r - 1);
    } else if (pix < npix - ncap) {
        const int64_t ip = pix - ncap;
        ring = ip / (4 * nside) + nside;
        k = ip % (4 * nside);
    } else {
        const int64_t ip = npix - pix;  // in [1, ncap]
        int64_t r = (int64_t)((1.0 + sqrt(2.0 * (double)ip - 1.0)) * 0.5);
        while (2 * r * (r + 1) < ip) ++r;
        while (r > 1 && 2 * (r - 1) * r >= ip) --r;
        ring = 4 * nside - r;
        k = pix - (npix - 2 * r * (r + 1));
    }
}

// The pixel interval of ring iz inside the disc: [lo, hi] ring-local, where
// lo may be negative (wrap: pixels lo+nr .. nr-1 and 0 .. hi).  Returns
// false if the ring contributes no pixel.  Mirrors the per-ring step of the
// disc query (non-inclusive: pixel centres within the radius).
__device__ __forceinline__ bool hp_ring_interval(int64_t nside, int64_t iz, double z0,
                                                 double xa, double cosr, double phi0,
                                                 int64_t &lo, int64_t &hi)
{
    const double z = hp_ring_z(nside, iz);
    const double x = (cosr - z * z0) * xa;
    const double ysq = 1.0 - z * z - x * x;
    double dphi;
    if (ysq <= 0.0)
        dphi = (x < 0.0) ? kPi : 0.0;
    else
        dphi = atan2(sqrt(ysq), x);
    if (!(dphi > 0.0))
        return false;
    int64_t start, nr;
    double shift;
    hp_ring_info(nside, iz, start, nr, shift);
    const double inv2pi = 1.0 / (2.0 * kPi);
    int64_t a = (int64_t)floor((double)nr * inv2pi * (phi0 - dphi) - shift) + 1;
    int64_t b = (int64_t)floor((double)nr * inv2pi * (phi0 + dphi) - shift);
    if (a > b)
        return false;
    if (b >= nr) {
        a -= nr;
        b -= nr;
    }
    lo = a;
    hi = b;
    return true;
}

// A ring lying wholly inside the disc (polar cap): all nr pixels, taken from
// the first one at or after phi0 - pi (lo may be negative, as above), so that
// phi - phi0 stays in [-pi, pi) along the chord and the pixel loops need no
// wrap of it.  The same set as [0, nr - 1], in another order.
__device__ __forceinline__ void hp_full_chord(int64_t nr, double shift, double phi0,
                                              int64_t &lo, int64_t &hi)
{
    const double inv2pi = 1.0 / (2.0 * kPi);
    int64_t a = (int64_t)floor((double)nr * inv2pi * (phi0 - kPi) - shift) + 1;
    int64_t b = a + nr - 1;
    if (b >= nr) {
        a -= nr;
        b -= nr;
    }
    lo = a;
    hi = b;
}

// ---------------------------------------------------------------------------
// setup
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_hp_setup(const HpArgs a, HpRec *__restrict__ rec,
                                                  int64_t *__restrict__ nunits)
{
    const T *__restrict__ pos = (const T *)a.pos;
    const T *__restrict__ hsml = (const T *)a.hsml;
    const T *__restrict__ m = (const T *)a.m;
    const T *__restrict__ rho = (const T *)a.rho;
    const int64_t n = a.n, nside = a.nside;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n)
        return;
    HpRec r;
    r.z0 = r.phi0 = r.theta0 = r.sth0 = r.xa = r.theta = r.s = 0.0;
    r.ch0 = r.sh0 = r.cosr = 0.0;
    r.cring = r.ck = 0;
    r.ra = 1;
    r.rb = 0;
    r.irmin = 1;
    r.irmax = 0;
    r.cpix = -1;
    r.fb = 0;
    r.ok = 0;
    r.unit0 = 0;
    // main.jl:160-170: without calc_mean a zero quantity skips the particle;
    // here a particle is left out only if it is skipped in every map
    bool qskip = !(a.flags & S2G_FLAG_CALC_MEAN);
    for (int k = 0; k < a.nmaps && qskip; ++k)
        qskip = ld((const T *)a.binq[k], p) == 0.0;
    const double h = ld(hsml, p);
    const double px = ld(pos, 3 * p), py = ld(pos, 3 * p + 1), pz = ld(pos, 3 * p + 2);
    // shared.jl:1-10: sqrt((0 + x^2) + y^2) + z^2)
    const double dx = sqrt(add_rn(add_rn(mul_rn(px, px), mul_rn(py, py)), mul_rn(pz, pz)));
    bool skip = qskip || dx < h || !(dx == dx) || !(h == h);
    if (!skip) {
        const double npix = (double)(12 * nside * nside);
        const double ang_pix = sqrt(4.0 * kPi / npix);  // main.jl:144
        const double theta = asin(h / dx);               // main.jl:182
        const double dz0 = 2.0 * h;                       // main.jl:59-60
        const double area = (ld(m, p) / ld(rho, p)) / dz0;
        const double ap = ang_pix * dx;
        const double dz = dz0 / (ap * ap);  // main.jl:193
        r.s = mul_rn(area, dz);
        r.theta = theta;
        // vec2ang
        const double th0 = atan2(sqrt(add_rn(mul_rn(px, px), mul_rn(py, py))), pz);
        double ph0 = atan2(py, px);
        if (ph0 < 0.0)
            ph0 += 2.0 * kPi;
        r.theta0 = th0;
        r.ch0 = cos(0.5 * th0);
        r.sh0 = sin(0.5 * th0);
        r.cosr = cos(theta);
        r.phi0 = ph0;
        r.z0 = cos(th0);
        r.sth0 = sin(th0);
        r.xa = 1.0 / sqrt((1.0 - r.z0) * (1.0 + r.z0));
        // ring range of the disc query
        const double rlat1 = th0 - theta;
        const int64_t irmin = hp_ring_above(nside, cos(rlat1)) + 1;
        const double rlat2 = th0 + theta;
        const int64_t irmax = hp_ring_above(nside, cos(rlat2));
        r.irmin = (int32_t)irmin;
        r.irmax = (int32_t)irmax;
        r.ra = (rlat1 <= 0.0 && irmin > 1) ? 1 : (int32_t)irmin;
        r.rb = (rlat2 >= kPi && irmax + 1 < 4 * nside) ? (int32_t)(4 * nside - 1) : (int32_t)irmax;
        r.cpix = hp_ang2pix(nside, r.z0, ph0);
        int64_t cring, ck;
        hp_pix2ring(nside, r.cpix, cring, ck);
        r.cring = (int32_t)cring;
        r.ck = (int32_t)ck;
        r.ok = 1;
    }
    rec[p] = r;
    const int64_t nrings = r.ok ? (int64_t)(r.rb - r.ra + 1) : 0;
    // one extra unit when only the centre pixel contributes
    int64_t u = nrings > 0 ? (nrings + kHpRingsPerUnit - 1) / kHpRingsPerUnit : (r.ok ? 1 : 0);
    nunits[p] = u;
    if (p == n - 1)
        nunits[n] = 0;  // becomes the unit total after the scan
}

__global__ void k_hp_units(const int64_t *__restrict__ ubase, int64_t n, HpRec *__restrict__ rec,
                           int2 *__restrict__ units)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n)
        return;
    const int64_t b = ubase[p], e = ubase[p + 1];
    rec[p].unit0 = b;
    for (int64_t k = b; k < e; ++k)
        units[k] = make_int2((int)p, (int)(k - b));
}

// ---------------------------------------------------------------------------
// per-pixel factors
// ---------------------------------------------------------------------------
struct RingTerms {
    float s1, s2;      // sin^2(dtheta/2), sin(theta_ring) sin(theta0)
    double dphi_pix;   // 2 pi / nr
    double phase;      // shift * dphi_pix - phi0
    int64_t start;
    int64_t nr;
};

// Per-ring half-angle terms (ring iz at rtab[iz]): sin(theta_ring / 2) =
// sqrt((1 - z) / 2), cos(theta_ring / 2) = sqrt((1 + z) / 2) and sin(theta_ring)
// = sqrt((1 - z)(1 + z)), z = cos(theta_ring); formed once per call instead of
// once per (particle, ring), with the same FP64 operations.
__global__ __launch_bounds__(256) void k_hp_ring_table(int64_t nside, double4 *__restrict__ rtab)
{
    const int64_t iz = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (iz < 1 || iz > 4 * nside - 1)
        return;
    double z, omz, opz;
    hp_ring_zz(nside, iz, z, omz, opz);
    rtab[iz] = make_double4(sqrt(0.5 * omz), sqrt(0.5 * opz), sqrt(omz * opz), z);
}

void launch_hp_ring_table(int64_t nside, double4 *rtab, hipStream_t s)
{
    hipLaunchKernelGGL(k_hp_ring_table, dim3((unsigned)((4 * nside + 255) / 256)), dim3(256), 0,
                       s, nside, rtab);
}

__device__ __forceinline__ RingTerms ring_terms(int64_t nside, int64_t iz, const HpRec &r,
                                                const double4 *__restrict__ rtab)
{
    RingTerms t;
    double shift;
    hp_ring_info(nside, iz, t.start, t.nr, shift);
    // sin((theta_ring - theta0) / 2) from the half-angle terms of both colatitudes
    const double4 q = rtab[iz];
    const double sh = q.x * r.ch0 - q.y * r.sh0;
    t.s1 = (float)(sh * sh);
    t.s2 = (float)(q.z * r.sth0);
    t.dphi_pix = 2.0 * kPi / (double)t.nr;
    t.phase = shift * t.dphi_pix - r.phi0;
    return t;
}

// phi_a - phi0 of a chord's first pixel a.  Its true value lies in
// (-pi, 2 pi / nr] (the chord starts within dphi of phi0 - dphi, dphi <= pi,
// hp_ring_interval / hp_full_chord), and the index shift of a wrapped chord
// adds -2 pi or +2 pi; the representative is taken in [-pi - 1e-9, pi / 2),
// so a start a rounding below -pi keeps its place and the chord's offsets
// run up to < pi (nr >= 4).
__device__ __forceinline__ double hp_chord_start(int64_t a, const RingTerms &t)
{
    double d0 = (double)a * t.dphi_pix + t.phase;
    d0 = d0 >= 0.5 * kPi ? d0 - 2.0 * kPi : d0;
    d0 = d0 < -kPi - 1e-9 ? d0 + 2.0 * kPi : d0;
    return d0;
}

// sin(y) for |y| <= pi/2: y (1 + y^2 P(y^2)) with P of degree 3 fitted for
// the least maximum relative error on the interval (7.6e-9; the FP32
// evaluation adds ~1e-7, as any FP32 formula does).  One branch-free formula
// for every pixel.
__device__ __forceinline__ float hp_sin(float y)
{
    const float y2 = y * y;
    float p = fmaf(y2, 2.60893264e-06f, -1.98111127e-04f);
    p = fmaf(y2, p, 8.33308817e-03f);
    p = fmaf(y2, p, -1.66666605e-01f);
    return fmaf(y * y2, p, y);
}

// asin(x) for 0 <= x <= 1: x (1 + x^2 P(x^2)) with P of degree 4 fitted
// for the least maximum relative error on [0, 1/2] (4.9e-9; the FP32
// evaluation adds ~6e-8), which covers every pixel closer than pi/3 to the
// particle; the library function above.  The choice is per pixel, so a
// pixel's weight does not depend on which pixels share its wave.
__device__ __forceinline__ float hp_asin(float x)
{
    if (x <= 0.5f) {
        const float t = x * x;
        float p = fmaf(t, 4.216631e-02f, 2.417951e-02f);
        p = fmaf(t, p, 4.547038e-02f);
        p = fmaf(t, p, 7.495297e-02f);
        p = fmaf(t, p, 1.6666752e-01f);
        return fmaf(x * t, p, x);
    }
    return asinf(x);
}

// angular distance from the haversine: hav(d) = sin^2(dtheta / 2) +
// sin(theta) sin(theta0) sin^2(dphi / 2), d = 2 asin(sqrt(hav)).  The one
// formula for every pixel of the norm and deposit passes, so both see the
// same weights.
__device__ __forceinline__ float hp_angle(float dphi, float s1, float s2)
{
    const float sx = hp_sin(0.5f * dphi);
    const float hav = fminf(fmaf(s2, sx * sx, s1), 1.0f);
    return 2.0f * hp_asin(__builtin_amdgcn_sqrtf(hav));
}

// angular distance (radians) from the particle to pixel k of the ring
__device__ __forceinline__ float hp_dist(const RingTerms &t, int64_t k)
{
    double dphi = (double)k * t.dphi_pix + t.phase;  // phi_k - phi0
    dphi = dphi > kPi ? dphi - 2.0 * kPi : dphi;
    dphi = dphi < -kPi ? dphi + 2.0 * kPi : dphi;
    return hp_angle((float)dphi, t.s1, t.s2);
}

// W = shape(d / theta) and the area fraction (pixel_weights.jl:6-8, without
// the constant 1/(ang_pix*dx)^2 that cancels):
// A = max(0, min(ang, |theta - (d - ang / 2)|)) / ang, with thh = theta +
// ang / 2 formed once per particle (the clamp is one v_med3_f32)
template <int K>
__device__ __forceinline__ void hp_weights(float d, float thh, float inv_theta, float ang,
                                           float inv_ang, float &W, float &A)
{
    W = shape_row<K>(d * inv_theta, 1.0f);
    A = __builtin_amdgcn_fmed3f(fabsf(thh - d), 0.0f, ang) * inv_ang;
}

// One ring of a work unit, as the pixel loop reads it (per-wave LDS table).
struct HpRing {
    int off;     // first position of the ring's chord in the unit's pixel list
    int lo;      // ring-local index of the chord's first pixel (< 0: wraps)
    int start;   // RING index of the ring's first pixel
    int nr;      // pixels in the ring
    float s1, s2;
    float base;  // phi of the chord's first pixel - phi0, in [-pi, pi)
    float step;  // 2 pi / nr
};

__device__ __forceinline__ int wave_incl_scan(int v, int ln)
{
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int t = __shfl_up(v, o, kWave);
        v += ln >= o ? t : 0;
    }
    return v;
}

// Iterate the pixels of unit `ui` of particle r: rings [ra + 64*ui, ...].
// Phase 1: lane l computes the pixel interval and the terms of ring
// ra + 64*ui + l and writes them to the wave's LDS table `tab`, with its
// chord's offset in the unit's pixel list (wave prefix sum).  Phase 2: lane
// l takes list positions l, l + 64, ..., walking forward through the table,
// so every lane holds a pixel until the list ends.
// F(pix, W, A) is called per pixel on every lane that holds a pixel.
template <int K, typename F>
__device__ __forceinline__ void hp_unit_pixels(int64_t nside, const HpRec &r, int ui, int ln,
                                               float ang, HpRing *tab,
                                               const double4 *__restrict__ rtab, F &&f)
{
    const float thh = (float)r.theta + 0.5f * ang, inv_theta = (float)(1.0 / r.theta);
    const float inv_ang = 1.0f / ang;
    const double cosr = r.cosr;
    const int64_t r0 = (int64_t)r.ra + (int64_t)ui * kHpRingsPerUnit;
    const int64_t r1 = r0 + kHpRingsPerUnit - 1 < r.rb ? r0 + kHpRingsPerUnit - 1 : r.rb;
    const int nrings = r1 >= r0 ? (int)(r1 - r0 + 1) : 0;
    const int64_t cring = r.cring, ck = r.ck;

    // phase 1: one ring per lane
    HpRing e;
    e.lo = 0;
    e.start = 0;
    e.nr = 1;
    e.s1 = e.s2 = e.base = e.step = 0.0f;
    int len = 0;
    bool cov = false;
    if (ln < nrings) {
        const int64_t iz = r0 + ln;
        int64_t a = 0, b = -1;
        bool any;
        if (iz < r.irmin || iz > r.irmax) {  // a ring fully inside a polar cap
            int64_t st, n_r;
            double sh;
            hp_ring_info(nside, iz, st, n_r, sh);
            hp_full_chord(n_r, sh, r.phi0, a, b);
            any = true;
        } else {
            any = hp_ring_interval(nside, iz, r.z0, r.xa, cosr, r.phi0, a, b);
        }
        if (any) {
            const RingTerms t = ring_terms(nside, iz, r, rtab);
            e.lo = (int)a;
            len = (int)(b - a + 1);
            e.start = (int)t.start;
            e.nr = (int)t.nr;
            e.s1 = t.s1;
            e.s2 = t.s2;
            e.base = (float)hp_chord_start(a, t);
            e.step = (float)t.dphi_pix;
            if (iz == cring)
                cov = (a >= 0 && ck >= a && ck <= b) || (a < 0 && (ck <= b || ck >= a + t.nr));
        }
    }
    const bool covered_cp = __ballot(cov) != 0ull;
    const int incl = wave_incl_scan(len, ln);
    const int total = __shfl(incl, kWave - 1, kWave);
    e.off = incl - len;
    __builtin_amdgcn_wave_barrier();
    tab[ln] = e;
    __builtin_amdgcn_wave_barrier();

    // phase 2: the unit's pixel list, 64 positions at a time
    int j = -1, next = 0;
    HpRing cur = e;
    for (int q = ln; q < total; q += kWave) {
        while (q >= next) {
            ++j;
            cur = tab[j];
            next = j + 1 < kWave ? tab[j + 1].off : total;
        }
        const int c = q - cur.off;
        int k = cur.lo + c;
        k = k < 0 ? k + cur.nr : k;
        // in [-pi, pi] up to FP32 rounding: every chord starts at or after
        // phi0 - pi and spans less than 2 pi (hp_ring_interval, hp_full_chord)
        const float dphi = fmaf((float)c, cur.step, cur.base);
        const float d = hp_angle(dphi, cur.s1, cur.s2);
        float W, A;
        hp_weights<K>(d, thh, inv_theta, ang, inv_ang, W, A);
        f((int64_t)cur.start + k, W, A);
    }
    __builtin_amdgcn_wave_barrier();
    // constributing_pixels.jl:16-19: the pixel containing the particle is
    // added if the disc query did not return it (handled by the unit that
    // owns its ring, or by unit 0 if its ring is outside [ra, rb])
    const bool own = (cring >= r0 && cring <= r1) ||
                     (ui == 0 && (cring < r.ra || cring > r.rb || r.rb < r.ra));
    if (own && !covered_cp && ln == 0) {
        const RingTerms t = ring_terms(nside, cring, r, rtab);
        const float d = hp_dist(t, ck);
        float W, A;
        hp_weights<K>(d, thh, inv_theta, ang, inv_ang, W, A);
        f(r.cpix, W, A);
    }
}

// Per unit: sum of W*A in FP64 and of A in FP32 (A is needed only in the
// fallback, i.e. when no pixel centre lies inside the kernel, which happens
// only for discs of a few pixels).
template <int K>
__device__ __forceinline__ void k_hp_norm_one(int64_t gw, const int2 *__restrict__ units,
                                              const HpRec *__restrict__ rec, int64_t nside,
                                              HpRing *tab, const double4 *__restrict__ rtab,
                                              double4 *__restrict__ part)
{
    const int ln = threadIdx.x % kWave;
    const int2 u = units[gw];
    const HpRec r = rec[u.x];
    const float ang = (float)sqrt(4.0 * kPi / (double)(12 * nside * nside));
    double swa = 0.0;
    float sa = 0.0f;
    hp_unit_pixels<K>(nside, r, u.y, ln, ang, tab, rtab, [&](int64_t, float W, float A) {
        swa += (double)(W * A);
        sa += A;
    });
    swa = wave_sum(swa);
    const double sad = wave_sum((double)sa);
    if (ln == 0)
        part[gw] = make_double4(swa, sad, 0.0, 0.0);
}
// Held to 128 VGPRs (four waves per SIMD): the spills sit in the per-unit
// ring set-up, and the pixel loop gains more from the occupancy (norm pass
// 87 -> 61 ms per map on the all-sky workload).
template <int K>
__global__ __launch_bounds__(256, 4) void k_hp_norm(const int2 *__restrict__ units, int64_t nu,
                                                 const HpRec *__restrict__ rec, int64_t nside,
                                                 const double4 *__restrict__ rtab,
                                                 double4 *__restrict__ part)
{
    __shared__ HpRing s_tab[256 / kWave][kWave];
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x / kWave);
    for (int64_t gw = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; gw < nu;
         gw += nw)
        k_hp_norm_one<K>(gw, units, rec, nside, s_tab[threadIdx.x / kWave], rtab, part);
}


// ordered sum of a particle's unit partials -> final scale and fallback
__global__ void k_hp_finish(const int64_t *__restrict__ ubase, int64_t n,
                            const double4 *__restrict__ part, HpRec *__restrict__ rec)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n)
        return;
    HpRec r = rec[p];
    if (!r.ok)
        return;
    double swa = 0.0, sa = 0.0;
    for (int64_t k = ubase[p]; k < ubase[p + 1]; ++k) {
        swa += part[k].x;
        sa += part[k].y;
    }
    // pixel_weights.jl:120-137: fall back to pure area weights when no pixel
    // centre lies inside the kernel
    if (swa != 0.0) {
        r.fb = 0;
        r.s = r.s / swa;
    } else {
        r.fb = 1;
        r.s = sa != 0.0 ? r.s / sa : 0.0;
    }
    rec[p] = r;
}

// per-map pixel scale of particle p: s times the map's weight, 0 where the
// map skips the particle (zero quantity without calc_mean, main.jl:160-170)
template <typename T>
__device__ __forceinline__ void hp_map_scale(const HpGroup &g, int k, int64_t p, double s,
                                             uint32_t flags, double &sw, double &q)
{
    q = ld((const T *)g.binq[k], p);
    const bool skip = !(flags & S2G_FLAG_CALC_MEAN) && q == 0.0;
    sw = skip ? 0.0 : s * ld((const T *)g.w[k], p);
}

template <typename T, int K, int NM>
__device__ __forceinline__ void k_hp_deposit_one(int64_t gw, const int2 *__restrict__ units,
                                                 const HpRec *__restrict__ rec, const HpGroup &g,
                                                 uint32_t flags, int64_t nside, HpRing *tab)
{
    const int ln = threadIdx.x % kWave;
    const int2 u = units[gw];
    const HpRec r = rec[u.x];
    if (r.s == 0.0)
        return;
    double sw[NM], bq[NM];
#pragma unroll
    for (int k = 0; k < NM; ++k)
        hp_map_scale<T>(g, k, u.x, r.s, flags, sw[k], bq[k]);
    const float ang = (float)sqrt(4.0 * kPi / (double)(12 * nside * nside));
    const bool fb = r.fb != 0;
    hp_unit_pixels<K>(nside, r, u.y, ln, ang, tab, g.rtab, [&](int64_t pix, float W, float A) {
        const double wa = (double)((fb ? 1.0f : W) * A);
#pragma unroll
        for (int k = 0; k < NM; ++k) {
            const double pw = wa * sw[k];
            if (pw != 0.0) {
                unsafeAtomicAdd(&g.wmap[k][pix], pw);
                unsafeAtomicAdd(&g.map[k][pix], bq[k] * pw);
            }
        }
    });
}
template <typename T, int K, int NM>
__global__ __launch_bounds__(256, 4) void k_hp_deposit(const int2 *__restrict__ units, int64_t nu,
                                                    const HpRec *__restrict__ rec,
                                                    const HpGroup g, uint32_t flags,
                                                    int64_t nside)
{
    __shared__ HpRing s_tab[256 / kWave][kWave];
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x / kWave);
    for (int64_t gw = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; gw < nu;
         gw += nw)
        k_hp_deposit_one<T, K, NM>(gw, units, rec, g, flags, nside, s_tab[threadIdx.x / kWave]);
}


// ---------------------------------------------------------------------------
// Ring-major deposit (Nside <= kHpRingLds / 4): one block per ring holds the
// ring's two map rows in LDS, takes the particles whose disc (or centre
// pixel) touches the ring from a per-ring list, adds W*A*scale with LDS
// atomics and writes the ring with plain stores.  No global atomics; every
// pixel of the map is written, so no memset either.  The pixel weights are
// formed exactly as in the norm pass (same chord, ring terms and FP32
// arithmetic), so the deposited weights sum to the normalisation.
// ---------------------------------------------------------------------------
constexpr int kHpRingLds = 4096;  // pixels of the longest ring held in LDS
// One map: 8 waves per ring block, 32 chords per wave and round (two blocks
// per CU); two maps (four rows, 128 KB): 16 waves, 28 chords per wave, one
// block per CU.  Either way four waves per SIMD.
template <int NM>
constexpr int kHpRingThreads = NM == 1 ? 512 : 1024;
template <int NM>
constexpr int kHpRingPairs = NM == 1 ? 32 : 28;

// The rings a particle deposits to: its disc's rings [ra, rb] and the ring of
// its centre pixel (constributing_pixels.jl:16-19), which lies outside that
// range only when the disc query returns no pixel of it.  False for
// particles that deposit nothing.
// Every particle kept by the set-up is listed, whatever its scale: the lists
// are built beside the normalisation pass, which writes the scale (and sets
// it to 0 only when no pixel of the disc has an area weight; such a particle
// then adds zeros), so they must not read it.
__device__ __forceinline__ bool hp_dep_rings(int64_t nside, const HpRec &r, int64_t &cring,
                                             int64_t &ck, bool &cring_extra)
{
    if (!r.ok)
        return false;
    cring = r.cring;
    ck = r.ck;
    cring_extra = r.rb < r.ra || cring < r.ra || cring > r.rb;
    return true;
}

// Per-ring particle lists without atomics.  Entries are the work units
// (rings [ra + 64 ui, min(ra + 64 ui + 63, rb)]) and, per particle, the ring
// of its centre pixel when that lies outside [ra, rb]; an entry's key is
// first ring * kHpSpan + last ring (sentinel for entries without a ring).
// After a radix sort, the entries that can touch ring R are those whose first
// ring lies in [R - 63, R], one contiguous key range; those whose last ring
// is >= R touch it.  One block per ring counts them, a scan gives the list
// offsets, and one block per ring writes them in sorted order.
constexpr uint32_t kHpSpan = 8192;  // > 4 Nside for Nside <= kHpRingLds / 4

__device__ __forceinline__ int hp_entry_particle(int64_t e, int64_t nu,
                                                 const int2 *__restrict__ units)
{
    return e < nu ? units[e].x : (int)(e - nu);
}

__global__ void k_hp_entry_keys(const int2 *__restrict__ units, int64_t nu,
                                const HpRec *__restrict__ rec, int64_t n, int64_t nside,
                                uint32_t *__restrict__ keys, int32_t *__restrict__ vals)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nu + n)
        return;
    const uint32_t none = (uint32_t)(4 * nside + 1) * kHpSpan;
    uint32_t key = none;
    const int p = hp_entry_particle(e, nu, units);
    const HpRec r = rec[p];
    int64_t cring, ck;
    bool extra;
    if (hp_dep_rings(nside, r, cring, ck, extra)) {
        if (e < nu) {
            const int64_t a = (int64_t)r.ra + (int64_t)units[e].y * kHpRingsPerUnit;
            const int64_t bb = a + kHpRingsPerUnit - 1 < r.rb ? a + kHpRingsPerUnit - 1 : r.rb;
            if (bb >= a)
                key = (uint32_t)a * kHpSpan + (uint32_t)bb;
        } else if (extra) {
            key = (uint32_t)cring * kHpSpan + (uint32_t)cring;
        }
    }
    keys[e] = key;
    vals[e] = p;  // the sort is stable: equal keys keep the entries' order
}

// off[t] = first sorted entry with first ring >= t, t in [0, 4 Nside]
__global__ void k_hp_entry_offsets(const uint32_t *__restrict__ keys, int64_t ne, int64_t nt,
                                   int64_t *__restrict__ off)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t > nt)
        return;
    int64_t lo = 0, hi = ne;
    const uint32_t key = (uint32_t)t * kHpSpan;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (keys[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    off[t] = lo;
}

// count[iz] = sorted entries touching ring iz (count[0] = count[4 Nside] = 0)
__global__ __launch_bounds__(256) void k_hp_ring_count(const uint32_t *__restrict__ keys,
                                                       const int64_t *__restrict__ eoff,
                                                       int64_t nside,
                                                       int64_t *__restrict__ count)
{
    __shared__ int s_part[256 / kWave];
    const int64_t iz = blockIdx.x;  // 0 .. 4 Nside
    if (iz == 0 || iz == 4 * nside) {
        if (threadIdx.x == 0)
            count[iz] = 0;
        return;
    }
    const int64_t b = eoff[iz >= kHpRingsPerUnit - 1 ? iz - (kHpRingsPerUnit - 1) : 0];
    const int64_t e = eoff[iz + 1];
    int c = 0;
    for (int64_t i = b + threadIdx.x; i < e; i += 256)
        c += (int64_t)(keys[i] % kHpSpan) >= iz ? 1 : 0;
    for (int o = 32; o > 0; o >>= 1)
        c += __shfl_xor(c, o, kWave);
    if (threadIdx.x % kWave == 0)
        s_part[threadIdx.x / kWave] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < 256 / kWave; ++w)
            t += s_part[w];
        count[iz] = t;
    }
}

// list[loff[iz] ...] = the particles of the entries touching ring iz, in
// sorted order (block-wide ordered compaction; the sorted values are the
// entries' particles)
__global__ __launch_bounds__(256) void k_hp_ring_fill(const uint32_t *__restrict__ keys,
                                                      const int32_t *__restrict__ vals,
                                                      const int64_t *__restrict__ eoff,
                                                      const int64_t *__restrict__ loff,
                                                      int64_t nside,
                                                      int32_t *__restrict__ list)
{
    __shared__ int s_wc[256 / kWave];
    const int64_t iz = blockIdx.x + 1;  // rings 1 .. 4 Nside - 1
    const int64_t b = eoff[iz >= kHpRingsPerUnit - 1 ? iz - (kHpRingsPerUnit - 1) : 0];
    const int64_t e = eoff[iz + 1];
    const int wv = threadIdx.x / kWave, ln = threadIdx.x % kWave;
    const uint64_t below = ln == 0 ? 0ull : (~0ull >> (kWave - ln));
    int64_t out = loff[iz];
    for (int64_t c0 = b; c0 < e; c0 += 256) {
        const int64_t i = c0 + threadIdx.x;
        const bool take = i < e && (int64_t)(keys[i] % kHpSpan) >= iz;
        const uint64_t bt = __ballot(take);
        if (ln == 0)
            s_wc[wv] = (int)__popcll(bt);
        __syncthreads();
        int before = 0, total = 0;
        for (int w = 0; w < 256 / kWave; ++w) {
            before += w < wv ? s_wc[w] : 0;
            total += s_wc[w];
        }
        if (take)
            list[out + before + (int)__popcll(bt & below)] = vals[i];
        out += total;
        __syncthreads();
    }
}

// one (particle, ring) chord of the ring-major deposit (per-wave LDS table)
template <int NM>
struct HpPair {
    int off;    // first position of the chord in the wave's pixel list
    int lo;     // ring-local index of the chord's first pixel (< 0: wraps)
    int ck;     // ring-local index of the centre pixel, or -1 if not added
    int fb;     // fallback: area weights only
    float s1, s2, base, step;
    float theta, inv_theta;  // theta + ang / 2 (hp_weights's thh), 1 / theta
    double sw[NM], bq[NM];   // per map: pixel scale and quantity
};

// Ring of block i.  The hardware hands consecutive blocks to the eight XCDs
// in turn (block i to XCD i mod 8), each with its own L2; here every XCD
// takes runs of kHpXcdRun consecutive rings, so the blocks resident on it
// read the particle records of neighbouring rings (a particle spans many)
// through one L2, and the runs go to the XCDs in turn, which keeps their
// work balanced (one band per XCD would give the polar ones a quarter of
// the pixels).  The last G mod (8 kHpXcdRun) blocks keep their own index.
constexpr int kXcds = 8, kHpXcdRun = 16;
__device__ __forceinline__ int64_t hp_xcd_ring(int64_t i, int64_t G)
{
    const int64_t span = (int64_t)kXcds * kHpXcdRun;
    if (i >= G / span * span)
        return i;
    const int64_t x = i % kXcds, t = i / kXcds;
    return t / kHpXcdRun * span + x * kHpXcdRun + t % kHpXcdRun;
}

template <typename T, int K, int NM>
__global__ __launch_bounds__(kHpRingThreads<NM>, 4) void k_hp_ring_deposit(
    const int32_t *__restrict__ list, const int64_t *__restrict__ off,
    const HpRec *__restrict__ rec, const HpGroup g, uint32_t flags, int64_t nside)
{
    constexpr int kThreads = kHpRingThreads<NM>, kPairs = kHpRingPairs<NM>;
    __shared__ double s_m[NM][kHpRingLds], s_w[NM][kHpRingLds];
    __shared__ HpPair<NM> s_tab[kThreads / kWave][kPairs];
    const int64_t iz = hp_xcd_ring(blockIdx.x, gridDim.x) + 1;  // ring, 1-based
    int64_t start, nr;
    double shift;
    hp_ring_info(nside, iz, start, nr, shift);
    for (int k = threadIdx.x; k < nr; k += kThreads) {
#pragma unroll
        for (int q = 0; q < NM; ++q) {
            s_m[q][k] = 0.0;
            s_w[q][k] = 0.0;
        }
    }
    __syncthreads();
    const int wv = threadIdx.x / kWave, ln = threadIdx.x % kWave;
    HpPair<NM> *tab = s_tab[wv];
    const float ang = (float)sqrt(4.0 * kPi / (double)(12 * nside * nside));
    const float inv_ang = 1.0f / ang;
    const int64_t b = off[iz], e = off[iz + 1];
    for (int64_t c0 = b + (int64_t)wv * kPairs; c0 < e; c0 += (kThreads / kWave) * kPairs) {
        // phase 1: lane l < kPairs takes pair c0 + l: its chord on this ring
        HpPair<NM> q;
        q.lo = 0;
        q.ck = -1;
        q.fb = 0;
        q.s1 = q.s2 = q.base = q.step = q.theta = q.inv_theta = 0.0f;
#pragma unroll
        for (int k = 0; k < NM; ++k)
            q.sw[k] = q.bq[k] = 0.0;
        int len = 0;
        if (ln < kPairs && c0 + ln < e) {
            const int p = list[c0 + ln];
            const HpRec r = rec[p];
            int64_t cring, ck;
            bool extra;
            hp_dep_rings(nside, r, cring, ck, extra);
            int64_t a = 0, bb = -1;
            bool any = false;
            if (iz >= r.ra && iz <= r.rb) {
                if (iz < r.irmin || iz > r.irmax) {  // a ring fully inside a polar cap
                    hp_full_chord(nr, shift, r.phi0, a, bb);
                    any = true;
                } else {
                    any = hp_ring_interval(nside, iz, r.z0, r.xa, r.cosr, r.phi0, a, bb);
                }
            }
            const RingTerms t = ring_terms(nside, iz, r, g.rtab);
            if (any) {
                q.lo = (int)a;
                len = (int)(bb - a + 1);
                q.base = (float)hp_chord_start(a, t);
                q.step = (float)t.dphi_pix;
            }
            q.s1 = t.s1;
            q.s2 = t.s2;
            // constributing_pixels.jl:16-19: the centre pixel, if the disc
            // query did not return it
            if (iz == cring) {
                const bool cov = any && ((a >= 0 && ck >= a && ck <= bb) ||
                                         (a < 0 && (ck <= bb || ck >= a + nr)));
                if (!cov) {
                    q.ck = (int)ck;
                    len += 1;
                }
            }
            q.fb = r.fb;
            q.theta = (float)r.theta + 0.5f * ang;  // theta + ang / 2, as the norm pass
            q.inv_theta = (float)(1.0 / r.theta);
#pragma unroll
            for (int k = 0; k < NM; ++k)
                hp_map_scale<T>(g, k, p, r.s, flags, q.sw[k], q.bq[k]);
        }
        const int incl = wave_incl_scan(len, ln);
        const int total = __shfl(incl, kWave - 1, kWave);
        q.off = incl - len;
        __builtin_amdgcn_wave_barrier();
        if (ln < kPairs)
            tab[ln] = q;
        __builtin_amdgcn_wave_barrier();
        // phase 2: the chords' pixels, 64 list positions at a time
        int j = -1, next = 0;
        HpPair<NM> cur = q;
        int clen = 0;
        for (int pos = ln; pos < total; pos += kWave) {
            while (pos >= next) {
                ++j;
                cur = tab[j];
                next = j + 1 < kPairs ? tab[j + 1].off : total;
                clen = next - cur.off - (cur.ck >= 0 ? 1 : 0);
            }
            const int c = pos - cur.off;
            int k;
            float d;
            if (c < clen) {
                k = cur.lo + c;
                k = k < 0 ? k + (int)nr : k;
                // in [-pi, pi] up to FP32 rounding, as in hp_unit_pixels
                d = hp_angle(fmaf((float)c, cur.step, cur.base), cur.s1, cur.s2);
            } else {
                // the centre pixel, as the norm pass computes it
                k = cur.ck;
                const HpRec &r = rec[list[c0 + j]];
                d = hp_dist(ring_terms(nside, iz, r, g.rtab), k);
            }
            float W, A;
            hp_weights<K>(d, cur.theta, cur.inv_theta, ang, inv_ang, W, A);
            const double wa = (double)((cur.fb ? 1.0f : W) * A);
#pragma unroll
            for (int m = 0; m < NM; ++m) {
                const double pw = wa * cur.sw[m];
                if (pw != 0.0) {
                    unsafeAtomicAdd(&s_w[m][k], pw);
                    unsafeAtomicAdd(&s_m[m][k], cur.bq[m] * pw);
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    for (int k = threadIdx.x; k < nr; k += kThreads) {
#pragma unroll
        for (int m = 0; m < NM; ++m) {
            g.map[m][start + k] = s_m[m][k];
            g.wmap[m][start + k] = s_w[m][k];
        }
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
#define S2G_HP_KERNEL_SWITCH(kernel, CALL)                                      \
    switch (kernel) {                                                           \
    case S2G_KERNEL_CUBIC: { constexpr int KK = S2G_KERNEL_CUBIC; CALL; } break; \
    case S2G_KERNEL_QUINTIC: { constexpr int KK = S2G_KERNEL_QUINTIC; CALL; } break; \
    case S2G_KERNEL_WENDLAND_C2: { constexpr int KK = S2G_KERNEL_WENDLAND_C2; CALL; } break; \
    case S2G_KERNEL_WENDLAND_C4: { constexpr int KK = S2G_KERNEL_WENDLAND_C4; CALL; } break; \
    case S2G_KERNEL_WENDLAND_C6: { constexpr int KK = S2G_KERNEL_WENDLAND_C6; CALL; } break; \
    default: { constexpr int KK = S2G_KERNEL_WENDLAND_C8; CALL; } break;       \
    }

void launch_hp_setup(const HpArgs &a, HpRec *rec, int64_t *nunits, hipStream_t s)
{
    const dim3 g((unsigned)((a.n + 255) / 256)), blk(256);
    if (a.dtype == S2G_F32)
        hipLaunchKernelGGL(k_hp_setup<float>, g, blk, 0, s, a, rec, nunits);
    else
        hipLaunchKernelGGL(k_hp_setup<double>, g, blk, 0, s, a, rec, nunits);
}

void launch_hp_units(const int64_t *ubase, int64_t n, HpRec *rec, int2 *units, hipStream_t s)
{
    hipLaunchKernelGGL(k_hp_units, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ubase, n,
                       rec, units);
}

void launch_hp_norm(int kernel, const int2 *units, int64_t nu, const HpRec *rec, int64_t nside,
                    const double4 *rtab, double4 *part, hipStream_t s)
{
    if (nu == 0)
        return;
    const dim3 g((unsigned)wave_blocks(nu, 256)), blk(256);
    S2G_HP_KERNEL_SWITCH(kernel, hipLaunchKernelGGL(k_hp_norm<KK>, g, blk, 0, s, units, nu, rec,
                                                    nside, rtab, part));
}

void launch_hp_finish(const int64_t *ubase, int64_t n, const double4 *part, HpRec *rec,
                      hipStream_t s)
{
    hipLaunchKernelGGL(k_hp_finish, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ubase, n,
                       part, rec);
}

template <typename T, int K>
static void hp_deposit_k(const HpArgs &a, const HpGroup &g, const int2 *units, int64_t nu,
                         const HpRec *rec, hipStream_t s)
{
    const dim3 grid((unsigned)wave_blocks(nu, 256)), blk(256);
    if (g.nm == 1)
        hipLaunchKernelGGL((k_hp_deposit<T, K, 1>), grid, blk, 0, s, units, nu, rec, g, a.flags,
                           a.nside);
    else
        hipLaunchKernelGGL((k_hp_deposit<T, K, 2>), grid, blk, 0, s, units, nu, rec, g, a.flags,
                           a.nside);
}

void launch_hp_deposit(const HpArgs &a, const HpGroup &g, const int2 *units, int64_t nu,
                       const HpRec *rec, hipStream_t s)
{
    if (nu == 0)
        return;
    if (a.dtype == S2G_F32) {
        S2G_HP_KERNEL_SWITCH(a.kernel, (hp_deposit_k<float, KK>(a, g, units, nu, rec, s)));
    } else {
        S2G_HP_KERNEL_SWITCH(a.kernel, (hp_deposit_k<double, KK>(a, g, units, nu, rec, s)));
    }
}

bool hp_ring_deposit_fits(int64_t nside) { return 4 * nside <= kHpRingLds; }

void launch_hp_entry_keys(const int2 *units, int64_t nu, const HpRec *rec, int64_t n,
                          int64_t nside, uint32_t *keys, int32_t *vals, hipStream_t s)
{
    const int64_t ne = nu + n;
    hipLaunchKernelGGL(k_hp_entry_keys, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s,
                       units, nu, rec, n, nside, keys, vals);
}

int hp_entry_key_bits(int64_t nside)
{
    const uint64_t top = (uint64_t)(4 * nside + 1) * kHpSpan;
    int b = 1;
    while (b < 32 && ((uint64_t)1 << b) <= top)
        ++b;
    return b;
}

void launch_hp_entry_offsets(const uint32_t *keys, int64_t ne, int64_t nside, int64_t *off,
                             hipStream_t s)
{
    const int64_t nt = 4 * nside;
    hipLaunchKernelGGL(k_hp_entry_offsets, dim3((unsigned)((nt + 1 + 255) / 256)), dim3(256), 0,
                       s, keys, ne, nt, off);
}

void launch_hp_ring_count(const uint32_t *keys, const int64_t *eoff, int64_t nside,
                          int64_t *count, hipStream_t s)
{
    hipLaunchKernelGGL(k_hp_ring_count, dim3((unsigned)(4 * nside + 1)), dim3(256), 0, s, keys,
                       eoff, nside, count);
}

void launch_hp_ring_fill(const uint32_t *keys, const int32_t *vals, const int64_t *eoff,
                         const int64_t *loff, int64_t nside, int32_t *list, hipStream_t s)
{
    hipLaunchKernelGGL(k_hp_ring_fill, dim3((unsigned)(4 * nside - 1)), dim3(256), 0, s, keys,
                       vals, eoff, loff, nside, list);
}

template <typename T, int K>
static void hp_ring_deposit_k(const HpArgs &a, const HpGroup &g, const int32_t *list,
                              const int64_t *off, const HpRec *rec, hipStream_t s)
{
    const dim3 grid((unsigned)(4 * a.nside - 1));
    if (g.nm == 1)
        hipLaunchKernelGGL((k_hp_ring_deposit<T, K, 1>), grid, dim3(kHpRingThreads<1>), 0, s,
                           list, off, rec, g, a.flags, a.nside);
    else
        hipLaunchKernelGGL((k_hp_ring_deposit<T, K, 2>), grid, dim3(kHpRingThreads<2>), 0, s,
                           list, off, rec, g, a.flags, a.nside);
}

void launch_hp_ring_deposit(const HpArgs &a, const HpGroup &g, const int32_t *list,
                            const int64_t *off, const HpRec *rec, hipStream_t s)
{
    if (a.dtype == S2G_F32) {
        S2G_HP_KERNEL_SWITCH(a.kernel, (hp_ring_deposit_k<float, KK>(a, g, list, off, rec, s)));
    } else {
        S2G_HP_KERNEL_SWITCH(a.kernel, (hp_ring_deposit_k<double, KK>(a, g, list, off, rec, s)));
    }
}

}  // namespace s2g
