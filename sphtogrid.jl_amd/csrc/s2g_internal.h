// s2g_internal.h -- shared device helpers and host-side declarations of
// libsph2grid (gfx950, wave64).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/sph2grid.h"

namespace s2g {

constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// Particle classes of the 2D/3D deposits
// ---------------------------------------------------------------------------
enum : int32_t {
    kClsSkip = 0,      // zero quantity without calc_mean, or empty footprint
    kClsSmall = 1,     // footprint box fits the scatter tile halo
    kClsLargeAn = 2,   // gather path, closed-form normalisation
    kClsLargeEx = 3,   // gather path, explicit normalisation sum
};

// Largest footprint side (pixels) of the scatter path.  A box of side <= 34
// corresponds to hsml <= 16 pixels (side = floor(x+h) - floor(x-h) + 1).
constexpr int kSmallSide2D = 34;
// side of the small-footprint tiles (a particle's bucket is the tile holding
// its footprint's first pixel)
constexpr int kST = 32;
// 2D gather tile: 64 x 64 pixels, 4 waves x 16 rows x 64 columns.
constexpr int kGT = 64;
constexpr int kGRows = 16;

// 3D: footprint box side <= 8 voxels goes to the scatter path; bricks of 16^3.
constexpr int kSmallSide3D = 8;
constexpr int kBT = 16;

// Per-particle record of the 2D deposit, built once per call.
struct Rec2 {
    double x, y;     // pixel coordinates, cic_shared.jl:85-100
    double h;        // hsml in pixels, cic_2D.jl:83
    double s;        // area*w*dz (small / exact) or the final pixel scale (large)
    int32_t i0, i1, j0, j1;  // footprint, cic_shared.jl:46-52
    int16_t cls;
    int16_t fb;      // fallback (no pixel centre inside the kernel), cic_2D.jl:51-69
    int32_t map;     // map of the deposit pass the particle belongs to
};

// Per-particle record of the 3D deposit (cic_3D.jl:87-97, 150-155).
struct Rec3 {
    double x, y, z;  // voxel coordinates
    double h;        // hsml in voxels
    double s;        // vol*w*len2pix (small / exact) or the final voxel scale (large)
    int32_t i0, i1, j0, j1, k0, k1;
    int32_t cls;
    int32_t fb;
};

// ---------------------------------------------------------------------------
// Kernel shapes (SPHKernels.jl closed forms).  shape(u) for u < 1, 0 beyond.
// The norm * h_inv^dim factor cancels in every deposit (pw ~ wk*A/sum(wk*A)).
// ---------------------------------------------------------------------------
template <int K, typename R>
__device__ __forceinline__ R shape_of(R u)
{
    const R one = R(1);
    R t = one - u;
    t = t > R(0) ? t : R(0);
    if constexpr (K == S2G_KERNEL_CUBIC) {
        // u < 1/2: 1 + 6 (u-1) u^2 ; 1/2 <= u < 1: 2 (1-u)^3
        R a = R(1) + R(6) * (u - one) * u * u;
        R b = R(2) * t * t * t;
        return u < R(0.5) ? a : b;
    } else if constexpr (K == S2G_KERNEL_QUINTIC) {
        R t2 = t * t;
        R v = t2 * t2 * t;
        R a = R(2.0 / 3.0) - u;
        a = a > R(0) ? a : R(0);
        R a2 = a * a;
        R b = R(1.0 / 3.0) - u;
        b = b > R(0) ? b : R(0);
        R b2 = b * b;
        return v - R(6) * (a2 * a2 * a) + R(15) * (b2 * b2 * b);
    } else if constexpr (K == S2G_KERNEL_WENDLAND_C2) {
        R t2 = t * t;
        return t2 * t2 * (R(1) + R(4) * u);
    } else if constexpr (K == S2G_KERNEL_WENDLAND_C4) {
        R t2 = t * t;
        R p = R(1) + u * (R(6) + u * R(35.0 / 3.0));
        return t2 * t2 * t2 * p;
    } else if constexpr (K == S2G_KERNEL_WENDLAND_C6) {
        R t2 = t * t;
        R t4 = t2 * t2;
        R p = R(1) + u * (R(8) + u * (R(25) + u * R(32)));
        return t4 * t4 * p;
    } else {  // WENDLAND_C8
        R t2 = t * t;
        R t4 = t2 * t2;
        R p = R(1) + u * (R(10) + u * (R(42) + u * (R(90) + u * R(429.0 / 5.0))));
        return t4 * t4 * t2 * p;
    }
}

// Power n of the (1 - u) factor of each kernel shape.  The gather folds a
// row's x overlap ax into that factor as a = ax^(1/n), so the pixel weight
// ax * shape(u) costs no multiply of its own; a^n equals ax to a few FP32
// ulps.  ax < 1 only in the first and last rows of a footprint
// (cic_shared.jl:60-76), ax = 0 outside it.
template <int K>
constexpr int kTPow = K == S2G_KERNEL_CUBIC          ? 3
                      : K == S2G_KERNEL_QUINTIC      ? 5
                      : K == S2G_KERNEL_WENDLAND_C2  ? 4
                      : K == S2G_KERNEL_WENDLAND_C4  ? 6
                      : K == S2G_KERNEL_WENDLAND_C6  ? 8
                                                     : 10;

// ax * shape(u) with a = ax^(1/n): the SPHKernels closed forms with every
// (c - u)_+ factor written as a (c - u) clamped to [0, 1] (the clamp is the
// VALU output modifier)
template <int K>
__device__ __forceinline__ float shape_row(float u, float a)
{
    const float t = __builtin_amdgcn_fmed3f(fmaf(-a, u, a), 0.0f, 1.0f);
    if constexpr (K == S2G_KERNEL_CUBIC) {
        // 2 (1-u)^3 - 8 (1/2-u)^3_+ (= 1 + 6 (u-1) u^2 below u = 1/2)
        const float v = a * __builtin_amdgcn_fmed3f(0.5f - u, 0.0f, 1.0f);
        return fmaf(-8.0f, v * v * v, 2.0f * (t * t * t));
    } else if constexpr (K == S2G_KERNEL_QUINTIC) {
        const float p = a * __builtin_amdgcn_fmed3f(2.0f / 3.0f - u, 0.0f, 1.0f);
        const float q = a * __builtin_amdgcn_fmed3f(1.0f / 3.0f - u, 0.0f, 1.0f);
        const float t2 = t * t, p2 = p * p, q2 = q * q;
        return fmaf(15.0f, q2 * q2 * q, fmaf(-6.0f, p2 * p2 * p, t2 * t2 * t));
    } else if constexpr (K == S2G_KERNEL_WENDLAND_C2) {
        const float t2 = t * t;
        return t2 * t2 * fmaf(4.0f, u, 1.0f);
    } else if constexpr (K == S2G_KERNEL_WENDLAND_C4) {
        const float t2 = t * t;
        return t2 * t2 * t2 * fmaf(u, fmaf(u, 35.0f / 3.0f, 6.0f), 1.0f);
    } else if constexpr (K == S2G_KERNEL_WENDLAND_C6) {
        const float t2 = t * t, t4 = t2 * t2;
        return t4 * t4 * fmaf(u, fmaf(u, fmaf(u, 32.0f, 25.0f), 8.0f), 1.0f);
    } else {  // WENDLAND_C8
        const float t2 = t * t, t4 = t2 * t2;
        return t4 * t4 * t2 *
               fmaf(u, fmaf(u, fmaf(u, fmaf(u, 429.0f / 5.0f, 90.0f), 42.0f), 10.0f), 1.0f);
    }
}

// a = ax^(1/n) of a row (or column) overlap ax, as shape_row takes it
template <int K>
__device__ __forceinline__ float overlap_root(float ax)
{
    // v_log_f32 / v_exp_f32 directly: ax is 0 or a normal number here
    return ax >= 1.0f   ? 1.0f
           : ax > 0.0f ? __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(ax) * (1.0f / kTPow<K>))
                       : 0.0f;
}

// integral of shape(|v|) over the unit disc (= 1 / norm_2D), used for the
// closed-form normalisation of large unclipped footprints
__host__ __device__ inline double shape_disc_integral(int kernel)
{
    const double pi = 3.14159265358979323846;
    switch (kernel) {
    case S2G_KERNEL_CUBIC: return 7.0 * pi / 40.0;
    case S2G_KERNEL_QUINTIC: return 478.0 * pi / 15309.0;
    case S2G_KERNEL_WENDLAND_C2: return pi / 7.0;
    case S2G_KERNEL_WENDLAND_C4: return pi / 9.0;
    case S2G_KERNEL_WENDLAND_C6: return 7.0 * pi / 78.0;
    default: return 3.0 * pi / 40.0;  // WENDLAND_C8
    }
}

// Smallest hsml (pixels) at which the closed-form normalisation h^2 * I_k
// matches the discrete footprint sum to better than ~1e-9 relative
// (midpoint-rule error of a compactly supported C^k function; measured over
// random sub-pixel offsets, see tests/test_norm_closed_form.py).
__host__ __device__ inline double closed_form_min_h(int kernel)
{
    switch (kernel) {
    case S2G_KERNEL_CUBIC: return 64.0;
    case S2G_KERNEL_QUINTIC: return 24.0;
    case S2G_KERNEL_WENDLAND_C2: return 64.0;
    case S2G_KERNEL_WENDLAND_C4: return 28.0;
    case S2G_KERNEL_WENDLAND_C6: return 16.0;
    default: return 12.0;  // WENDLAND_C8
    }
}

// integral of shape(|v|) over the unit ball (= 1 / norm_3D)
__host__ __device__ inline double shape_ball_integral(int kernel)
{
    const double pi = 3.14159265358979323846;
    switch (kernel) {
    case S2G_KERNEL_CUBIC: return pi / 8.0;
    case S2G_KERNEL_QUINTIC: return 40.0 * pi / 2187.0;
    case S2G_KERNEL_WENDLAND_C2: return 2.0 * pi / 21.0;
    case S2G_KERNEL_WENDLAND_C4: return 32.0 * pi / 495.0;
    case S2G_KERNEL_WENDLAND_C6: return 64.0 * pi / 1365.0;
    default: return 64.0 * pi / 1785.0;  // WENDLAND_C8
    }
}

// 3D counterpart of closed_form_min_h (voxels; same 2e-9 bound)
__host__ __device__ inline double closed_form_min_h3(int kernel)
{
    switch (kernel) {
    case S2G_KERNEL_CUBIC: return 40.0;
    case S2G_KERNEL_QUINTIC: return 24.0;
    case S2G_KERNEL_WENDLAND_C2: return 32.0;
    case S2G_KERNEL_WENDLAND_C4: return 20.0;
    case S2G_KERNEL_WENDLAND_C6: return 14.0;
    default: return 11.0;  // WENDLAND_C8
    }
}

// ---------------------------------------------------------------------------
// Products and sums rounded on their own (cic_shared.jl:85-100 rounds
// x*len2pix and then +0.5*Npix).  HIP's __dmul_rn / __fmul_rn are plain
// operators here and hipcc contracts them into FMAs; an operation emitted
// under `fp contract(off)` carries no contract flag, so it stays separate
// after inlining.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double mul_rn(double a, double b)
{
#pragma clang fp contract(off)
    return a * b;
}
__device__ __forceinline__ double add_rn(double a, double b)
{
#pragma clang fp contract(off)
    return a + b;
}
__device__ __forceinline__ double sub_rn(double a, double b)
{
#pragma clang fp contract(off)
    return a - b;
}
// an FP32 square summed with another must equal the sum of the squares held
// in a table (no FMA between them)
__device__ __forceinline__ float fmul_rn(float a, float b)
{
#pragma clang fp contract(off)
    return a * b;
}

// overlap of [x-h, x+h] with pixel [i, i+1] (cic_shared.jl:60-62)
__device__ __forceinline__ double overlap1(double x, double h, double fi)
{
    double a = add_rn(x, h), b = fi + 1.0;
    double c = sub_rn(x, h);
    return sub_rn(fmin(a, b), fmax(c, fi));
}
// offset of x from the centre of pixel i (cic_shared.jl:70-76)
__device__ __forceinline__ double coff1(double x, double fi)
{
    return sub_rn(sub_rn(x, fi), 0.5);
}

template <typename T>
__device__ __forceinline__ double ld(const T *p, int64_t i)
{
    return (double)p[i];
}

// Blocks for a one-wave-per-item kernel with a grid-stride loop.  The grid is
// capped so that blocks * threads stays far below 2^32 work-items (the
// dispatch grid size is 32-bit); larger counts loop.
inline int64_t wave_blocks(int64_t items, int threads_per_block)
{
    const int64_t waves_per_block = threads_per_block / kWave;
    int64_t b = (items + waves_per_block - 1) / waves_per_block;
    const int64_t cap = 1 << 18;
    return b < 1 ? 1 : (b > cap ? cap : b);
}

// v + the value of another lane of the same 16-lane row (DPP lane pattern)
template <int kCtrl>
__device__ __forceinline__ double dpp_add(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), kCtrl, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), kCtrl, 0xF, 0xF, true);
    return v + __hiloint2double(hi, lo);
}

// Sum over the 64 lanes of a fully active wave, returned on every lane.  DPP
// steps inside each 16-lane row (quad swaps, half-row and row mirrors), then
// the four row sums read from lanes 0, 16, 32, 48.  The order is fixed, so
// the result is reproducible.
__device__ __forceinline__ double wave_sum(double v)
{
    v = dpp_add<0xB1>(v);   // quad_perm [1, 0, 3, 2]
    v = dpp_add<0x4E>(v);   // quad_perm [2, 3, 0, 1]
    v = dpp_add<0x141>(v);  // row_half_mirror
    v = dpp_add<0x140>(v);  // row_mirror
    double r[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        r[k] = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 16 * k),
                                __builtin_amdgcn_readlane(__double2loint(v), 16 * k));
    return (r[0] + r[1]) + (r[2] + r[3]);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
#define S2G_HIP(call)                                                        \
    do {                                                                     \
        hipError_t e_ = (call);                                              \
        if (e_ != hipSuccess)                                                \
            throw ::s2g::Error(S2G_EHIP, std::string(#call) + ": " +         \
                                             hipGetErrorString(e_));         \
    } while (0)

struct Error {
    int code;
    std::string msg;
    Error(int c, std::string m) : code(c), msg(std::move(m)) {}
};

// grow-only device buffer
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    void *get(size_t need)
    {
        if (need > bytes) {
            if (p)
                (void)hipFree(p);
            p = nullptr;
            bytes = 0;
            size_t b = need < 256 ? 256 : need;
            b = (b + 255) & ~size_t(255);
            if (hipMalloc(&p, b) != hipSuccess) {
                p = nullptr;
                throw Error(S2G_ENOMEM, "device allocation of " + std::to_string(b) +
                                            " bytes failed");
            }
            bytes = b;
        }
        return p;
    }
    void release()
    {
        if (p)
            (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
};

// grow-only pinned host buffer (staging of small tables and read-backs).
// Only grown while no copy from or into it is in flight.
struct PinnedBuf {
    void *p = nullptr;
    size_t bytes = 0;
    void *get(size_t need)
    {
        if (need > bytes) {
            if (p)
                (void)hipHostFree(p);
            p = nullptr;
            bytes = 0;
            size_t b = need < 4096 ? 4096 : 2 * need;
            if (hipHostMalloc(&p, b, hipHostMallocDefault) != hipSuccess) {
                p = nullptr;
                throw Error(S2G_ENOMEM, "pinned host allocation of " + std::to_string(b) +
                                            " bytes failed");
            }
            bytes = b;
        }
        return p;
    }
    void release()
    {
        if (p)
            (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
    }
};

// exclusive scan of n int64 values in place; returns nothing, device only.
// tmp must hold scan_tmp_elems(n) int64.
size_t scan_tmp_elems(int64_t n);
void exclusive_scan_i64(int64_t *d, int64_t n, int64_t *tmp, hipStream_t s);

}  // namespace s2g
