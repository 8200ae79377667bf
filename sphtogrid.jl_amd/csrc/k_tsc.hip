// k_tsc.hip -- triangular-shaped-cloud deposit onto an npix^3 grid (textbook
// TSC, Hockney & Eastwood; the reference's TSC path is commented out,
// tsc_interpolation/tsc_interpolation.jl:1-182).
//
// Grid units follow the 3D deposit (get_xyz, cic_shared.jl:85-100):
// g = x*len2pix + npix/2, cell c has its centre at c + 1/2.  Per axis the
// weights of the nearest cell and its two neighbours are
//   W(d) = 3/4 - d^2          |d| < 1/2
//        = (3/2 - |d|)^2 / 2  1/2 <= |d| < 3/2
// and they sum to one.  Plane 0 accumulates q*W, plane 1 accumulates W.
#include "k_tsc.h"

namespace s2g {

__device__ __forceinline__ double tsc_w(double d)
{
    const double a = fabs(d);
    if (a < 0.5)
        return 0.75 - a * a;
    if (a < 1.5) {
        const double b = 1.5 - a;
        return 0.5 * b * b;
    }
    return 0.0;
}

// nearest cell of particle p along each axis, as k_tsc forms it; false if
// the particle is dropped (a non-finite coordinate, or far outside a
// non-periodic grid)
template <typename T>
__device__ __forceinline__ bool tsc_cell(const T *__restrict__ pos, int64_t p, int64_t npix,
                                         double len2pix, int periodic, int64_t c[3])
{
    for (int d = 0; d < 3; ++d) {
        const double g = add_rn(mul_rn(ld(pos, 3 * p + d), len2pix), 0.5 * (double)npix);
        const double fg = floor(g);
        if (!isfinite(fg) || ((fg < -2.0 || fg > (double)npix + 2.0) && !periodic))
            return false;
        c[d] = (int64_t)fg;
    }
    return true;
}

// Sort key: the 8^3 brick of the particle's nearest cell (row-major bricks;
// dropped particles last), so that a wave's 27-cell stencils share cache
// lines and the atomics stay in L2 instead of scattering over the grid.
template <typename T>
__global__ __launch_bounds__(256) void k_tsc_keys(const T *__restrict__ pos, int64_t n,
                                                  int64_t npix, double len2pix, int periodic,
                                                  uint32_t *__restrict__ keys,
                                                  int32_t *__restrict__ vals)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n)
        return;
    const int64_t nb = (npix + 7) / 8;
    int64_t c[3];
    uint32_t key = (uint32_t)(nb * nb * nb);
    if (tsc_cell(pos, p, npix, len2pix, periodic, c)) {
        int64_t b[3];
        for (int d = 0; d < 3; ++d) {
            int64_t x = periodic ? ((c[d] % npix) + npix) % npix : c[d];
            x = x < 0 ? 0 : x > npix - 1 ? npix - 1 : x;
            b[d] = x / 8;
        }
        key = (uint32_t)((b[0] * nb + b[1]) * nb + b[2]);
    }
    keys[p] = key;
    vals[p] = (int32_t)p;
}

constexpr int kTscRows = 3;  // threads per particle

template <typename T>
__global__ __launch_bounds__(256) void k_tsc(const T *__restrict__ pos, const T *__restrict__ binq,
                                             int64_t n, int64_t npix, double len2pix,
                                             int periodic, const int32_t *__restrict__ order,
                                             double *__restrict__ grid)
{
    // three threads per particle, one per x slab of the 27-cell stencil:
    // more atomics in flight than one thread issuing all 54 (2^22 particles
    // on 512^3: 8.9 ms with one thread, 8.4 with three, 8.5 with nine)
    const int64_t tt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t t = tt / kTscRows;
    const int slab = (int)(tt - kTscRows * t);
    if (t >= n)
        return;
    const int64_t p = order ? (int64_t)order[t] : t;
    const int64_t nd = npix * npix * npix;
    int64_t c[3];
    double w[3][3];
    bool bad = false;
    for (int d = 0; d < 3; ++d) {
        const double g = add_rn(mul_rn(ld(pos, 3 * p + d), len2pix), 0.5 * (double)npix);
        const double fg = floor(g);
        if (!isfinite(fg) || fg < -2.0 || fg > (double)npix + 2.0) {
            if (!periodic || !isfinite(fg))
                bad = true;
        }
        c[d] = bad ? 0 : (int64_t)fg;
        for (int o = -1; o <= 1; ++o)
            w[d][o + 1] = tsc_w(sub_rn(g, (double)(c[d] + o) + 0.5));
    }
    if (bad)
        return;
    const double q = ld(binq, p);
    int64_t i = c[0] + slab - 1;
    if (periodic)
        i = ((i % npix) + npix) % npix;
    else if (i < 0 || i >= npix)
        return;
    for (int b = -1; b <= 1; ++b) {
        int64_t j = c[1] + b;
        if (periodic)
            j = ((j % npix) + npix) % npix;
        else if (j < 0 || j >= npix)
            continue;
        const double wab = w[0][slab] * w[1][b + 1];
        for (int e = -1; e <= 1; ++e) {
            int64_t k = c[2] + e;
            if (periodic)
                k = ((k % npix) + npix) % npix;
            else if (k < 0 || k >= npix)
                continue;
            const double ww = wab * w[2][e + 1];
            if (ww == 0.0)
                continue;
            const int64_t idx = (i * npix + j) * npix + k;
            unsafeAtomicAdd(&grid[idx], q * ww);
            unsafeAtomicAdd(&grid[nd + idx], ww);
        }
    }
}

int tsc_key_bits(int64_t npix)
{
    const int64_t nb = (npix + 7) / 8;
    int bits = 1;
    while (bits < 32 && ((int64_t)1 << bits) <= nb * nb * nb)
        ++bits;
    return bits;
}

void launch_tsc_keys(int dtype, const void *pos, int64_t n, int64_t npix, double len2pix,
                     int periodic, uint32_t *keys, int32_t *vals, hipStream_t s)
{
    if (n == 0)
        return;
    const dim3 g((unsigned)((n + 255) / 256)), blk(256);
    if (dtype == S2G_F32)
        hipLaunchKernelGGL(k_tsc_keys<float>, g, blk, 0, s, (const float *)pos, n, npix,
                           len2pix, periodic, keys, vals);
    else
        hipLaunchKernelGGL(k_tsc_keys<double>, g, blk, 0, s, (const double *)pos, n, npix,
                           len2pix, periodic, keys, vals);
}

void launch_tsc(int dtype, const void *pos, const void *binq, int64_t n, int64_t npix,
                double len2pix, int periodic, const int32_t *order, double *grid, hipStream_t s)
{
    if (n == 0)
        return;
    const dim3 g((unsigned)((kTscRows * n + 255) / 256)), blk(256);
    if (dtype == S2G_F32)
        hipLaunchKernelGGL(k_tsc<float>, g, blk, 0, s, (const float *)pos, (const float *)binq,
                           n, npix, len2pix, periodic, order, grid);
    else
        hipLaunchKernelGGL(k_tsc<double>, g, blk, 0, s, (const double *)pos,
                           (const double *)binq, n, npix, len2pix, periodic, order, grid);
}

}  // namespace s2g
