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
//
// Pipeline (a gather: every cell is written once with a plain store, no
// atomics and no memset, and the result is bit-reproducible):
//   k_tsc_keys     brick of each particle's nearest cell (bricks of
//                  kTX x kTY x kTZ cells, row-major; dropped particles last)
//   (radix sort)   stable, so particles keep their input order in a brick
//   k_tsc_offsets  first sorted particle of every brick
//   k_tsc_prep     grid coordinates (FP64) and quantity in sorted order, and
//                  each particle's nearest cell packed into one word
//   k_tsc_gather   one wave per brick: scans the particles of the (up to) 27
//                  neighbouring bricks, queues those whose 3^3 stencil
//                  reaches the brick and deposits them into LDS seven at a
//                  time (9 lanes per particle, one stencil row each), in scan
//                  order; then writes the brick's cells.
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

__host__ __device__ __forceinline__ int64_t tsc_wrap(int64_t c, int64_t npix)
{
    return ((c % npix) + npix) % npix;
}

// nearest cell of particle p along each axis; false if the particle is
// dropped (a non-finite coordinate, or far outside a non-periodic grid)
template <typename T>
__device__ __forceinline__ bool tsc_cell(const T *__restrict__ pos, int64_t p, int64_t npix,
                                         double len2pix, int periodic, double g[3],
                                         int64_t c[3])
{
    for (int d = 0; d < 3; ++d) {
        g[d] = add_rn(mul_rn(ld(pos, 3 * p + d), len2pix), 0.5 * (double)npix);
        const double fg = floor(g[d]);
        if (!isfinite(fg) || ((fg < -2.0 || fg > (double)npix + 2.0) && !periodic))
            return false;
        c[d] = (int64_t)fg;
    }
    return true;
}

// the brick grid of an npix^3 grid
__host__ __device__ __forceinline__ void tsc_bricks(int64_t npix, int64_t nb[3])
{
    nb[0] = (npix + kTX - 1) / kTX;
    nb[1] = (npix + kTY - 1) / kTY;
    nb[2] = (npix + kTZ - 1) / kTZ;
}

int64_t tsc_nbricks(int64_t npix)
{
    int64_t nb[3];
    tsc_bricks(npix, nb);
    return nb[0] * nb[1] * nb[2];
}

// Sort key: the brick of the particle's nearest cell (clamped into the grid,
// or wrapped when periodic); dropped particles get nbricks, after all bricks.
template <typename T>
__global__ __launch_bounds__(256) void k_tsc_keys(const T *__restrict__ pos, int64_t n,
                                                  int64_t npix, double len2pix, int periodic,
                                                  uint32_t *__restrict__ keys,
                                                  int32_t *__restrict__ vals)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n)
        return;
    int64_t nb[3];
    tsc_bricks(npix, nb);
    double g[3];
    int64_t c[3];
    uint32_t key = (uint32_t)(nb[0] * nb[1] * nb[2]);
    if (tsc_cell(pos, p, npix, len2pix, periodic, g, c)) {
        const int64_t side[3] = {kTX, kTY, kTZ};
        int64_t b[3];
        for (int d = 0; d < 3; ++d) {
            int64_t x = periodic ? tsc_wrap(c[d], npix) : c[d];
            x = x < 0 ? 0 : x > npix - 1 ? npix - 1 : x;
            b[d] = x / side[d];
        }
        key = (uint32_t)((b[0] * nb[1] + b[1]) * nb[2] + b[2]);
    }
    keys[p] = key;
    vals[p] = (int32_t)p;
}

// off[b] = first sorted entry with key >= b, b in [0, nbricks]
__global__ __launch_bounds__(256) void k_tsc_offsets(const uint32_t *__restrict__ keys,
                                                     int64_t n, int64_t nbricks,
                                                     int64_t *__restrict__ off)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b > nbricks)
        return;
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if ((int64_t)keys[mid] < b)
            lo = mid + 1;
        else
            hi = mid;
    }
    off[b] = lo;
}

// packed nearest cell: 21 bits per axis, each stored as c + kCellBias (c is
// in [-2, npix + 2] when not periodic, wrapped into [0, npix) when periodic)
constexpr int64_t kCellBias = 4;
__device__ __forceinline__ uint64_t tsc_pack(const int64_t c[3])
{
    return (uint64_t)(c[0] + kCellBias) << 42 | (uint64_t)(c[1] + kCellBias) << 21 |
           (uint64_t)(c[2] + kCellBias);
}
__device__ __forceinline__ int64_t tsc_unpack(uint64_t v, int d)
{
    return (int64_t)((v >> (42 - 21 * d)) & 0x1FFFFF) - kCellBias;
}

// sorted grid coordinates and quantity (gq) and packed cells (cell) of the
// particles that are not dropped (the first off[nbricks] sorted entries)
template <typename T>
__global__ __launch_bounds__(256) void k_tsc_prep(const T *__restrict__ pos,
                                                  const T *__restrict__ binq,
                                                  const int32_t *__restrict__ order,
                                                  const int64_t *__restrict__ off,
                                                  int64_t nbricks, int64_t npix, double len2pix,
                                                  int periodic, double4 *__restrict__ gq,
                                                  uint64_t *__restrict__ cell)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= off[nbricks])
        return;
    const int64_t p = order[k];
    double g[3];
    int64_t c[3];
    tsc_cell(pos, p, npix, len2pix, periodic, g, c);
    if (periodic)
        for (int d = 0; d < 3; ++d)
            c[d] = tsc_wrap(c[d], npix);
    gq[k] = make_double4(g[0], g[1], g[2], ld(binq, p));
    cell[k] = tsc_pack(c);
}

// bitmask of the stencil offsets o in {-1, 0, 1} (bit o + 1) whose cell
// c + o (wrapped when periodic) lies in [b0, b0 + side) and in the grid
__device__ __forceinline__ int tsc_reach(int64_t c, int64_t b0, int64_t side, int64_t npix,
                                         int periodic)
{
    int m = 0;
#pragma unroll
    for (int o = -1; o <= 1; ++o) {
        int64_t x = c + o;
        if (periodic)
            x = x < 0 ? x + npix : x >= npix ? x - npix : x;
        const bool in = x >= b0 && x < b0 + side && x >= 0 && x < npix;
        m |= in ? 1 << (o + 1) : 0;
    }
    return m;
}

// particles per deposit step: 9 lanes each, one stencil row (o0, o1) per
// lane and its three cells along z
constexpr int kTscPerStep = kWave / 9;
// Deposit queue of the gather: the candidates whose stencil reaches the
// brick, in scan order (a ring of LDS slots holding sorted particle
// indices; the deposit step loads their records).  It holds at most
// kTscPerStep - 1 leftovers plus one chunk of kWave candidates.  The LDS of
// a block, not the VGPRs, bounds the waves per CU: with 32-byte records in
// 128 slots 19 waves, in 72 slots 23, with 4-byte indices 33.
constexpr int kTscQueue = 72;
static_assert(kTscQueue >= kTscPerStep - 1 + kWave, "queue holds one chunk and the leftovers");

// One deposit step: particles slots [h, h + cnt) of the queue.  The
// weights are formed as the oracle forms them: per axis W(g - (c + o) - 1/2)
// with c = floor(g), and the product (wx * wy) * wz.
__device__ __forceinline__ void tsc_step(const double4 *__restrict__ gq, const int32_t *queue,
                                         int h, int cnt, int ln,
                                         const int64_t b0[3], int64_t npix, int periodic,
                                         double (*acc)[kTX * kTY * kTZ])
{
    const int slot = ln / 9, row = ln - 9 * slot;
    if (slot >= cnt)
        return;
    const double4 e = gq[queue[(h + slot) % kTscQueue]];
    const double g[3] = {e.x, e.y, e.z};
    int64_t c[3];
    for (int d = 0; d < 3; ++d)
        c[d] = (int64_t)floor(g[d]);
    const int o0 = row / 3 - 1, o1 = row % 3 - 1;
    const double wx = tsc_w(sub_rn(g[0], (double)(c[0] + o0) + 0.5));
    const double wy = tsc_w(sub_rn(g[1], (double)(c[1] + o1) + 0.5));
    int64_t x = c[0] + o0, y = c[1] + o1;
    if (periodic) {
        x = tsc_wrap(x, npix);
        y = tsc_wrap(y, npix);
    }
    const int64_t lx = x - b0[0], ly = y - b0[1];
    if (x < 0 || x >= npix || lx < 0 || lx >= kTX || y < 0 || y >= npix || ly < 0 ||
        ly >= kTY)
        return;
    const double wxy = wx * wy;
#pragma unroll
    for (int o2 = -1; o2 <= 1; ++o2) {
        int64_t z = c[2] + o2;
        if (periodic)
            z = tsc_wrap(z, npix);
        const int64_t lz = z - b0[2];
        const double ww = wxy * tsc_w(sub_rn(g[2], (double)(c[2] + o2) + 0.5));
        if (z >= 0 && z < npix && lz >= 0 && lz < kTZ && ww != 0.0) {
            const int cell = (int)((lx * kTY + ly) * kTZ + lz);
            unsafeAtomicAdd(&acc[0][cell], e.w * ww);
            unsafeAtomicAdd(&acc[1][cell], ww);
        }
    }
}

// One wave per brick.  Lane j < 27 takes neighbouring brick j (offsets
// (j/9 - 1, (j/3)%3 - 1, j%3 - 1), each brick once) and reads its particle
// range; the ranges are laid end to end (wave prefix sum) and scanned 64
// candidates at a time (packed cells only; the next chunk's cells are loaded
// before the current chunk is handled).  The candidates whose stencil
// reaches the brick load their coordinates into the LDS queue in scan order,
// and the queue is deposited kTscPerStep particles per step.  The order of
// the additions to a cell follows the scan order, so the sums are
// reproducible.
__global__ __launch_bounds__(kWave) void k_tsc_gather(const double4 *__restrict__ gq,
                                                      const uint64_t *__restrict__ cell,
                                                      const int64_t *__restrict__ off,
                                                      int64_t npix, int periodic,
                                                      double *__restrict__ grid)
{
    __shared__ double acc[2][kTX * kTY * kTZ];
    __shared__ int32_t queue[kTscQueue];
    __shared__ int64_t s_start[32];
    __shared__ int s_base[32];
    int64_t nb[3];
    tsc_bricks(npix, nb);
    const int64_t side[3] = {kTX, kTY, kTZ};
    const int64_t b = blockIdx.x;
    const int64_t bi[3] = {b / (nb[1] * nb[2]), (b / nb[2]) % nb[1], b % nb[2]};
    int64_t b0[3];
    for (int d = 0; d < 3; ++d)
        b0[d] = bi[d] * side[d];
    const int ln = threadIdx.x;
    // neighbour brick of lane ln, if it exists and is not a repeat of an
    // earlier offset on some axis (periodic grids of one or two bricks)
    int len = 0;
    int64_t start = 0;
    if (ln < 27) {
        const int o[3] = {ln / 9 - 1, (ln / 3) % 3 - 1, ln % 3 - 1};
        bool valid = true;
        int64_t x[3];
        for (int d = 0; d < 3; ++d) {
            auto nbr = [&](int oo) {
                int64_t v = bi[d] + oo;
                if (periodic)
                    v = v < 0 ? v + nb[d] : v >= nb[d] ? v - nb[d] : v;
                return v;
            };
            x[d] = nbr(o[d]);
            if (!periodic && (x[d] < 0 || x[d] >= nb[d]))
                valid = false;
            for (int oo = -1; oo < o[d]; ++oo)
                valid = valid && nbr(oo) != x[d];
        }
        if (valid) {
            const int64_t nbrick = (x[0] * nb[1] + x[1]) * nb[2] + x[2];
            start = off[nbrick];
            len = (int)(off[nbrick + 1] - start);
        }
    }
    // exclusive prefix of the range lengths over the lanes
    int incl = len;
#pragma unroll
    for (int sh = 1; sh < kWave; sh <<= 1) {
        const int t = __shfl_up(incl, sh);
        incl += ln >= sh ? t : 0;
    }
    const int total = __shfl(incl, 26);
    const int64_t nd = npix * npix * npix;
    // lane ln writes cells (t, ln / kTZ, ln % kTZ) of the brick, t < kTX
    static_assert(kTY * kTZ == kWave, "one x plane of a brick per wave");
    const int64_t j = b0[1] + ln / kTZ, k = b0[2] + ln % kTZ;
    if (total == 0) {  // wave-uniform: no particle reaches the brick
        for (int64_t i = b0[0]; i < b0[0] + kTX; ++i)
            if (i < npix && j < npix && k < npix) {
                const int64_t idx = (i * npix + j) * npix + k;
                grid[idx] = 0.0;
                grid[nd + idx] = 0.0;
            }
        return;
    }
    for (int e = ln; e < 2 * kTX * kTY * kTZ; e += kWave)
        (&acc[0][0])[e] = 0.0;
    if (ln < 27) {
        s_start[ln] = start;
        s_base[ln] = incl - len;
    }
    __syncthreads();

    // candidate v of the concatenated ranges -> sorted particle index: the
    // last range whose base is <= v (bases are exclusive prefix sums, so an
    // empty range after the one holding v has a base > v)
    auto locate = [&](int v) {
        int lo = 0;
#pragma unroll
        for (int step = 16; step >= 1; step >>= 1)
            if (lo + step < 27 && s_base[lo + step] <= v)
                lo += step;
        return s_start[lo] + (int64_t)(v - s_base[lo]);
    };
    const uint64_t below = ln == 0 ? 0ull : (~0ull >> (kWave - ln));
    int head = 0, tail = 0;  // queue slots consumed / filled (wave-uniform)
    int64_t kc = -1;
    uint64_t cv = 0;
    if (ln < total) {
        kc = locate(ln);
        cv = cell[kc];
    }
    for (int v0 = 0; v0 < total; v0 += kWave) {
        bool use = kc >= 0;
        if (use)
            for (int d = 0; d < 3; ++d)
                use = use && tsc_reach(tsc_unpack(cv, d), b0[d], side[d], npix, periodic) != 0;
        const int64_t kuse = kc;
        // the next chunk's cells
        kc = -1;
        if (v0 + kWave + ln < total) {
            kc = locate(v0 + kWave + ln);
            cv = cell[kc];
        }
        const uint64_t mask = __ballot(use);
        if (use)
            queue[(tail + (int)__popcll(mask & below)) % kTscQueue] = (int32_t)kuse;
        tail += (int)__popcll(mask);
        __builtin_amdgcn_wave_barrier();
        while (tail - head >= kTscPerStep) {
            tsc_step(gq, queue, head, kTscPerStep, ln, b0, npix, periodic, acc);
            head += kTscPerStep;
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (tail > head)
        tsc_step(gq, queue, head, tail - head, ln, b0, npix, periodic, acc);
    __syncthreads();
    for (int t = 0; t < kTX; ++t) {
        const int64_t i = b0[0] + t;
        if (i < npix && j < npix && k < npix) {
            const int64_t idx = (i * npix + j) * npix + k;
            grid[idx] = acc[0][t * kWave + ln];
            grid[nd + idx] = acc[1][t * kWave + ln];
        }
    }
}

int tsc_key_bits(int64_t npix)
{
    const int64_t nbr = tsc_nbricks(npix);
    int bits = 1;
    while (bits < 32 && ((int64_t)1 << bits) <= nbr)
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

void launch_tsc_offsets(const uint32_t *sorted_keys, int64_t n, int64_t npix, int64_t *off,
                        hipStream_t s)
{
    const int64_t nbr = tsc_nbricks(npix);
    hipLaunchKernelGGL(k_tsc_offsets, dim3((unsigned)((nbr + 1 + 255) / 256)), dim3(256), 0, s,
                       sorted_keys, n, nbr, off);
}

void launch_tsc_prep(int dtype, const void *pos, const void *binq, const int32_t *order,
                     const int64_t *off, int64_t n, int64_t npix, double len2pix, int periodic,
                     double4 *gq, uint64_t *cell, hipStream_t s)
{
    if (n == 0)
        return;
    const int64_t nbr = tsc_nbricks(npix);
    const dim3 g((unsigned)((n + 255) / 256)), blk(256);
    if (dtype == S2G_F32)
        hipLaunchKernelGGL(k_tsc_prep<float>, g, blk, 0, s, (const float *)pos,
                           (const float *)binq, order, off, nbr, npix, len2pix, periodic, gq,
                           cell);
    else
        hipLaunchKernelGGL(k_tsc_prep<double>, g, blk, 0, s, (const double *)pos,
                           (const double *)binq, order, off, nbr, npix, len2pix, periodic, gq,
                           cell);
}

void launch_tsc_gather(const double4 *gq, const uint64_t *cell, const int64_t *off,
                       int64_t npix, int periodic, double *grid, hipStream_t s)
{
    const int64_t nbr = tsc_nbricks(npix);
    hipLaunchKernelGGL(k_tsc_gather, dim3((unsigned)nbr), dim3(kWave), 0, s, gq, cell, off,
                       npix, periodic, grid);
}

}  // namespace s2g
