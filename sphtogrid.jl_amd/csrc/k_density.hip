// k_density.hip -- mass_density (effects/mass_density.jl:30-99) on gfx950:
// SPH density and smoothing length of an arbitrary particle set from its
// k = Nneighbors + 1 nearest neighbours (the particle itself included).
//
// Neighbour search without a tree build: particles are sorted by a 63-bit
// Morton key (21 bits per axis over the domain), so the cells of every level
// of the implicit octree are contiguous ranges of the sorted order.  A table
// of cell offsets at one level (8 at most, 8^8 cells) turns a cell into its range
// with two loads; finer cells are found by binary search inside their table
// cell.
//
//   k_md_prep      FP64 (x, y, z, m), positions wrapped into a periodic box,
//                  per-block bounding boxes
//   k_md_keys      Morton key per particle -> radix sort
//   k_md_sorted    particles in key order
//   k_md_table     cell offsets at the table level (binary search per cell)
//   k_md_knn       one wave per particle:
//                  1. a first bound T0 >= the true k-th distance: the k-th
//                     smallest distance to the particles next to it in key
//                     order (they are k real particles);
//                  2. the level whose cells are between T0/2 and T0 wide and
//                     the cube of <= 5^3 cells around the particle that covers
//                     the ball of radius T0;
//                  3. every particle in the cube at distance <= T goes to a
//                     per-wave LDS buffer (T shrinks as the buffer is cut back
//                     to the k nearest), the cells within T0/2 first: the
//                     k-th nearest of those gives T1, and of the other cells
//                     only those within T1 are read; the search works on
//                     squared distances, and the exact k-th smallest is a
//                     radix select on their FP64 bit patterns (its square
//                     root is hsml);
//                  4. the kernel sum over the buffer and the bias correction
//                     of Dehnen & Aly (2012, eqs. 18-19).
// The cube holds the whole ball of radius T0 and T0 is an upper bound, so the
// k nearest neighbours are all found: the result is exact, and the work per
// particle follows the local density (clustered sets cost what uniform ones
// cost).  Distances are FP64 without contraction, minimum image per
// dimension in a periodic box.
#include "k_density.h"

namespace s2g {

constexpr int kMdWaves = 4;     // waves per block of the search
constexpr int kMdCap = 600;     // candidate buffer per wave
constexpr int kMdBits = 21;     // Morton bits per axis
constexpr int kMdMaxCells = 216;  // cube cells per particle (at most 6^3)

template <typename T>
__global__ __launch_bounds__(256) void k_md_prep(const T *__restrict__ pos,
                                                 const T *__restrict__ mass, int64_t n,
                                                 MdGrid g, double4 *__restrict__ P,
                                                 double *__restrict__ bbox)
{
    __shared__ double s_lo[3][256], s_hi[3][256];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    if (i < n) {
        double x[3];
        for (int d = 0; d < 3; ++d) {
            double v = (double)pos[3 * i + d];
            if (g.periodic) {
                v = v - g.box[d] * floor(v / g.box[d]);
                if (v >= g.box[d])
                    v -= g.box[d];
            }
            x[d] = v;
            lo[d] = hi[d] = v;
            if (!isfinite(v)) {  // reported by the host as an invalid input
                lo[d] = -INFINITY;
                hi[d] = INFINITY;
            }
        }
        P[i] = make_double4(x[0], x[1], x[2], (double)mass[i]);
    }
    for (int d = 0; d < 3; ++d) {
        s_lo[d][threadIdx.x] = lo[d];
        s_hi[d][threadIdx.x] = hi[d];
    }
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int d = 0; d < 3; ++d) {
                s_lo[d][threadIdx.x] = fmin(s_lo[d][threadIdx.x], s_lo[d][threadIdx.x + s]);
                s_hi[d][threadIdx.x] = fmax(s_hi[d][threadIdx.x], s_hi[d][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 3) {
        bbox[6 * blockIdx.x + threadIdx.x] = s_lo[threadIdx.x][0];
        bbox[6 * blockIdx.x + 3 + threadIdx.x] = s_hi[threadIdx.x][0];
    }
}

// spread the low 21 bits of a over every third bit
__host__ __device__ __forceinline__ uint64_t md_split3(uint32_t a)
{
    uint64_t x = a & 0x1fffffu;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

__device__ __forceinline__ uint64_t md_morton(uint32_t x, uint32_t y, uint32_t z)
{
    return md_split3(x) | md_split3(y) << 1 | md_split3(z) << 2;
}

__device__ __forceinline__ uint32_t md_quant(double v, const MdGrid &g, int d)
{
    const double t = (v - g.lo[d]) * g.inv_ext[d] * (double)(1u << kMdBits);
    const double c = floor(t);
    return c < 0.0 ? 0u : (c >= (double)((1u << kMdBits) - 1) ? (1u << kMdBits) - 1 : (uint32_t)c);
}

__global__ void k_md_keys(const double4 *__restrict__ P, int64_t n, MdGrid g,
                          uint64_t *__restrict__ keys, int32_t *__restrict__ vals)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const double4 p = P[i];
    keys[i] = md_morton(md_quant(p.x, g, 0), md_quant(p.y, g, 1), md_quant(p.z, g, 2));
    vals[i] = (int32_t)i;
}

__global__ void k_md_sorted(const double4 *__restrict__ P, const int32_t *__restrict__ idx,
                            int64_t n, double4 *__restrict__ S)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        S[i] = P[idx[i]];
}

__device__ __forceinline__ int64_t md_lower_bound(const uint64_t *__restrict__ K, int64_t lo,
                                                  int64_t hi, uint64_t key)
{
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (K[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// offsets of the 8^lt cells of level lt (+ the end)
__global__ void k_md_table(const uint64_t *__restrict__ K, int64_t n, int lt,
                           int32_t *__restrict__ tab)
{
    const int64_t ncell = (int64_t)1 << (3 * lt);
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e > ncell)
        return;
    tab[e] = e == ncell ? (int32_t)n
                        : (int32_t)md_lower_bound(K, 0, n, (uint64_t)e << (3 * (kMdBits - lt)));
}

// minimum-image component (mass_density.jl:45-52, Distances.jl PeriodicEuclidean)
__device__ __forceinline__ double md_comp(double a, double b, double box, int periodic)
{
    double x = fabs(a - b);
    if (periodic && x > 0.5 * box)
        x = box - x;
    return x;
}

// squared distance, rounded as the reference rounds it before the square
// root (no contraction).  The search selects on these: the k-th smallest
// squared distance is the square of the k-th smallest distance, because a
// correctly rounded sqrt is monotone, so sqrt of it is the reference's hsml.
__device__ __forceinline__ double md_dist2(const double4 &a, const double4 &b, const MdGrid &g)
{
    const double dx = md_comp(a.x, b.x, g.box[0], g.periodic);
    const double dy = md_comp(a.y, b.y, g.box[1], g.periodic);
    const double dz = md_comp(a.z, b.z, g.box[2], g.periodic);
    return add_rn(add_rn(mul_rn(dx, dx), mul_rn(dy, dy)), mul_rn(dz, dz));
}

// exclusive prefix sum over the lanes of a wave
__device__ __forceinline__ int wave_excl_scan(int v, int ln)
{
    int x = v;
    for (int o = 1; o < kWave; o <<= 1) {
        const int y = __shfl_up(x, o, kWave);
        if (ln >= o)
            x += y;
    }
    return x - v;
}

// value of another lane of the same 16-lane row (DPP lane pattern)
template <int kCtrl>
__device__ __forceinline__ double dpp_get(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), kCtrl, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), kCtrl, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}

// minimum and maximum over the 64 lanes of a fully active wave, on every lane
__device__ __forceinline__ void wave_min_max(double &mn, double &mx)
{
    mn = fmin(mn, dpp_get<0xB1>(mn));  // quad_perm [1, 0, 3, 2]
    mx = fmax(mx, dpp_get<0xB1>(mx));
    mn = fmin(mn, dpp_get<0x4E>(mn));  // quad_perm [2, 3, 0, 1]
    mx = fmax(mx, dpp_get<0x4E>(mx));
    mn = fmin(mn, dpp_get<0x141>(mn));  // row_half_mirror
    mx = fmax(mx, dpp_get<0x141>(mx));
    mn = fmin(mn, dpp_get<0x140>(mn));  // row_mirror
    mx = fmax(mx, dpp_get<0x140>(mx));
    double a = mn, b = mx;
#pragma unroll
    for (int r = 16; r < kWave; r += 16) {
        a = fmin(a, __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(mn), r),
                                     __builtin_amdgcn_readlane(__double2loint(mn), r)));
        b = fmax(b, __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(mx), r),
                                     __builtin_amdgcn_readlane(__double2loint(mx), r)));
    }
    mn = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(a), 0),
                          __builtin_amdgcn_readlane(__double2loint(a), 0));
    mx = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(b), 0),
                          __builtin_amdgcn_readlane(__double2loint(b), 0));
}

// k-th smallest (1-based) of the doubles buf[0..cnt) (each +0 or positive),
// exactly: most-significant-bits-first radix select on the bit patterns.
// Zeros (the particle itself, coincident particles) are counted apart; the
// bits above the highest one in which the smallest and largest positive
// entries differ are common to all of them, so the first 8-bit pass starts
// there and spreads the entries over many bins (a pass from the top byte
// sends them all to one bin, and its LDS atomics serialise).  Once the chosen
// bin holds a single entry, that entry is the k-th: one scan finds it.
// kExact == false: an upper bound of the k-th smallest instead, for
// thresholds that only have to keep at least k entries: the upper edge of
// the bin that holds it once the bins are narrower than 1/16 of a binade
// (the exponent and four mantissa bits chosen; entries that straddle a power
// of two where many exponent bits flip can need a second pass for that), or
// the largest entry if smaller.
template <bool kExact>
__device__ double md_select(const double *buf, int cnt, int k, int *hist, int ln)
{
    double mn = __builtin_inf(), mx = 0.0;
    int nz = 0;  // uniform
    for (int e0 = 0; e0 < cnt; e0 += kWave) {
        const int e = e0 + ln;
        const double v = e < cnt ? buf[e] : -1.0;
        nz += (int)__popcll(__ballot(v == 0.0));
        if (v > 0.0) {
            mn = fmin(mn, v);
            mx = fmax(mx, v);
        }
    }
    if (k <= nz)
        return 0.0;
    k -= nz;
    wave_min_max(mn, mx);
    const uint64_t lo = (uint64_t)__double_as_longlong(mn);
    const uint64_t hi = (uint64_t)__double_as_longlong(mx);
    if (lo == hi)
        return mn;
    int top = 64 - __clzll((long long)(lo ^ hi));  // bits from `top` up are common (top < 64)
    uint64_t prefix = lo >> top << top;
    for (;;) {
        const int w = top < 8 ? top : 8;
        const int shift = top - w;
        const uint64_t mask = (1ull << w) - 1;
        for (int b = ln; b < 256; b += kWave)
            hist[b] = 0;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        for (int e = ln; e < cnt; e += kWave) {
            const uint64_t key = (uint64_t)__double_as_longlong(buf[e]);
            if (key != 0 && (key >> top) == (prefix >> top))
                atomicAdd(&hist[(key >> shift) & mask], 1);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        // lane l owns bins 4l .. 4l+3
        int c4[4], loc = 0;
        for (int j = 0; j < 4; ++j) {
            c4[j] = hist[4 * ln + j];
            loc += c4[j];
        }
        const int before = wave_excl_scan(loc, ln);
        const bool mine = before < k && k <= before + loc;
        const uint64_t owner = __ballot(mine);
        const int src = __ffsll((unsigned long long)owner) - 1;
        int bin = 0, below = 0, nbin = 0;
        if (mine) {
            int acc = before;
            for (int j = 0; j < 4; ++j) {
                if (k <= acc + c4[j]) {
                    bin = 4 * ln + j;
                    below = acc;
                    nbin = c4[j];
                    break;
                }
                acc += c4[j];
            }
        }
        bin = __shfl(bin, src, kWave);
        below = __shfl(below, src, kWave);
        nbin = __shfl(nbin, src, kWave);
        k -= below;
        prefix |= (uint64_t)bin << shift;
        top = shift;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        if (!kExact && top <= 48)
            return fmin(__longlong_as_double((long long)(prefix | ((1ull << top) - 1))), mx);
        if (top == 0)
            return __longlong_as_double((long long)prefix);
        if (nbin == 1) {
            // the one positive entry whose bits from `top` up equal the prefix's
            for (int e0 = 0; e0 < cnt; e0 += kWave) {
                const int e = e0 + ln;
                const uint64_t key = e < cnt ? (uint64_t)__double_as_longlong(buf[e]) : 0;
                const uint64_t bh = __ballot(key != 0 && (key >> top) == (prefix >> top));
                if (bh)  // wave-uniform
                    return __shfl(__longlong_as_double((long long)key),
                                  __ffsll((unsigned long long)bh) - 1, kWave);
            }
        }
    }
}

// keep the entries below T and, of those equal to T, as many as make k in
// total (they are at u = 1, W = 0); returns the new count
__device__ int md_compact(double *buf, int *jb, int cnt, double T, int k, int ln)
{
    int nless = 0;
    for (int e0 = 0; e0 < cnt; e0 += kWave) {
        const int e = e0 + ln;
        nless += __popcll(__ballot(e < cnt && buf[e] < T));
    }
    int quota = k - nless;  // entries equal to T to keep
    int out = 0;
    const uint64_t below = ln == 0 ? 0ull : (~0ull >> (kWave - ln));
    for (int e0 = 0; e0 < cnt; e0 += kWave) {
        const int e = e0 + ln;
        const double d = e < cnt ? buf[e] : 0.0;
        const int j = e < cnt ? jb[e] : 0;
        const bool less = e < cnt && d < T;
        const bool eq = e < cnt && d == T;
        const uint64_t beq = __ballot(eq);
        const bool keep_eq = eq && (int)__popcll(beq & below) < quota;
        const bool keep = less || keep_eq;
        const uint64_t bk = __ballot(keep);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        if (keep) {
            const int at = out + (int)__popcll(bk & below);
            buf[at] = d;
            jb[at] = j;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        out += (int)__popcll(bk);
        const int taken = (int)__popcll(beq);
        quota -= taken < quota ? taken : (quota > 0 ? quota : 0);
    }
    return out;
}

// Append the candidates (lane's j, squared distance d, valid flag) with
// d <= T to the buffer; when it is about to overflow, cut it back to the k
// nearest.
__device__ __forceinline__ void md_offer(bool valid, double d, int j, double &T, int &cnt,
                                         double *buf, int *jb, int *hist, int k1, int ln,
                                         uint64_t below)
{
    const bool take = valid && d <= T;
    const uint64_t bt = __ballot(take);
    if (take) {
        const int at = cnt + (int)__popcll(bt & below);
        buf[at] = d;
        jb[at] = j;
    }
    cnt += (int)__popcll(bt);
    if (cnt > kMdCap - kWave) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        T = md_select<true>(buf, cnt, k1, hist, ln);
        cnt = md_compact(buf, jb, cnt, T, k1, ln);
    }
}

template <int K>
__global__ __launch_bounds__(kMdWaves * kWave, 4) void k_md_knn(
    const double4 *__restrict__ S, const uint64_t *__restrict__ SK,
    const int32_t *__restrict__ order, const int32_t *__restrict__ tab, int lt, int64_t n,
    MdGrid g, int k1, int dim, int nneigh, double norm, double *__restrict__ rho_out,
    double *__restrict__ hsml_out)
{
    __shared__ double s_d[kMdWaves][kMdCap];
    __shared__ int s_j[kMdWaves][kMdCap];
    __shared__ int s_h[kMdWaves][256];
    __shared__ int s_cs[kMdWaves][kMdMaxCells];  // cube cell: first particle
    __shared__ int s_cp[kMdWaves][kMdMaxCells + 1];  // prefix of the cells' counts
    const int wv = threadIdx.x / kWave, ln = threadIdx.x % kWave;
    double *buf = s_d[wv];
    int *jb = s_j[wv];
    int *hist = s_h[wv];
    int *cstart = s_cs[wv];
    int *cpre = s_cp[wv];
    const uint64_t below = ln == 0 ? 0ull : (~0ull >> (kWave - ln));
    const int64_t nw = (int64_t)gridDim.x * kMdWaves;
    for (int64_t q = (int64_t)blockIdx.x * kMdWaves + wv; q < n; q += nw) {
        const double4 me = S[q];
        // 1. first bound: the k-th nearest of the particles next in key
        //    order, k on each side (where the particles on one side are far,
        //    as at the edge of a cluster in key order, the other side alone
        //    still holds k)
        int cnt = 0;
        double T = 1e300;  // squared distances from here on
        {
            const int64_t a = q - k1 > 0 ? q - k1 : 0;
            const int64_t b = q + k1 < n - 1 ? q + k1 : n - 1;
            for (int64_t t0 = a; t0 <= b; t0 += kWave) {
                const int64_t t = t0 + ln;
                const bool valid = t <= b;
                const double d = valid ? md_dist2(me, S[t], g) : 0.0;
                md_offer(valid, d, (int)t, T, cnt, buf, jb, hist, k1, ln, below);
            }
            T = md_select<false>(buf, cnt, k1, hist, ln);  // >= the k-th of these
        }
        const double T20 = T, T0 = sqrt(T20);
        // 2. level with cells in [T0/2, T0) along the narrowest axis; per
        //    axis the cells that meet [x - T0, x + T0], found with the
        //    particles' own quantisation widened by one quantum: the span is
        //    2 T0 + 2 quanta < 5 cells (cells are >= T0 / 2 wide and a
        //    quantum is at most half a cell below level 21, where the cell
        //    exceeds T0), so at most six cells per axis
        int lev = kMdBits;
        if (T0 > 0.0) {
            const int l = (int)floor(log2(g.min_ext / T0)) + 1;
            lev = l < 0 ? 0 : (l > kMdBits ? kMdBits : l);
        }
        const uint32_t ncl = 1u << lev;
        const int sh = kMdBits - lev;
        int ca[3], cl[3];  // first cell (unwrapped) and cell count per axis
        const double mv[3] = {me.x, me.y, me.z};
        for (int d = 0; d < 3; ++d) {
            const double qs = (double)(1u << kMdBits);
            const int64_t qlo = (int64_t)floor((mv[d] - T0 - g.lo[d]) * g.inv_ext[d] * qs) - 1;
            const int64_t qhi = (int64_t)floor((mv[d] + T0 - g.lo[d]) * g.inv_ext[d] * qs) + 1;
            int64_t clo = qlo >> sh, chi = qhi >> sh;  // floor division
            if (g.periodic) {
                if (chi - clo + 1 >= (int64_t)ncl) {
                    clo = 0;
                    chi = (int64_t)ncl - 1;
                }
            } else {
                clo = clo < 0 ? 0 : clo;
                chi = chi > (int64_t)ncl - 1 ? (int64_t)ncl - 1 : chi;
            }
            ca[d] = (int)clo;
            cl[d] = (int)(chi - clo + 1);
        }
        // cells whose nearest point is farther than T0 (with a margin of two
        // quanta per axis) hold no candidate
        const double T0m = T0 * (1.0 + 1e-12);
        const int ncube = cl[0] * cl[1] * cl[2];
        // 3. the cube's particles; cells are ranges of the sorted order.  The
        //    cells whose nearest point lies within T0/2 go first: the k-th
        //    nearest of their particles is a bound T1 close to the true k-th
        //    distance (T0 is often half as large again), and of the other
        //    cells only those within T1 are read
        // per axis: unwrapped cell o spans [o, o + 1) * cw around x; the
        // distance to it is taken eps = two quanta short
        double cw[3], xlo[3], xhi[3];
        float rcl[3];
        for (int d = 0; d < 3; ++d) {
            const double eps = 2.0 * g.ext[d] / (double)(1u << kMdBits);
            const double xr = mv[d] - g.lo[d];
            cw[d] = g.ext[d] / (double)ncl;
            xlo[d] = xr + eps;
            xhi[d] = xr - eps;
            rcl[d] = 1.0f / (float)cl[d];
        }
        int ncells = 0;  // uniform
        auto gather_cells = [&](double lo2, double hi2) -> int {
            // the non-empty cells with lo2 < (nearest distance)^2 <= hi2, in
            // cube order, into cstart/cpre (first particle, number of the
            // first candidate) and ncells; returns their particle count
            // (uniform)
            int C = 0, nc = 0;
            for (int t0 = 0; t0 < ncube; t0 += kWave) {
                const int t = t0 + ln;
                int s = 0, e = 0;
                bool near = false;
                uint32_t cc[3];
                if (t < ncube) {
                    int r = t;
                    double dd = 0.0;
                    for (int d = 0; d < 3; ++d) {
                        // r = q cl + rem with r < 216, cl <= 6: the float
                        // quotient of r + 1/2 is at least 1/12 from an integer
                        const int qd = d < 2 ? (int)(((float)r + 0.5f) * rcl[d]) : 0;
                        int o = ca[d] + (r - qd * cl[d]);
                        r = qd;
                        if (cl[d] < (int)ncl || !g.periodic) {
                            const double a = (double)o * cw[d] - xlo[d];
                            const double b = xhi[d] - (double)(o + 1) * cw[d];
                            const double dx = a > 0.0 ? a : (b > 0.0 ? b : 0.0);
                            dd += dx * dx;
                        }
                        o = o < 0 ? o + (int)ncl : (o >= (int)ncl ? o - (int)ncl : o);
                        cc[d] = (uint32_t)o;
                    }
                    near = dd > lo2 && dd <= hi2;
                }
                if (near) {
                    const uint64_t pre = md_morton(cc[0], cc[1], cc[2]);  // 3*lev bits
                    if (lev <= lt) {
                        s = tab[pre << (3 * (lt - lev))];
                        e = tab[(pre + 1) << (3 * (lt - lev))];
                    } else {
                        const uint64_t par = pre >> (3 * (lev - lt));
                        const int64_t ps = tab[par], pe = tab[par + 1];
                        const int ksh = 3 * (kMdBits - lev);
                        s = (int)md_lower_bound(SK, ps, pe, pre << ksh);
                        e = (int)md_lower_bound(SK, s, pe, (pre + 1) << ksh);
                    }
                }
                const int len = e - s;
                const int ex = wave_excl_scan(len, ln);
                const uint64_t bne = __ballot(len > 0);
                if (len > 0) {
                    const int at = nc + (int)__popcll(bne & below);
                    cstart[at] = s;
                    cpre[at] = C + ex;
                }
                nc += (int)__popcll(bne);
                C += __shfl(ex + len, kWave - 1, kWave);
            }
            ncells = nc;
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            return C;
        };
        auto offer_cells = [&](int C) {
            // candidate u lies in the last cell c with cpre[c] <= u.  Of a
            // chunk of 64 candidates from u0, the cells after cp (the cell
            // of u0 - 1) that start in it are the next ones in the list:
            // their starts are flagged (in the histogram's first 64 ints,
            // free here), and a lane's cell is cp plus the flags at or
            // below its position
            int cp = -1;  // uniform
            for (int u0 = 0; u0 < C; u0 += kWave) {
                hist[ln] = 0;
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                const int ci = cp + 1 + ln;
                if (ci < ncells) {
                    const int st = cpre[ci] - u0;  // >= 0
                    if (st < kWave)
                        hist[st] = 1;
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                const uint64_t bf = __ballot(hist[ln] != 0);
                const int u = u0 + ln;
                const bool valid = u < C;
                double d = 0.0;
                int j = 0;
                if (valid) {
                    const int c = cp + (int)__popcll(bf & (below | (1ull << ln)));
                    j = cstart[c] + (u - cpre[c]);
                    d = md_dist2(me, S[j], g);
                }
                cp += (int)__popcll(bf);
                md_offer(valid, d, j, T, cnt, buf, jb, hist, k1, ln, below);
            }
            // the next pass rewrites cstart/cpre
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        };
        cnt = 0;  // fresh buffer, threshold T0 (inclusive)
        T = T20;
        const double inner2 = 0.25 * T0m * T0m;
        for (int pass = 0; pass < 2; ++pass) {
            double lo2 = -1.0, hi2 = inner2;
            if (pass == 1) {
                const double T1m = sqrt(T) * (1.0 + 1e-12);
                lo2 = inner2;
                hi2 = T1m * T1m;
            }
            offer_cells(gather_cells(lo2, hi2));
            // after the inner cells a bound T1 <= T0 of the k-th (at least
            // k entries are within it); at the end the exact k-th, with
            // every particle nearer than it in the buffer, each once
            if (cnt >= k1)
                T = pass == 0 ? md_select<false>(buf, cnt, k1, hist, ln)
                              : md_select<true>(buf, cnt, k1, hist, ln);
        }
        // 4. mass_density.jl:70-86: kernel sum over the neighbours, bias correction
        const double h = sqrt(T), h_inv = 1.0 / h;
        double rho = 0.0;
        for (int e = ln; e < cnt; e += kWave) {
            const double d2 = buf[e];
            if (d2 <= T) {
                const double d = sqrt(d2);
                if (d < h)
                    rho += S[jb[e]].w * shape_of<K, double>(d * h_inv);
            }
        }
        rho = wave_sum(rho);
        double hd = h_inv;
        for (int d = 1; d < dim; ++d)
            hd *= h_inv;
        rho *= norm * hd;
        rho = md_bias_correction(K, rho, me.w, norm * hd, nneigh);
        if (ln == 0) {
            const int oi = order[q];
            rho_out[oi] = rho;
            hsml_out[oi] = h;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
int64_t md_prep_blocks(int64_t n) { return (n + 255) / 256; }

int md_max_neighbours() { return kMdCap - 2 * kWave - 1; }

int md_bits() { return kMdBits; }

void launch_md_prep(int dtype, const void *pos, const void *mass, int64_t n, const MdGrid &g,
                    double4 *P, double *bbox, hipStream_t s)
{
    const dim3 grid((unsigned)md_prep_blocks(n)), blk(256);
    if (dtype == S2G_F32)
        hipLaunchKernelGGL(k_md_prep<float>, grid, blk, 0, s, (const float *)pos,
                           (const float *)mass, n, g, P, bbox);
    else
        hipLaunchKernelGGL(k_md_prep<double>, grid, blk, 0, s, (const double *)pos,
                           (const double *)mass, n, g, P, bbox);
}

void launch_md_keys(const double4 *P, int64_t n, const MdGrid &g, uint64_t *keys, int32_t *vals,
                    hipStream_t s)
{
    hipLaunchKernelGGL(k_md_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, P, n, g,
                       keys, vals);
}

void launch_md_sorted(const double4 *P, const int32_t *idx, int64_t n, double4 *S,
                      hipStream_t s)
{
    hipLaunchKernelGGL(k_md_sorted, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, P, idx,
                       n, S);
}

void launch_md_table(const uint64_t *K, int64_t n, int lt, int32_t *tab, hipStream_t s)
{
    const int64_t ne = ((int64_t)1 << (3 * lt)) + 1;
    hipLaunchKernelGGL(k_md_table, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s, K, n, lt,
                       tab);
}

void launch_md_knn(int kernel, const double4 *S, const uint64_t *SK, const int32_t *order,
                   const int32_t *tab, int lt, int64_t n, const MdGrid &g, int k1, int dim,
                   int nneigh, double norm, double *rho, double *hsml, hipStream_t s)
{
    int64_t blocks = (n + kMdWaves - 1) / kMdWaves;
    if (blocks > (1 << 18))
        blocks = 1 << 18;
    const dim3 grid((unsigned)blocks), blk(kMdWaves * kWave);
#define S2G_MD_CASE(KK)                                                                      \
    hipLaunchKernelGGL(k_md_knn<KK>, grid, blk, 0, s, S, SK, order, tab, lt, n, g, k1, dim, \
                       nneigh, norm, rho, hsml)
    switch (kernel) {
    case S2G_KERNEL_CUBIC: S2G_MD_CASE(S2G_KERNEL_CUBIC); break;
    case S2G_KERNEL_QUINTIC: S2G_MD_CASE(S2G_KERNEL_QUINTIC); break;
    case S2G_KERNEL_WENDLAND_C2: S2G_MD_CASE(S2G_KERNEL_WENDLAND_C2); break;
    case S2G_KERNEL_WENDLAND_C4: S2G_MD_CASE(S2G_KERNEL_WENDLAND_C4); break;
    case S2G_KERNEL_WENDLAND_C6: S2G_MD_CASE(S2G_KERNEL_WENDLAND_C6); break;
    default: S2G_MD_CASE(S2G_KERNEL_WENDLAND_C8); break;
    }
#undef S2G_MD_CASE
}

}  // namespace s2g
