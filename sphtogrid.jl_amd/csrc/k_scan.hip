// k_scan.hip -- deterministic device-wide exclusive scan of int64 values.
// Three phases (block scan, scan of block sums, add), recursive on the block
// sums.  The result is independent of scheduling, so offsets built from it
// (class lists, per-tile pair lists) are bit-reproducible.
#include "s2g_internal.h"

namespace s2g {

constexpr int kScanThreads = 256;
constexpr int kScanPer = 4;
constexpr int kScanBlock = kScanThreads * kScanPer;  // elements per block

__global__ __launch_bounds__(kScanThreads) void k_scan_block(int64_t *d, int64_t n,
                                                             int64_t *block_sums)
{
    __shared__ int64_t sh[kScanThreads];
    const int64_t base = (int64_t)blockIdx.x * kScanBlock + threadIdx.x * kScanPer;
    int64_t v[kScanPer];
    int64_t run = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        int64_t idx = base + k;
        int64_t x = idx < n ? d[idx] : 0;
        v[k] = run;  // exclusive within the thread
        run += x;
    }
    sh[threadIdx.x] = run;
    __syncthreads();
    // Hillis-Steele inclusive scan of the per-thread totals
    for (int off = 1; off < kScanThreads; off <<= 1) {
        int64_t t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    int64_t excl = threadIdx.x ? sh[threadIdx.x - 1] : 0;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        int64_t idx = base + k;
        if (idx < n)
            d[idx] = v[k] + excl;
    }
    if (threadIdx.x == kScanThreads - 1)
        block_sums[blockIdx.x] = sh[kScanThreads - 1];
}

__global__ __launch_bounds__(kScanThreads) void k_scan_add(int64_t *d, int64_t n,
                                                           const int64_t *block_off)
{
    const int64_t base = (int64_t)blockIdx.x * kScanBlock;
    const int64_t add = block_off[blockIdx.x];
    for (int k = threadIdx.x; k < kScanBlock; k += kScanThreads) {
        int64_t idx = base + k;
        if (idx < n)
            d[idx] += add;
    }
}

size_t scan_tmp_elems(int64_t n)
{
    size_t total = 0;
    while (n > 1) {
        int64_t nb = (n + kScanBlock - 1) / kScanBlock;
        total += (size_t)nb;
        n = nb;
    }
    return total + 1;
}

void exclusive_scan_i64(int64_t *d, int64_t n, int64_t *tmp, hipStream_t s)
{
    if (n <= 0)
        return;
    int64_t nb = (n + kScanBlock - 1) / kScanBlock;
    hipLaunchKernelGGL(k_scan_block, dim3((unsigned)nb), dim3(kScanThreads), 0, s, d, n,
                       tmp);
    if (nb > 1) {
        exclusive_scan_i64(tmp, nb, tmp + nb, s);
        hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(kScanThreads), 0, s, d,
                           n, tmp);
    }
}

}  // namespace s2g
