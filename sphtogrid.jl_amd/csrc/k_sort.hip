// k_sort.hip -- key/value radix sort used to bucket small-footprint particles
// by tile (rocPRIM; its own translation unit because of its compile time).
#include <rocprim/device/device_radix_sort.hpp>

#include "k_sort.h"

namespace s2g {

size_t sort_pairs_tmp_bytes(int64_t n, int end_bit)
{
    size_t bytes = 0;
    const hipError_t e = rocprim::radix_sort_pairs(
        nullptr, bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr, (const int32_t *)nullptr,
        (int32_t *)nullptr, (unsigned int)n, 0u, (unsigned int)end_bit, (hipStream_t)0);
    if (e != hipSuccess)
        throw Error(S2G_EHIP, std::string("radix sort size query: ") + hipGetErrorString(e));
    return bytes;
}

void sort_pairs(void *tmp, size_t bytes, const uint32_t *kin, uint32_t *kout, const int32_t *vin,
                int32_t *vout, int64_t n, int end_bit, hipStream_t s)
{
    const hipError_t e = rocprim::radix_sort_pairs(tmp, bytes, kin, kout, vin, vout,
                                                   (unsigned int)n, 0u, (unsigned int)end_bit, s);
    if (e != hipSuccess)
        throw Error(S2G_EHIP, std::string("radix sort: ") + hipGetErrorString(e));
}

size_t sort_pairs64_tmp_bytes(int64_t n, int end_bit)
{
    size_t bytes = 0;
    const hipError_t e = rocprim::radix_sort_pairs(
        nullptr, bytes, (const uint64_t *)nullptr, (uint64_t *)nullptr, (const int32_t *)nullptr,
        (int32_t *)nullptr, (unsigned int)n, 0u, (unsigned int)end_bit, (hipStream_t)0);
    if (e != hipSuccess)
        throw Error(S2G_EHIP, std::string("radix sort size query: ") + hipGetErrorString(e));
    return bytes;
}

void sort_pairs64(void *tmp, size_t bytes, const uint64_t *kin, uint64_t *kout,
                  const int32_t *vin, int32_t *vout, int64_t n, int end_bit, hipStream_t s)
{
    const hipError_t e = rocprim::radix_sort_pairs(tmp, bytes, kin, kout, vin, vout,
                                                   (unsigned int)n, 0u, (unsigned int)end_bit, s);
    if (e != hipSuccess)
        throw Error(S2G_EHIP, std::string("radix sort: ") + hipGetErrorString(e));
}

}  // namespace s2g
