// k_sort.h -- stable LSD radix sort of (uint32 key, int32 value) pairs.
#pragma once

#include "s2g_internal.h"

namespace s2g {

// temporary storage for sort_pairs of n pairs with keys below 2^end_bit
size_t sort_pairs_tmp_bytes(int64_t n, int end_bit);
void sort_pairs(void *tmp, size_t bytes, const uint32_t *kin, uint32_t *kout, const int32_t *vin,
                int32_t *vout, int64_t n, int end_bit, hipStream_t s);

// same with uint64 keys
size_t sort_pairs64_tmp_bytes(int64_t n, int end_bit);
void sort_pairs64(void *tmp, size_t bytes, const uint64_t *kin, uint64_t *kout,
                  const int32_t *vin, int32_t *vout, int64_t n, int end_bit, hipStream_t s);

}  // namespace s2g
