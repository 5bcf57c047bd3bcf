// Stable device radix sort of (uint64 key, int32 value) pairs.
#pragma once
#include <hip/hip_runtime.h>

namespace dem {
// int32 scratch needed by radix_sort_u64 for n keys.
size_t radix_sort_workspace_ints(int n);
// Sorts the low `bits` bits of keys (ascending, stable); the result ends in
// keys/vals.  keys_alt/vals_alt are same-size scratch buffers.
int radix_sort_u64(unsigned long long* keys, int* vals, unsigned long long* keys_alt,
                   int* vals_alt, int n, int bits, int* ws, hipStream_t st);
}  // namespace dem
