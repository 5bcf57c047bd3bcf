// Stable device radix sort of (uint64 key, int32 value) pairs.
#pragma once
#include <hip/hip_runtime.h>

namespace dem {
// int32 scratch needed by radix_sort_u64 for n keys.
size_t radix_sort_workspace_ints(int n);
// Stable ascending sort by key bits [begin_bit, end_bit) (8-bit digits from
// begin_bit up; bits beyond 63 read as zero); the values end in vals.  The
// keys end in keys, or, when sorted_keys is not null, in whichever of keys /
// keys_alt holds them (*sorted_keys), which saves a copy after an odd number
// of passes.  keys_alt/vals_alt are same-size scratch buffers.
int radix_sort_u64(unsigned long long* keys, int* vals, unsigned long long* keys_alt,
                   int* vals_alt, int n, int begin_bit, int end_bit, int* ws, hipStream_t st,
                   unsigned long long** sorted_keys = nullptr);
}  // namespace dem
