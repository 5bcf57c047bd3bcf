// Device-wide exclusive scan used by the hash and octree passes.
#pragma once
#include <hip/hip_runtime.h>

namespace dem {
// int32 scratch needed by exclusive_scan for n elements.
size_t scan_workspace_ints(int n);
// out[i] = sum(in[0..i)); optional *total_out = sum(in[0..n)) (device pointer).
// `in` and `out` may alias.
int exclusive_scan(const int* in, int* out, int n, int* ws, int* total_out, hipStream_t st);
}  // namespace dem
