// Device-wide exclusive prefix sum of int32 (three launches, no host sync).
// Replaces the reference's serial scans (spatial_hash_collision_detector.py:114-123,
// flat_tree.py:370-385).
#include "dem_common.h"
#include "scan.h"

namespace dem {

namespace {
constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr int kScanBlock = kScanThreads * kScanItems;  // 1024 elements per block

__device__ __forceinline__ int block_exclusive(int v, int* sh, int& total) {
  // Hillis-Steele over 256 threads in LDS.
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    int add = t >= off ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += add;
    __syncthreads();
  }
  total = sh[kScanThreads - 1];
  int excl = sh[t] - v;
  __syncthreads();
  return excl;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_blocks(const int* in, int* out, int n,
                                                              int* __restrict__ block_sums) {
  __shared__ int sh[kScanThreads];
  const int base = blockIdx.x * kScanBlock + threadIdx.x * kScanItems;
  int v[kScanItems];
  int local = 0;
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0;
    local += v[k];
  }
  int total;
  int excl = block_exclusive(local, sh, total);
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) out[base + k] = excl;
    excl += v[k];
  }
  if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// Single workgroup: exclusive scan of the block sums in chunks of 1024.
__global__ __launch_bounds__(kScanThreads) void k_scan_sums(int* __restrict__ sums, int nb,
                                                            int* __restrict__ total_out) {
  __shared__ int sh[kScanThreads];
  int carry = 0;
  for (int base = 0; base < nb; base += kScanBlock) {
    const int b = base + threadIdx.x * kScanItems;
    int v[kScanItems];
    int local = 0;
    for (int k = 0; k < kScanItems; ++k) {
      v[k] = (b + k < nb) ? sums[b + k] : 0;
      local += v[k];
    }
    int total;
    int excl = block_exclusive(local, sh, total) + carry;
    for (int k = 0; k < kScanItems; ++k) {
      if (b + k < nb) sums[b + k] = excl;
      excl += v[k];
    }
    carry += total;
  }
  if (threadIdx.x == 0 && total_out) *total_out = carry;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_add(int* __restrict__ out, int n,
                                                           const int* __restrict__ sums) {
  const int base = blockIdx.x * kScanBlock + threadIdx.x * kScanItems;
  const int add = sums[blockIdx.x];
  for (int k = 0; k < kScanItems; ++k)
    if (base + k < n) out[base + k] += add;
}
}  // namespace

size_t scan_workspace_ints(int n) { return (size_t)((n + kScanBlock - 1) / kScanBlock) + 1; }

int exclusive_scan(const int* in, int* out, int n, int* ws, int* total_out, hipStream_t st) {
  if (n <= 0) return DEM_OK;
  const int nb = (n + kScanBlock - 1) / kScanBlock;
  hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(kScanThreads), 0, st, in, out, n, ws);
  DEM_CHECK_LAUNCH("k_scan_blocks");
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanThreads), 0, st, ws, nb, total_out);
  DEM_CHECK_LAUNCH("k_scan_sums");
  hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(kScanThreads), 0, st, out, n, (const int*)ws);
  DEM_CHECK_LAUNCH("k_scan_add");
  return DEM_OK;
}

}  // namespace dem
