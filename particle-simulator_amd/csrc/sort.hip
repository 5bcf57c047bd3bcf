// Stable LSD radix sort of (uint64 key, int32 value) pairs, 8-bit digits.
//
// Per pass: (1) per-tile digit histograms in LDS, stored digit-major;
// (2) one device-wide exclusive scan turns them into each tile's output
// offset per digit; (3) each tile re-reads its keys in order and scatters
// them stably: inside a wave a key's rank among equal digits comes from
// eight ballots (one per digit bit), waves of the workgroup are ordered via
// per-wave digit counts in LDS, tiles via the scanned offsets.  Used to
// Morton-order particles for the octree.
#include "dem_common.h"
#include "scan.h"
#include "sort.h"

namespace dem {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kRounds = 8;
constexpr int kTile = kThreads * kRounds;  // keys per workgroup
constexpr int kRadix = 256;

__global__ __launch_bounds__(kThreads) void k_hist(const unsigned long long* __restrict__ keys,
                                                   int n, int shift, int nb,
                                                   int* __restrict__ hist) {
  __shared__ int h[kRadix];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int base = blockIdx.x * kTile;
  for (int r = 0; r < kRounds; ++r) {
    const int e = base + r * kThreads + threadIdx.x;
    if (e < n) atomicAdd(&h[(int)((keys[e] >> shift) & 0xffu)], 1);
  }
  __syncthreads();
  hist[threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(kThreads) void k_scatter(const unsigned long long* __restrict__ keys,
                                                      const int* __restrict__ vals, int n,
                                                      int shift, int nb,
                                                      const int* __restrict__ offs,
                                                      unsigned long long* __restrict__ okeys,
                                                      int* __restrict__ ovals) {
  __shared__ int base_d[kRadix];
  __shared__ int cnt[kWaves][kRadix];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  base_d[threadIdx.x] = offs[threadIdx.x * nb + blockIdx.x];
  for (int w = 0; w < kWaves; ++w) cnt[w][threadIdx.x] = 0;
  __syncthreads();
  const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int tbase = blockIdx.x * kTile;
  for (int r = 0; r < kRounds; ++r) {
    const int e = tbase + r * kThreads + threadIdx.x;
    const bool valid = e < n;
    unsigned long long k = valid ? keys[e] : 0ull;
    const int v = valid ? vals[e] : 0;
    const int d = (int)((k >> shift) & 0xffu);
    unsigned long long m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const unsigned long long bb = __ballot((d >> b) & 1);
      m &= ((d >> b) & 1) ? bb : ~bb;
    }
    const int rank = __popcll(m & lt);
    // the lowest lane of each equal-digit group publishes the group size
    if (valid && rank == 0) cnt[wave][d] = __popcll(m);
    __syncthreads();
    // per digit: prefix over waves, then advance the digit's running base
    {
      const int dd = threadIdx.x;
      int run = base_d[dd];
      for (int w = 0; w < kWaves; ++w) {
        const int c = cnt[w][dd];
        cnt[w][dd] = run;
        run += c;
      }
      base_d[dd] = run;
    }
    __syncthreads();
    if (valid) {
      const int pos = cnt[wave][d] + rank;
      okeys[pos] = k;
      ovals[pos] = v;
    }
    __syncthreads();
    for (int w = 0; w < kWaves; ++w) cnt[w][threadIdx.x] = 0;
    __syncthreads();
  }
}

}  // namespace

size_t radix_sort_workspace_ints(int n) {
  const int nb = (n + kTile - 1) / kTile;
  const size_t h = (size_t)kRadix * (nb > 0 ? nb : 1);
  return h + scan_workspace_ints((int)h) + 64;
}

int radix_sort_u64(unsigned long long* keys, int* vals, unsigned long long* keys_alt,
                   int* vals_alt, int n, int bits, int* ws, hipStream_t st) {
  if (n <= 1) return DEM_OK;
  const int nb = (n + kTile - 1) / kTile;
  int* hist = ws;
  int* scan_ws = ws + (size_t)kRadix * nb + 32;
  unsigned long long* src_k = keys;
  int* src_v = vals;
  unsigned long long* dst_k = keys_alt;
  int* dst_v = vals_alt;
  int passes = 0;
  for (int shift = 0; shift < bits; shift += 8, ++passes) {
    hipLaunchKernelGGL(k_hist, dim3(nb), dim3(kThreads), 0, st, (const unsigned long long*)src_k, n,
                       shift, nb, hist);
    DEM_CHECK_LAUNCH("k_hist");
    int rc = exclusive_scan(hist, hist, kRadix * nb, scan_ws, nullptr, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(kThreads), 0, st, (const unsigned long long*)src_k,
                       (const int*)src_v, n, shift, nb, (const int*)hist, dst_k, dst_v);
    DEM_CHECK_LAUNCH("k_scatter");
    unsigned long long* tk = src_k;
    src_k = dst_k;
    dst_k = tk;
    int* tv = src_v;
    src_v = dst_v;
    dst_v = tv;
  }
  if (passes & 1) {
    // result sits in the alternate buffers: copy back
    DEM_CHECK_HIP(hipMemcpyAsync(keys, src_k, sizeof(unsigned long long) * (size_t)n,
                                 hipMemcpyDeviceToDevice, st));
    DEM_CHECK_HIP(hipMemcpyAsync(vals, src_v, sizeof(int) * (size_t)n, hipMemcpyDeviceToDevice, st));
  }
  return DEM_OK;
}

}  // namespace dem
