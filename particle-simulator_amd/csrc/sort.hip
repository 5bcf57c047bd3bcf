// Stable LSD radix sort of (uint64 key, int32 value) pairs, 8-bit digits.
//
// Per pass, three launches: (1) per-tile digit histograms in LDS, stored
// digit-major; (2) a single-launch tile scan of those histograms (each
// scatter workgroup adds the prefix of the scan-tile sums itself), giving
// each tile's output offset per digit; (3) each tile re-reads its keys in
// order and scatters them stably: inside a wave a key's rank among equal
// digits comes from eight ballots (one per digit bit), waves of the workgroup
// are ordered via per-wave digit counts in LDS, tiles via the scanned
// offsets.  Used to Morton-order particles for the octree.
#include "dem_common.h"
#include "sort.h"
#include "tile_scan.h"

namespace dem {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kRounds = 8;
constexpr int kTile = kThreads * kRounds;  // keys per workgroup
constexpr int kDigitBits = 8;
constexpr int kRadix = 1 << kDigitBits;
constexpr unsigned kDigitMask = kRadix - 1;

__global__ __launch_bounds__(kThreads) void k_hist(const unsigned long long* __restrict__ keys,
                                                   int n, int shift, int nb,
                                                   int* __restrict__ hist) {
  __shared__ int h[kRadix];
  for (int d = threadIdx.x; d < kRadix; d += kThreads) h[d] = 0;
  __syncthreads();
  const int base = blockIdx.x * kTile;
  for (int r = 0; r < kRounds; ++r) {
    const int e = base + r * kThreads + threadIdx.x;
    if (e < n) atomicAdd(&h[(int)((keys[e] >> shift) & kDigitMask)], 1);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < kRadix; d += kThreads) hist[d * nb + blockIdx.x] = h[d];
}

__global__ __launch_bounds__(kThreads) void k_hist_scan(int* __restrict__ hist, int m,
                                                        int* __restrict__ tile_sums) {
  ts_tile_scan([=](int k) { return hist[k]; }, m, hist, tile_sums, nullptr);
}

__global__ __launch_bounds__(kThreads) void k_scatter(const unsigned long long* __restrict__ keys,
                                                      const int* __restrict__ vals, int n,
                                                      int shift, int nb,
                                                      const int* __restrict__ offs,
                                                      const int* __restrict__ tile_sums, int nt,
                                                      unsigned long long* __restrict__ okeys,
                                                      int* __restrict__ ovals) {
  extern __shared__ int pre[];
  __shared__ int base_d[kRadix];
  __shared__ int cnt[kWaves][kRadix];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  ts_tile_prefix(tile_sums, nt, pre);
  for (int d = threadIdx.x; d < kRadix; d += kThreads) {
    const int idx = d * nb + blockIdx.x;
    base_d[d] = offs[idx] + pre[idx / kScanTile];
    for (int w = 0; w < kWaves; ++w) cnt[w][d] = 0;
  }
  __syncthreads();
  const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int tbase = blockIdx.x * kTile;
  for (int r = 0; r < kRounds; ++r) {
    const int e = tbase + r * kThreads + threadIdx.x;
    const bool valid = e < n;
    unsigned long long k = valid ? keys[e] : 0ull;
    const int v = valid ? vals[e] : 0;
    const int d = (int)((k >> shift) & kDigitMask);
    unsigned long long m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < kDigitBits; ++b) {
      const unsigned long long bb = __ballot((d >> b) & 1);
      m &= ((d >> b) & 1) ? bb : ~bb;
    }
    const int rank = __popcll(m & lt);
    // the lowest lane of each equal-digit group publishes the group size
    if (valid && rank == 0) cnt[wave][d] = __popcll(m);
    __syncthreads();
    // per digit: prefix over waves, then advance the digit's running base
    for (int dd = threadIdx.x; dd < kRadix; dd += kThreads) {
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
    for (int dd = threadIdx.x; dd < kRadix; dd += kThreads)
      for (int w = 0; w < kWaves; ++w) cnt[w][dd] = 0;
    __syncthreads();
  }
}

}  // namespace

size_t radix_sort_workspace_ints(int n) {
  const int nb = (n + kTile - 1) / kTile;
  const size_t h = (size_t)kRadix * (nb > 0 ? nb : 1);
  return h + (size_t)scan_tiles((int)h) + 64;
}

int radix_sort_u64(unsigned long long* keys, int* vals, unsigned long long* keys_alt,
                   int* vals_alt, int n, int begin_bit, int end_bit, int* ws, hipStream_t st,
                   unsigned long long** sorted_keys) {
  if (sorted_keys) *sorted_keys = keys;
  if (n <= 1 || end_bit <= begin_bit) return DEM_OK;
  const int nb = (n + kTile - 1) / kTile;
  const int m = kRadix * nb;
  const int nt = scan_tiles(m);
  int* hist = ws;
  int* tile_sums = ws + (size_t)m + 32;
  unsigned long long* src_k = keys;
  int* src_v = vals;
  unsigned long long* dst_k = keys_alt;
  int* dst_v = vals_alt;
  int passes = 0;
  for (int shift = begin_bit; shift < end_bit; shift += kDigitBits, ++passes) {
    hipLaunchKernelGGL(k_hist, dim3(nb), dim3(kThreads), 0, st, (const unsigned long long*)src_k, n,
                       shift, nb, hist);
    DEM_CHECK_LAUNCH("k_hist");
    hipLaunchKernelGGL(k_hist_scan, dim3(nt), dim3(kThreads), 0, st, hist, m, tile_sums);
    DEM_CHECK_LAUNCH("k_hist_scan");
    hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(kThreads), nt * sizeof(int), st,
                       (const unsigned long long*)src_k, (const int*)src_v, n, shift, nb,
                       (const int*)hist, (const int*)tile_sums, nt, dst_k, dst_v);
    DEM_CHECK_LAUNCH("k_scatter");
    unsigned long long* tk = src_k;
    src_k = dst_k;
    dst_k = tk;
    int* tv = src_v;
    src_v = dst_v;
    dst_v = tv;
  }
  if (passes & 1) {
    // the result sits in the alternate buffers: values go back, keys too
    // unless the caller takes them where they are
    DEM_CHECK_HIP(hipMemcpyAsync(vals, src_v, sizeof(int) * (size_t)n, hipMemcpyDeviceToDevice, st));
    if (sorted_keys)
      *sorted_keys = src_k;
    else
      DEM_CHECK_HIP(hipMemcpyAsync(keys, src_k, sizeof(unsigned long long) * (size_t)n,
                                   hipMemcpyDeviceToDevice, st));
  }
  return DEM_OK;
}

}  // namespace dem
