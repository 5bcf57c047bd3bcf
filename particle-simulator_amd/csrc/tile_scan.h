// Single-launch exclusive scans for the build passes.
//
// A scan of n int32 values runs as one kernel over tiles of kScanTile
// entries: each workgroup writes its tile's local exclusive prefix and the
// tile's sum (and may add the sum to a grand total with an integer atomic:
// exact, order-free).  A consumer that needs the global prefix of entry k adds
// the sum of the tiles before k's tile, which it forms itself in LDS from the
// tile sums (ts_tile_prefix) — so no second and third launch are needed.
#pragma once
#include <hip/hip_runtime.h>

namespace dem {

constexpr int kScanThreads = 256;
constexpr int kScanTileItems = 16;
constexpr int kScanTile = kScanThreads * kScanTileItems;  // entries per tile

__host__ __device__ inline int scan_tiles(int n) { return (n + kScanTile - 1) / kScanTile; }

__device__ __forceinline__ int ts_wave_inclusive(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(v, o);
    if (lane >= o) v += u;
  }
  return v;
}

// Exclusive scan of one value per thread over a kScanThreads block.
__device__ __forceinline__ int ts_block_exclusive(int v, int* wsum, int* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int inc = ts_wave_inclusive(v);
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  int before = 0, all = 0;
#pragma unroll
  for (int k = 0; k < kScanThreads / 64; ++k) {
    const int s = wsum[k];
    before += k < w ? s : 0;
    all += s;
  }
  __syncthreads();
  *total = all;
  return before + inc - v;
}

// pre[t] = sum of tile_sums[0..t) for t < nt, in LDS (nt ints); returns the
// grand total.  Called by every thread of a kScanThreads block.
__device__ __forceinline__ int ts_tile_prefix(const int* __restrict__ tile_sums, int nt, int* pre) {
  __shared__ int wsum[kScanThreads / 64];
  const int per = (nt + kScanThreads - 1) / kScanThreads;
  const int b = threadIdx.x * per;
  int sum = 0;
  for (int k = 0; k < per; ++k)
    if (b + k < nt) sum += tile_sums[b + k];
  int total;
  int run = ts_block_exclusive(sum, wsum, &total);
  for (int k = 0; k < per; ++k)
    if (b + k < nt) {
      const int v = tile_sums[b + k];
      pre[b + k] = run;
      run += v;
    }
  __syncthreads();
  return total;
}

// One tile (blockIdx.x) of the scan of value(k), k < n: out[k] = exclusive
// prefix within the tile, tile_sums[tile] = its sum, and, when total is not
// null, atomicAdd(total, tile sum).
template <class Value>
__device__ __forceinline__ void ts_tile_scan(Value value, int n, int* __restrict__ out,
                                             int* __restrict__ tile_sums, int* __restrict__ total) {
  // padded LDS: entry e at e + e / 32
  __shared__ int s[kScanTile + kScanTile / 32];
  __shared__ int wsum[kScanThreads / 64];
  const int base = blockIdx.x * kScanTile, t = threadIdx.x;
#pragma unroll
  for (int k = 0; k < kScanTileItems; ++k) {
    const int e = k * kScanThreads + t;
    s[e + (e >> 5)] = base + e < n ? value(base + e) : 0;
  }
  __syncthreads();
  int v[kScanTileItems];
  int sum = 0;
#pragma unroll
  for (int k = 0; k < kScanTileItems; ++k) {
    const int e = t * kScanTileItems + k;
    v[k] = s[e + (e >> 5)];
    sum += v[k];
  }
  int tile_total;
  int run = ts_block_exclusive(sum, wsum, &tile_total);
#pragma unroll
  for (int k = 0; k < kScanTileItems; ++k) {
    const int e = t * kScanTileItems + k;
    s[e + (e >> 5)] = run;
    run += v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kScanTileItems; ++k) {
    const int e = k * kScanThreads + t;
    if (base + e < n) out[base + e] = s[e + (e >> 5)];
  }
  if (t == 0) {
    tile_sums[blockIdx.x] = tile_total;
    if (total) atomicAdd(total, tile_total);
  }
}

}  // namespace dem
