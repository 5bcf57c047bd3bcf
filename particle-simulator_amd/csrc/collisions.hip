// Collision counts between volume classes over all pairs, for the population
// balance coupling (DirectCollisionFrequency.compute,
// pbm/collision_frequency/direct_collision_freq.py:15-41):
//   for i < j:  d = x_j - x_i;  d -= L * rint(d / L)  (periodic, L > 0)
//               |d| = sqrt(dx^2 + dy^2 + dz^2) <= r_i + r_j
//               -> counts[bin_i][bin_j] += 1, counts[bin_j][bin_i] += 1
// float64 with contraction off, in the order numpy evaluates it (np.round is
// round-half-to-even = rint; the norm is the square root of the sum of
// squares).  Each workgroup stages 256 sources at a time in LDS; the pairs
// i < j of a target block against a source chunk are counted with integer
// atomics, so the result is exact and independent of the launch shape.
#include <math.h>

#include "dem_common.h"
#include "profile.h"

namespace {

constexpr int kCB = 256;

__global__ void k_fill_counts(int* p, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ __launch_bounds__(kCB) void k_pair_collisions(
    const double* __restrict__ pos, const double* __restrict__ r, int n, double L, int wrap,
    const int* __restrict__ bins, int n_bins, int jchunk, int* __restrict__ counts) {
#pragma clang fp contract(off)
  __shared__ double sx[kCB], sy[kCB], sz[kCB], sr[kCB];
  __shared__ int sb[kCB];
  const int t = threadIdx.x;
  const int ifirst = blockIdx.x * kCB;
  const int i = ifirst + t;
  const bool live = i < n;
  double xi = 0.0, yi = 0.0, zi = 0.0, ri = 0.0;
  int bi = -1;
  if (live) {
    xi = pos[3 * (size_t)i];
    yi = pos[3 * (size_t)i + 1];
    zi = pos[3 * (size_t)i + 2];
    ri = r[i];
    bi = bins[i];
  }
  const bool bi_ok = bi >= 0 && bi < n_bins;
  const int j0 = blockIdx.y * jchunk;
  const int j1 = min(n, j0 + jchunk);
  for (int jt = j0; jt < j1; jt += kCB) {
    const int cnt = min(kCB, j1 - jt);
    if (jt + cnt - 1 <= ifirst) continue;  // every j <= every i of the block
    __syncthreads();
    if (t < cnt) {
      const int j = jt + t;
      sx[t] = pos[3 * (size_t)j];
      sy[t] = pos[3 * (size_t)j + 1];
      sz[t] = pos[3 * (size_t)j + 2];
      sr[t] = r[j];
      sb[t] = bins[j];
    }
    __syncthreads();
    if (!live) continue;
    for (int k = max(0, i + 1 - jt); k < cnt; ++k) {
      double dx = sx[k] - xi, dy = sy[k] - yi, dz = sz[k] - zi;
      if (wrap) {
        dx -= L * rint(dx / L);
        dy -= L * rint(dy / L);
        dz -= L * rint(dz / L);
      }
      const double dist = sqrt(dx * dx + dy * dy + dz * dz);
      if (dist <= ri + sr[k]) {
        const int bj = sb[k];
        if (bi_ok && bj >= 0 && bj < n_bins) {
          atomicAdd(&counts[bi * n_bins + bj], 1);
          atomicAdd(&counts[bj * n_bins + bi], 1);
        }
      }
    }
  }
}

}  // namespace

extern "C" {

int dem_direct_collision_counts(const double* pos, const double* radii, int n, double L,
                                int periodic, const int32_t* bins, int n_bins, int32_t* counts,
                                dem_stream_t stream) {
  if (n < 0) return dem::set_error(DEM_ERR_INVALID, "n = %d < 0", n);
  if (n_bins < 1 || (long long)n_bins * n_bins > (1ll << 30))
    return dem::set_error(DEM_ERR_INVALID, "n_bins %d out of range", n_bins);
  if (!counts) return dem::set_error(DEM_ERR_INVALID, "null counts");
  if (n > 0 && (!pos || !radii || !bins))
    return dem::set_error(DEM_ERR_INVALID, "null positions, radii or bins");
  hipStream_t st = dem::as_stream(stream);
  const int nb2 = n_bins * n_bins;
  hipLaunchKernelGGL(k_fill_counts, dim3((nb2 + 255) / 256), dim3(256), 0, st, counts, nb2);
  DEM_CHECK_LAUNCH("k_fill_counts");
  if (n < 2) return DEM_OK;
  const int iblocks = (n + kCB - 1) / kCB;
  // about 2048 workgroups; source chunks are whole tiles
  int S = (2048 + iblocks - 1) / iblocks;
  const int tiles = (n + kCB - 1) / kCB;
  if (S > tiles) S = tiles;
  if (S < 1) S = 1;
  const int jchunk = ((tiles + S - 1) / S) * kCB;
  S = (n + jchunk - 1) / jchunk;
  const int wrap = (periodic && L > 0.0) ? 1 : 0;
  DEM_PROF_LAUNCH("k_pair_collisions", st,
                  hipLaunchKernelGGL(k_pair_collisions, dim3(iblocks, S), dim3(kCB), 0, st, pos,
                                     radii, n, L, wrap, bins, n_bins, jchunk, counts));
  DEM_CHECK_LAUNCH("k_pair_collisions");
  return DEM_OK;
}

}  // extern "C"
