// Explicit Euler step in float64, rounding element by element like the
// reference's numpy expressions (dem/solver/euler_droplet_solver.py):
//   direct path (:124-125):  mig = F / (stokes_factor * r[:, None])
//                            x  += (mig + conv) * dt
//   tree path   (:119):      x  += v_total * dt
//   periodic wrap (:128-131): x %= L   (numpy float remainder: fmod, then +L
//                             when the result is negative, +0.0 when zero)
// The step is HBM-bound (a few float64 loads/stores per coordinate); FMA
// contraction is disabled so results are bit-identical to numpy's.
#include <math.h>

#include "dem_common.h"
#include "euler_core.h"

namespace {

__global__ __launch_bounds__(256) void k_euler(double* __restrict__ pos,
                                               const double* __restrict__ F,
                                               const double* __restrict__ conv,
                                               const double* __restrict__ r, int n3, double sf,
                                               double dt, double L, int periodic) {
  int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n3) return;
  const int i = k / 3;
  pos[k] = dem::euler_coord(pos[k], F[k], conv ? conv[k] : 0.0, conv != nullptr, r[i], sf, dt, L,
                            periodic);
}

__global__ __launch_bounds__(256) void k_euler_total(double* __restrict__ pos,
                                                     const double* __restrict__ vel, int n3,
                                                     double dt, double L, int periodic) {
  int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n3) return;
  pos[k] = dem::euler_total_coord(pos[k], vel[k], dt, L, periodic);
}

// Merge compaction (droplet_solution.py:263-268 after the greedy plan):
// survivors keep their order, merged droplets are appended.  `removed` holds
// the n_removed distinct removed indices in ascending order, so a survivor's
// new index is i minus the number of removed indices below it (binary
// search).  Writes are bounded to [0, n_out), so malformed input (unsorted,
// repeated or out-of-range indices) cannot write outside the outputs.
__global__ __launch_bounds__(256) void k_merge_compact(
    const double* __restrict__ pos, const double* __restrict__ r, int n,
    const int32_t* __restrict__ removed, int n_removed, const double* __restrict__ merged,
    int n_merged, int n_out, double* __restrict__ out_pos, double* __restrict__ out_r) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) {
    int lo = 0, hi = n_removed;  // first removed index >= i
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (removed[mid] < i) lo = mid + 1; else hi = mid;
    }
    const bool gone = lo < n_removed && removed[lo] == i;
    const int o = i - lo;
    if (!gone && o >= 0 && o < n_out) {
      out_pos[3 * (size_t)o] = pos[3 * (size_t)i];
      out_pos[3 * (size_t)o + 1] = pos[3 * (size_t)i + 1];
      out_pos[3 * (size_t)o + 2] = pos[3 * (size_t)i + 2];
      out_r[o] = r[i];
    }
  }
  if (i < n_merged) {
    const int o = n - n_removed + i;
    if (o >= 0 && o < n_out) {
      out_pos[3 * (size_t)o] = merged[4 * (size_t)i];
      out_pos[3 * (size_t)o + 1] = merged[4 * (size_t)i + 1];
      out_pos[3 * (size_t)o + 2] = merged[4 * (size_t)i + 2];
      out_r[o] = merged[4 * (size_t)i + 3];
    }
  }
}

// (x, y, z, r) of the droplets idx[0 .. k) (the values a merge plan reads).
__global__ __launch_bounds__(256) void k_gather_droplets(const double* __restrict__ pos,
                                                         const double* __restrict__ r, int n,
                                                         const int32_t* __restrict__ idx, int k,
                                                         double* __restrict__ out) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= k) return;
  const int i = idx[q];
  const bool ok = i >= 0 && i < n;
  out[4 * (size_t)q] = ok ? pos[3 * (size_t)i] : NAN;
  out[4 * (size_t)q + 1] = ok ? pos[3 * (size_t)i + 1] : NAN;
  out[4 * (size_t)q + 2] = ok ? pos[3 * (size_t)i + 2] : NAN;
  out[4 * (size_t)q + 3] = ok ? r[i] : NAN;
}

}  // namespace

extern "C" {

int dem_gather_droplets(const double* pos, const double* radii, int n, const int32_t* idx, int k,
                        double* out, dem_stream_t stream) {
  if (n < 0 || k < 0) return dem::set_error(DEM_ERR_INVALID, "bad counts (n %d, k %d)", n, k);
  if (k == 0) return DEM_OK;
  if (!pos || !radii || !idx || !out) return dem::set_error(DEM_ERR_INVALID, "null argument");
  hipLaunchKernelGGL(k_gather_droplets, dim3((k + 255) / 256), dim3(256), 0,
                     dem::as_stream(stream), pos, radii, n, idx, k, out);
  DEM_CHECK_LAUNCH("k_gather_droplets");
  return DEM_OK;
}

int dem_merge_compact(const double* pos, const double* radii, int n, const int32_t* removed,
                      int n_removed, const double* merged, int n_merged, double* out_pos,
                      double* out_radii, dem_stream_t stream) {
  if (n < 0 || n_removed < 0 || n_removed > n || n_merged < 0)
    return dem::set_error(DEM_ERR_INVALID, "bad counts (n %d, removed %d, merged %d)", n,
                          n_removed, n_merged);
  const int n_out = n - n_removed + n_merged;
  if (n_out == 0) return DEM_OK;
  if (!out_pos || !out_radii || (n > 0 && (!pos || !radii)) || (n_removed > 0 && !removed) ||
      (n_merged > 0 && !merged))
    return dem::set_error(DEM_ERR_INVALID, "null argument");
  const int m = n > n_merged ? n : n_merged;
  hipLaunchKernelGGL(k_merge_compact, dim3((m + 255) / 256), dim3(256), 0, dem::as_stream(stream),
                     pos, radii, n, removed, n_removed, merged, n_merged, n_out, out_pos,
                     out_radii);
  DEM_CHECK_LAUNCH("k_merge_compact");
  return DEM_OK;
}

int dem_euler_step(double* pos, const double* forces, const double* conv, const double* radii,
                   int n, double stokes_factor, double dt, double L, int periodic,
                   dem_stream_t stream) {
  if (n < 0) return dem::set_error(DEM_ERR_INVALID, "n must be >= 0");
  if (n == 0) return DEM_OK;
  if (!pos || !forces || !radii) return dem::set_error(DEM_ERR_INVALID, "null argument");
  if (periodic && !(L > 0.0)) return dem::set_error(DEM_ERR_INVALID, "periodic wrap needs L > 0");
  const int n3 = 3 * n;
  hipLaunchKernelGGL(k_euler, dim3((n3 + 255) / 256), dim3(256), 0, dem::as_stream(stream), pos,
                     forces, conv, radii, n3, stokes_factor, dt, L, periodic ? 1 : 0);
  DEM_CHECK_LAUNCH("k_euler");
  return DEM_OK;
}

int dem_euler_step_total(double* pos, const double* vel, int n, double dt, double L, int periodic,
                         dem_stream_t stream) {
  if (n < 0) return dem::set_error(DEM_ERR_INVALID, "n must be >= 0");
  if (n == 0) return DEM_OK;
  if (!pos || !vel) return dem::set_error(DEM_ERR_INVALID, "null argument");
  if (periodic && !(L > 0.0)) return dem::set_error(DEM_ERR_INVALID, "periodic wrap needs L > 0");
  const int n3 = 3 * n;
  hipLaunchKernelGGL(k_euler_total, dim3((n3 + 255) / 256), dim3(256), 0, dem::as_stream(stream),
                     pos, vel, n3, dt, L, periodic ? 1 : 0);
  DEM_CHECK_LAUNCH("k_euler_total");
  return DEM_OK;
}

}  // extern "C"
