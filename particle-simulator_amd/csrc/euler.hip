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

namespace {

__device__ __forceinline__ double py_fmod_pos(double x, double L) {
#pragma clang fp contract(off)
  double m = fmod(x, L);
  if (m != 0.0) {
    if ((L < 0.0) != (m < 0.0)) m += L;
  } else {
    m = copysign(0.0, L);
  }
  return m;
}

__global__ __launch_bounds__(256) void k_euler(double* __restrict__ pos,
                                               const double* __restrict__ F,
                                               const double* __restrict__ conv,
                                               const double* __restrict__ r, int n3, double sf,
                                               double dt, double L, int periodic) {
#pragma clang fp contract(off)
  int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n3) return;
  const int i = k / 3;
  const double denom = sf * r[i];
  double v = F[k] / denom;
  if (conv) v = v + conv[k];
  double x = pos[k] + v * dt;
  if (periodic) x = py_fmod_pos(x, L);
  pos[k] = x;
}

__global__ __launch_bounds__(256) void k_euler_total(double* __restrict__ pos,
                                                     const double* __restrict__ vel, int n3,
                                                     double dt, double L, int periodic) {
#pragma clang fp contract(off)
  int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n3) return;
  double x = pos[k] + vel[k] * dt;
  if (periodic) x = py_fmod_pos(x, L);
  pos[k] = x;
}

}  // namespace

extern "C" {

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
