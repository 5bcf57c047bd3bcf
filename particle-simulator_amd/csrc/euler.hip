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
