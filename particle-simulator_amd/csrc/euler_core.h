// Float64 Euler update of one coordinate, rounding like the reference's
// element-wise numpy expressions (dem/solver/euler_droplet_solver.py):
//   direct path (:124-125): x += (F / (stokes_factor * r) + conv) * dt
//   tree path   (:119):     x += v_total * dt
//   periodic wrap (:128-131): x %= L with numpy's float remainder.
// FMA contraction is off so every operation rounds on its own.
#pragma once
#include <math.h>

#include <hip/hip_runtime.h>

namespace dem {

// numpy float remainder x % L: fmod, then +L when the signs differ, and a zero
// result takes the sign of L.
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

__device__ __forceinline__ double euler_coord(double x, double F, double conv, bool has_conv,
                                              double r, double sf, double dt, double L,
                                              int periodic) {
#pragma clang fp contract(off)
  const double denom = sf * r;
  double v = F / denom;
  if (has_conv) v = v + conv;
  x = x + v * dt;
  if (periodic) x = py_fmod_pos(x, L);
  return x;
}

__device__ __forceinline__ double euler_total_coord(double x, double v, double dt, double L,
                                                    int periodic) {
#pragma clang fp contract(off)
  x = x + v * dt;
  if (periodic) x = py_fmod_pos(x, L);
  return x;
}

}  // namespace dem
