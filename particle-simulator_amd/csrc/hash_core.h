// Cell, bucket and contact predicate of the spatial-hash contact pass,
// shared by the single-system (hash.hip) and batched (batch.hip) kernels.
//
// Reference: dem/collision_detector/spatial_hash_collision_detector.py
//   _hash (:73-76), _assign_cells (:78-112), contact test (:205-225).
#pragma once
#include <math.h>
#include <stdint.h>

#include <hip/hip_runtime.h>

namespace dem {

__device__ __forceinline__ int pymod(int a, int g) {
  int m = a % g;
  return m < 0 ? m + g : m;
}

__device__ __forceinline__ int cell_of(double x, double cell, int g, int periodic) {
  int c = (int)floor(x / cell);
  if (periodic) return pymod(c, g);
  return c < 0 ? 0 : (c >= g ? g - 1 : c);
}

// Neighbour coordinate c + d, d in {-1, 0, 1}, of a cell c in [0, g): the
// periodic wrap (equal to pymod(c + d, g) for these arguments, without an
// integer division), or -1 outside an open box.
__device__ __forceinline__ int nb_coord(int c, int d, int g, int periodic) {
  const int a = c + d;
  if (periodic) return a < 0 ? a + g : (a >= g ? a - g : a);
  return (a < 0 || a >= g) ? -1 : a;
}

__device__ __forceinline__ int hash_cell(int ix, int iy, int iz, int mask) {
  uint32_t h = ((uint32_t)ix * 73856093u) ^ ((uint32_t)iy * 19349663u) ^ ((uint32_t)iz * 83492791u);
  return (int)(h & (uint32_t)mask);
}

// Contact predicate in float64 with the reference's element-wise rounding
// (no contraction): d = x_j - x_i per axis, minimum image |d| > L/2 ->
// d -= sign(d) L, then sqrt((dx^2 + dy^2) + dz^2) <= r_i + r_j (:209-225).
__device__ __forceinline__ bool contact_pair(double xi, double yi, double zi, double ri,
                                             double xj, double yj, double zj, double rj,
                                             double L, int periodic) {
#pragma clang fp contract(off)
  double dx = xj - xi;
  double dy = yj - yi;
  double dz = zj - zi;
  if (periodic) {
    const double h = L * 0.5;
    if (fabs(dx) > h) dx -= copysign(L, dx);
    if (fabs(dy) > h) dy -= copysign(L, dy);
    if (fabs(dz) > h) dz -= copysign(L, dz);
  }
  double d2 = dx * dx;
  d2 = d2 + dy * dy;
  d2 = d2 + dz * dz;
  const double rs = ri + rj;
  // Far pairs skip the square root: with t = fl(rs * rs) (1 + 2^-40) >
  // rs^2 (1 + 2^-41), d2 > t makes the exact sqrt(d2) exceed rs (1 + 2^-43),
  // so its correctly rounded value (within 2^-53 relative) is above rs too:
  // the same answer as the test below.
  if (d2 > rs * rs * (1.0 + 0x1p-40)) return false;
  // Near pairs are rare: the empty volatile statement keeps the square root
  // behind a branch that waves skip, where the compiler would otherwise
  // compute it (a quarter-rate v_rsq_f64 and its refinement) for every pair.
  __asm__ volatile("");
  const double dist = sqrt(d2);
  return dist <= rs;
}

}  // namespace dem
