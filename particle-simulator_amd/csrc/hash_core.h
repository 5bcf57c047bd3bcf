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

__device__ __forceinline__ int hash_cell(int ix, int iy, int iz, int mask) {
  uint32_t h = ((uint32_t)ix * 73856093u) ^ ((uint32_t)iy * 19349663u) ^ ((uint32_t)iz * 83492791u);
  return (int)(h & (uint32_t)mask);
}

// Contact predicate in float64, element-wise rounding (no contraction).
__device__ __noinline__ bool contact(const double* __restrict__ pos, const double* __restrict__ r,
                                     int i, int j, double L, int periodic) {
#pragma clang fp contract(off)
  double dx = pos[3 * (size_t)j] - pos[3 * (size_t)i];
  double dy = pos[3 * (size_t)j + 1] - pos[3 * (size_t)i + 1];
  double dz = pos[3 * (size_t)j + 2] - pos[3 * (size_t)i + 2];
  if (periodic) {
    const double h = L * 0.5;
    if (fabs(dx) > h) dx -= copysign(L, dx);
    if (fabs(dy) > h) dy -= copysign(L, dy);
    if (fabs(dz) > h) dz -= copysign(L, dz);
  }
  double d2 = dx * dx;
  d2 = d2 + dy * dy;
  d2 = d2 + dz * dz;
  double dist = sqrt(d2);
  double rs = r[i] + r[j];
  return dist <= rs;
}

}  // namespace dem
