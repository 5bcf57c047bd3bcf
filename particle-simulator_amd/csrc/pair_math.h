// Per-pair FP32 physics shared by the direct, batched and tree kernels.
//
// Units: D = x_j - x_i in fixed-point units (2^22 fixed units = 1 work unit);
// radii in work units.  With d = sigma D (sigma = 2^-22, work units):
//   dipole force on i (reference direct_droplet_force_calculator.py:183-200):
//     F_i += m r_i^3 r_j^3 / |d|^7 (dx(4dz^2-dx^2-dy^2), dy(..), dz(2dz^2-3dx^2-3dy^2))
//   accumulated here as  r_j^3 g(D) / |d|^7  (the caller scales by
//   m lambda^2 sigma^3 r_i^3 once per target);
//   Oseen velocity at i (:401-411):  v_i += eta (F_j/|x| + x (x.F_j)/|x|^3),
//   accumulated as F_j/|D| + D (D.F_j)/|D|^3 in fixed units (the caller
//   scales by eta Fs / U).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dem {

constexpr float kSigma2f = 0x1p-44f;  // (work unit / fixed unit)^2

// Contact rule: tree leaves keep r^2 >= (r_i + r_j)^2 (flat_tree.py:835), the
// direct sum keeps |d| > r_i + r_j (direct...py:246).
template <bool INCLUSIVE>
__device__ __forceinline__ bool separated(float r2w, float rs) {
  return INCLUSIVE ? (r2w >= rs * rs) : (r2w > rs * rs);
}

// |d|^2 in work units from fixed-unit differences, in the order the dipole
// kernels evaluate it (so contact decisions agree wherever it is used).
__device__ __forceinline__ float r2_work(float dx, float dy, float dz) {
  const float dz2 = dz * dz;
  const float dxy2 = dx * dx + dy * dy;
  return (dxy2 + dz2) * kSigma2f;
}

// Dipole force of one source; excluded pairs (contact, keep == false for the
// caller's own exclusion such as j == i) add nothing.
template <bool INCLUSIVE>
__device__ __forceinline__ void dipole_pair(float dx, float dy, float dz, float ri, float rj,
                                            float rj3, bool keep, float& fx, float& fy,
                                            float& fz) {
  const float dz2 = dz * dz;
  const float dxy2 = dx * dx + dy * dy;
  const float rr = dxy2 + dz2;  // |D|^2, fixed units; r2_work's rounding
  const float r2 = rr * kSigma2f;
  const float ir = __builtin_amdgcn_rsqf(r2);
  const float ir2 = ir * ir;
  const float ir7 = (ir2 * ir2) * (ir2 * ir);
  // select, never 0*inf
  const float s7 = (keep && separated<INCLUSIVE>(r2, ri + rj)) ? rj3 * ir7 : 0.f;
  const float A = 4.f * dz2 - dxy2;
  const float txy = s7 * A;
  fx += txy * dx;
  fy += txy * dy;
  fz += (s7 * fmaf(-2.f, rr, A)) * dz;  // 2 dz^2 - 3 (dx^2 + dy^2)
}

template <bool INCLUSIVE>
__device__ __forceinline__ void dipole_pair(float dx, float dy, float dz, float ri, float rj,
                                            float rj3, float& fx, float& fy, float& fz) {
  dipole_pair<INCLUSIVE>(dx, dy, dz, ri, rj, rj3, true, fx, fy, fz);
}

// Dipole force of an aggregate (R^3 sum w3 at distance D); no contact rule.
__device__ __forceinline__ void dipole_monopole(float dx, float dy, float dz, float w3, float& fx,
                                                float& fy, float& fz) {
  const float dz2 = dz * dz;
  const float dxy2 = dx * dx + dy * dy;
  const float rr = dxy2 + dz2;
  const float ir = __builtin_amdgcn_rsqf(rr * kSigma2f);
  const float ir2 = ir * ir;
  const float s7 = w3 * ((ir2 * ir2) * (ir2 * ir));
  const float A = 4.f * dz2 - dxy2;
  const float txy = s7 * A;
  fx += txy * dx;
  fy += txy * dy;
  fz += (s7 * fmaf(-2.f, rr, A)) * dz;
}

// Oseen tensor applied to source force (Fx,Fy,Fz); keep == false adds nothing.
__device__ __forceinline__ void oseen_pair(float dx, float dy, float dz, float Fx, float Fy,
                                           float Fz, bool keep, float& vx, float& vy, float& vz) {
  const float r2 = dx * dx + dy * dy + dz * dz;
  const float ir = keep ? __builtin_amdgcn_rsqf(r2) : 0.f;
  const float t = (dx * Fx + dy * Fy + dz * Fz) * (ir * ir);
  vx += ir * (Fx + t * dx);
  vy += ir * (Fy + t * dy);
  vz += ir * (Fz + t * dz);
}

// Oseen term of a tree-leaf member (fixed units) with the leaf contact rule
// |D| >= r_i + r_j tested on fixed-unit radii (rsf = (r_i + r_j) 2^22): the
// work-unit test scaled by a power of two, so it decides the same.  The
// target itself (D = 0) fails it.  Returns whether the member was kept.
__device__ __forceinline__ bool oseen_member(float dx, float dy, float dz, float Fx, float Fy,
                                             float Fz, float rsf, bool keep, float& vx, float& vy,
                                             float& vz) {
  const float r2 = dx * dx + dy * dy + dz * dz;
  const bool sep = keep && r2 >= rsf * rsf;
  const float ir = sep ? __builtin_amdgcn_rsqf(r2) : 0.f;
  const float t = (dx * Fx + dy * Fy + dz * Fz) * (ir * ir);
  vx += ir * (Fx + t * dx);
  vy += ir * (Fy + t * dy);
  vz += ir * (Fz + t * dz);
  return sep;
}

// High word (floor) of the 64-bit minimum-image difference qj - qi, as float.
__device__ __forceinline__ float diff64(uint32_t jh, uint32_t jl, unsigned long long qi) {
  return (float)(int32_t)(((((unsigned long long)jh << 32) | jl) - qi) >> 32);
}

// The same difference from the 64-bit source position and the negated target
// position: qj + (-qi) is one 64-bit add (v_lshl_add_u64) instead of a
// subtract with borrow; bit-identical to diff64.
__device__ __forceinline__ float diff64n(unsigned long long qj, unsigned long long nqi) {
  return (float)(int32_t)((qj + nqi) >> 32);
}

// 32-bit minimum-image difference qj - qi, as float.
__device__ __forceinline__ float diff32(uint32_t qj, uint32_t qi) {
  return (float)(int32_t)(qj - qi);
}

// ---------------------------------------------------------------- COMSOL lookups

struct CorrGrid {
  const float4* grid;  // (u, v, w, 0) per node, ((ix*n)+iy)*n+iz
  int n;
  float nmax;          // n - 1
  float ga, gb;        // grid coordinate of x_i - x_j = -U D  is  ga * D + gb
};

// Clamped grid coordinate: index in [0, n-2] and fraction
// (direct_droplet_force_calculator.py:141-157).
__device__ __forceinline__ void grid_coord(float g, float nmax, int n, int& i, float& t) {
  g = fminf(fmaxf(g, 0.f), nmax);
  i = min((int)floorf(g), n - 2);
  t = g - (float)i;
}

// Three permuted trilinear lookups (:426-448) applied to force (fx,fy,fz):
// G_z at (x,y,z) -> (u,v,w), G_x at (z,y,x) -> (w,v,u), G_y at (x,z,y) -> (u,w,v).
// The lookups share the index math and the eight weights (permuted).
__device__ __forceinline__ void corr_lookup(const CorrGrid& cg, float dx, float dy, float dz,
                                            float fx, float fy, float fz, float& ux, float& uy,
                                            float& uz) {
  int gx, gy, gz;
  float tx, ty, tz;
  grid_coord(cg.ga * dx + cg.gb, cg.nmax, cg.n, gx, tx);
  grid_coord(cg.ga * dy + cg.gb, cg.nmax, cg.n, gy, ty);
  grid_coord(cg.ga * dz + cg.gb, cg.nmax, cg.n, gz, tz);
  float w[2][2][2];
  const float ox = 1.f - tx, oy = 1.f - ty, oz = 1.f - tz;
  w[0][0][0] = ox * oy * oz; w[0][0][1] = ox * oy * tz;
  w[0][1][0] = ox * ty * oz; w[0][1][1] = ox * ty * tz;
  w[1][0][0] = tx * oy * oz; w[1][0][1] = tx * oy * tz;
  w[1][1][0] = tx * ty * oz; w[1][1][1] = tx * ty * tz;
  const int nn = cg.n;
  float gzx = 0.f, gzy = 0.f, gzz = 0.f, gxx = 0.f, gxy = 0.f, gxz = 0.f, gyx = 0.f, gyy = 0.f,
        gyz = 0.f;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const float4 g1 = cg.grid[((gx + a) * nn + (gy + b)) * nn + (gz + c)];
        const float w1 = w[a][b][c];
        gzx += w1 * g1.x; gzy += w1 * g1.y; gzz += w1 * g1.z;
        const float4 g2 = cg.grid[((gz + a) * nn + (gy + b)) * nn + (gx + c)];
        const float w2 = w[c][b][a];
        gxx += w2 * g2.z; gxy += w2 * g2.y; gxz += w2 * g2.x;
        const float4 g3 = cg.grid[((gx + a) * nn + (gz + b)) * nn + (gy + c)];
        const float w3 = w[a][c][b];
        gyx += w3 * g3.x; gyy += w3 * g3.z; gyz += w3 * g3.y;
      }
  ux += gxx * fx + gyx * fy + gzx * fz;
  uy += gxy * fx + gyy * fy + gzy * fz;
  uz += gxz * fx + gyz * fy + gzz * fz;
}

}  // namespace dem
