// Records, scales and per-target sums of the direct dipole + Oseen evaluation,
// shared by the single-system kernels (direct.hip) and the batched-ensemble
// kernels (batch.hip).  Both call these functions with the same split of the
// source range, so a batched ensemble reproduces the single-system result bit
// for bit.
//
// Reference: dem/force_calculator/direct_droplet_force_calculator.py
//   _calculate_forces (:202-253), _calculate_convection_velocities (:354-468).
#pragma once
#include <math.h>
#include <stdint.h>

#include <hip/hip_runtime.h>

#include "pair_math.h"

namespace dem {

constexpr double kTwo32 = 4294967296.0;
constexpr double kSigma = 0x1p-22;  // work unit / fixed unit
constexpr double kSigma3 = 0x1p-66;

// Phase-1 source / target record (32 bytes, read whole by the force loop):
// 32-bit fixed-point position (rounded), radius and radius^3 in work units.
struct alignas(16) Rec {
  uint32_t x, y, z;
  float r;
  float r3;
  uint32_t pad[3];
};
static_assert(sizeof(Rec) == 32, "record layout");

// Phase-2 record (48 bytes): 64-bit fixed-point position split into high and
// low words, and the source force in units of Fs.  The Oseen sum decays as
// 1/r, so a pair whose separation lies within one 32-bit quantum of L/2 would
// change the velocity visibly if it took the other periodic image; 64-bit
// positions make the image choice agree with the float64 rule |d| > L/2.
struct alignas(16) RecC {
  uint32_t xh, yh, zh;
  float fx, fy, fz;
  uint32_t xl, yl, zl;
  uint32_t pad[3];
};
static_assert(sizeof(RecC) == 48, "record layout");

// Per-system scales.
struct DevScales {
  double origin[3];  // subtracted before quantisation (open boxes)
  double inv_U;      // fixed units per metre
  double U;          // metres per fixed unit
  double lambda;     // metres per work unit (= U * 2^22)
  double Fs;         // newtons per force unit of RecC::f
  int periodic;
  int pad;
};

// Scales of a periodic box [0,L)^3: the box maps onto the uint32 ring.
__host__ __device__ inline DevScales periodic_scales(double L, double m_const) {
  DevScales s;
  s.periodic = 1;
  s.pad = 0;
  s.origin[0] = s.origin[1] = s.origin[2] = 0.0;
  s.U = L / kTwo32;
  s.inv_U = kTwo32 / L;
  s.lambda = s.U / kSigma;
  double fs = m_const * s.lambda * s.lambda;
  s.Fs = (fs > 0.0 && isfinite(fs)) ? fs : 1.0;
  return s;
}

// Open box with bounding-box corner lo and largest extent ext: the box maps
// onto [0, 2^31 - 2^12], so differences never leave int32.
__host__ __device__ inline DevScales open_scales(const double lo[3], double ext, double m_const) {
  DevScales s;
  s.periodic = 0;
  s.pad = 0;
  if (!(ext > 0.0) || !isfinite(ext)) ext = 1.0;
  for (int a = 0; a < 3; ++a) s.origin[a] = isfinite(lo[a]) ? lo[a] : 0.0;
  s.U = ext / (2147483648.0 - 4096.0);
  s.inv_U = 1.0 / s.U;
  s.lambda = s.U / kSigma;
  double fs = m_const * s.lambda * s.lambda;
  s.Fs = (fs > 0.0 && isfinite(fs)) ? fs : 1.0;
  return s;
}

// Quantise one droplet (position p[3], radius r, optional source force f[3]).
__device__ __forceinline__ void pack_one(const double* p, double r, const double* f,
                                         const DevScales& sc, Rec& o, RecC& c) {
  unsigned long long q[3];
  for (int a = 0; a < 3; ++a) {
    double x = p[a];
    double v;
    if (sc.periodic) {
      double u = x * sc.inv_U * 0x1p-32;  // x / L
      u -= floor(u);
      v = rint(u * 0x1p64);
      if (v >= 0x1p64) v = 0.0;  // u rounded up to 1: wraps to the origin
    } else {
      v = rint((x - sc.origin[a]) * sc.inv_U * kTwo32);
    }
    q[a] = (unsigned long long)v;
  }
  const double rl = r / sc.lambda;
  // phase-1 position: the 64-bit value rounded to its high word
  o.x = (uint32_t)((q[0] + 0x80000000ull) >> 32);
  o.y = (uint32_t)((q[1] + 0x80000000ull) >> 32);
  o.z = (uint32_t)((q[2] + 0x80000000ull) >> 32);
  o.r = (float)rl;
  o.r3 = (float)(rl * rl * rl);
  o.pad[0] = o.pad[1] = o.pad[2] = 0u;
  c.xh = (uint32_t)(q[0] >> 32);
  c.yh = (uint32_t)(q[1] >> 32);
  c.zh = (uint32_t)(q[2] >> 32);
  c.xl = (uint32_t)q[0];
  c.yl = (uint32_t)q[1];
  c.zl = (uint32_t)q[2];
  if (f) {
    c.fx = (float)(f[0] / sc.Fs);
    c.fy = (float)(f[1] / sc.Fs);
    c.fz = (float)(f[2] / sc.Fs);
  } else {
    c.fx = c.fy = c.fz = 0.f;
  }
  c.pad[0] = c.pad[1] = c.pad[2] = 0u;
}

// Phase 1, one target against sources [j0, j1): FP32 sums over `tile`
// sources handed to float64 (blocked summation).  rec[j] is wave-uniform.
__device__ __forceinline__ void force_partial(const Rec* __restrict__ rec, uint32_t xi, uint32_t yi,
                                              uint32_t zi, float ri, int j0, int j1, int tile,
                                              double& ax, double& ay, double& az) {
  for (int jt = j0; jt < j1; jt += tile) {
    const int je = min(j1, jt + tile);
    float fx = 0.f, fy = 0.f, fz = 0.f;
#pragma unroll 4
    for (int j = jt; j < je; ++j) {
      const Rec sj = rec[j];  // wave-uniform: scalar loads
      // contact mask |d| > r_i + r_j also drops j == i
      dipole_pair<false>(diff32(sj.x, xi), diff32(sj.y, yi), diff32(sj.z, zi), ri, sj.r, sj.r3,
                         fx, fy, fz);
    }
    ax += (double)fx;
    ay += (double)fy;
    az += (double)fz;
  }
}

// Phase 2, one target (64-bit position xi, yi, zi) against sources [j0, j1):
// Oseen sum in m[], COMSOL lattice sum in c[] when CORR.
template <bool CORR>
__device__ __forceinline__ void conv_partial(const RecC* __restrict__ rec, unsigned long long xi,
                                             unsigned long long yi, unsigned long long zi,
                                             const CorrGrid& cg, int j0, int j1, int tile,
                                             double* m, double* c) {
  const unsigned long long nxi = 0ull - xi, nyi = 0ull - yi, nzi = 0ull - zi;
  for (int jt = j0; jt < j1; jt += tile) {
    const int je = min(j1, jt + tile);
    float vx = 0.f, vy = 0.f, vz = 0.f;
    float ux = 0.f, uy = 0.f, uz = 0.f;
#pragma unroll 4
    for (int j = jt; j < je; ++j) {
      const RecC sj = rec[j];
      const float dx = diff64n(((unsigned long long)sj.xh << 32) | sj.xl, nxi);
      const float dy = diff64n(((unsigned long long)sj.yh << 32) | sj.yl, nyi);
      const float dz = diff64n(((unsigned long long)sj.zh << 32) | sj.zl, nzi);
      const bool other = dx * dx + dy * dy + dz * dz > 0.f;  // drops j == i
      oseen_pair(dx, dy, dz, sj.fx, sj.fy, sj.fz, other, vx, vy, vz);
      if (CORR)
        corr_lookup(cg, dx, dy, dz, other ? sj.fx : 0.f, other ? sj.fy : 0.f,
                    other ? sj.fz : 0.f, ux, uy, uz);
    }
    m[0] += (double)vx;
    m[1] += (double)vy;
    m[2] += (double)vz;
    if (CORR) {
      c[0] += (double)ux;
      c[1] += (double)uy;
      c[2] += (double)uz;
    }
  }
}

// Force on one droplet from its summed partials s[] (units of the phase-1
// accumulator); also returns the force in units of Fs for the phase-2 record.
__device__ __forceinline__ void force_finish(const double* s, double radius, const DevScales& sc,
                                             double m_const, double* F, float* f_rec) {
#pragma clang fp contract(off)
  const double rl = radius / sc.lambda;
  const double k = m_const * sc.lambda * sc.lambda * kSigma3 * (rl * rl * rl);
  for (int a = 0; a < 3; ++a) {
    F[a] = k * s[a];
    f_rec[a] = (float)(F[a] / sc.Fs);
  }
}

struct CorrScalars {
  int enabled;
  double common;  // Fz_inv * eta_ratio * L_ratio
  double self_coeff;
};

// Velocity of one droplet from its summed Oseen partials m[] and lattice
// partials c[] (used when enabled), with the lattice self term g_self F_i.
__device__ __forceinline__ void conv_finish(const double* m, const double* c, const DevScales& sc,
                                            double eta_const, const CorrScalars& cs,
                                            const double* Fi, double* V) {
#pragma clang fp contract(off)
  const double km = eta_const * sc.Fs / sc.U;
  const double kc = sc.Fs * cs.common;
  for (int a = 0; a < 3; ++a) {
    double v = km * m[a];
    if (cs.enabled) v += kc * c[a] + cs.self_coeff * Fi[a];
    V[a] = v;
  }
}

}  // namespace dem
