// Hand-written adjoint of one direct DEM Euler step, and the near-contact
// loss used to differentiate trajectories.
//
// Forward (reference dem/force_calculator/direct_droplet_force_calculator.py
// :183-253, :354-411 and dem/solver/euler_droplet_solver.py:113-131):
//   F_i = m sum_{j != i, |d_ij| > r_i + r_j} r_i^3 r_j^3 f(d_ij),
//         f(d) = g(d)/|d|^7, g = (x(4z^2-x^2-y^2), y(4z^2-x^2-y^2), z(2z^2-3x^2-3y^2)),
//         d_ij = minimum image of x_j - x_i;
//   V_i = eta sum_{j != i} G(x_i - x_j) F_j,  G(x) = I/|x| + x x^T/|x|^3;
//   v_i = F_i/(sf r_i) + V_i,   x' = x + dt v   (mod L).
// For an adjoint vector a (N x 3) and J = sum_i a_i . v_i this file computes
//   grad_x J  (the vector-Jacobian product (dv/dx)^T a), and
//   S_E = sum_j c_j . F_j,  S_eta = sum_j F_j . u_j,  S_sf = sum_i b_i . F_i,
// from which dJ/dm = S_E/m, dJ/deta = S_eta/eta, dJ/dsf = -S_sf/sf, where
//   b_i = a_i/(sf r_i),  u_j = eta sum_{i != j} G(x_j - x_i) a_i  (the Oseen
//   sum with a in place of F),  c = b + u  (= dJ/dF);
//   grad_x J_i = sum_j w_ij J_f(d_ij) (c_j - c_i)                       (T1)
//              - eta sum_j [H(d_ij; a_i, F_j) + H(d_ij; a_j, F_i)]      (T2)
// with w_ij = m r_i^3 r_j^3 (zero in contact), J_f the symmetric Jacobian of
// f, and H(d; a, F) = grad_d (a^T G(d) F)
//   = [-(a.F) d + a (d.F) + F (a.d)]/|d|^3 - 3 (a.d)(d.F) d/|d|^5.
// The contact mask is treated as locally constant (its derivative is zero
// wherever it does not switch).
//
// MI355X design: the same fixed-point positions and work units as the
// forward kernels (direct_core.h); one target per lane, wave-uniform 64-byte
// source records, FP32 pair arithmetic summed in FP32 tiles handed to
// float64, j range split over blockIdx.y with deterministic float64
// reductions.  Adjoint, force and c vectors are stored divided by
// power-of-two scales (As, Fs, Cs) so their FP32 images stay in range.
#include <math.h>

#include <algorithm>

#include "dem_common.h"
#include "direct_core.h"
#include "profile.h"

namespace {

using dem::DevScales;
using dem::Rec;
using dem::RecC;

constexpr int WG = 256;
constexpr int kTile = 16;
constexpr int kSumBlocks = 256;

int split_for(int n) {  // same heuristic as the forward kernels
  const int iblocks = (n + WG - 1) / WG;
  long long s = (2048 + iblocks - 1) / iblocks;
  long long cap = (1ll << 21) / (n > 0 ? n : 1);
  long long s2 = (n + 63) / 64;
  if (cap < s2) s2 = cap;
  if (s2 > 64) s2 = 64;
  if (s > s2) s = s2;
  return (int)(s < 1 ? 1 : s);
}

// Source / target record of the pair VJP (64 bytes).
struct alignas(16) AdjRec {
  uint32_t x, y, z;  // 32-bit fixed-point position (rounded, as the force kernel)
  float r, r3;       // radius and radius^3 in work units
  float ax, ay, az;  // a / As
  float fx, fy, fz;  // F / Fs
  float cx, cy, cz;  // c / Cs
  float pad[2];
};
static_assert(sizeof(AdjRec) == 64, "record layout");

struct AdjScalars {
  DevScales sc;  // positions, lambda, Fs = m lambda^2
  double As;     // scale of a (power of two)
  double Cs;     // scale of c (power of two)
};

struct Layout {
  AdjScalars* s;
  unsigned long long* amax;  // [2] bit patterns of max|a|, max|c|
  Rec* rec;
  RecC* recc;
  AdjRec* arec;
  double* c;      // [n][3]
  double* part;   // [split][6][n]
  double* sums;   // [kSumBlocks][3]
  size_t bytes;
};

Layout carve(void* ws, int n) {
  dem::Carver c(ws);
  Layout l;
  const size_t nn = (size_t)(n > 0 ? n : 1);
  l.s = c.take<AdjScalars>(1);
  l.amax = c.take<unsigned long long>(2);
  l.rec = c.take<Rec>(nn);
  l.recc = c.take<RecC>(nn);
  l.arec = c.take<AdjRec>(nn);
  l.c = c.take<double>(3 * nn);
  l.part = c.take<double>((size_t)split_for(n) * 6 * nn);
  l.sums = c.take<double>(kSumBlocks * 3);
  l.bytes = dem::align_up(c.used, 256);
  return l;
}

__device__ __forceinline__ double pow2_at_least(unsigned long long bits) {
  const double v = __longlong_as_double((long long)bits);
  if (!(v > 0.0) || !isfinite(v)) return 1.0;
  int e;
  frexp(v, &e);  // v = f 2^e, f in [0.5, 1)
  return ldexp(1.0, e);
}

__global__ void k_adj_zero(unsigned long long* amax) {
  if (threadIdx.x < 2) amax[threadIdx.x] = 0ull;
}

// max |component| of a row-major (n,3) float64 array into *out (bit pattern of
// a non-negative double: unsigned order equals numeric order).
__global__ __launch_bounds__(256) void k_adj_absmax(const double* __restrict__ v, int n3,
                                                    unsigned long long* __restrict__ out) {
  double m = 0.0;
  for (int k = blockIdx.x * 256 + threadIdx.x; k < n3; k += gridDim.x * 256) {
    const double a = fabs(v[k]);
    if (a > m) m = a;  // NaN never wins
  }
  for (int d = 32; d > 0; d >>= 1) m = fmax(m, __shfl_down(m, d, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

__global__ void k_adj_setup(double L, double m_const, const unsigned long long* __restrict__ amax,
                            AdjScalars* __restrict__ s) {
  if (threadIdx.x != 0) return;
  AdjScalars o;
  o.sc = dem::periodic_scales(L, m_const);
  o.As = pow2_at_least(amax[0]);
  o.Cs = 1.0;
  *s = o;
}

// Records of the Oseen apply (64-bit positions, sources a / As) and the
// force-style 32-bit positions.
__global__ __launch_bounds__(256) void k_adj_pack(const double* __restrict__ pos,
                                                  const double* __restrict__ radii,
                                                  const double* __restrict__ a, int n,
                                                  const AdjScalars* __restrict__ sp,
                                                  Rec* __restrict__ rec, RecC* __restrict__ recc) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  DevScales sc = sp->sc;
  sc.Fs = sp->As;  // pack_one divides the source vector by sc.Fs
  Rec o;
  RecC c;
  dem::pack_one(pos + 3 * (size_t)i, radii[i], a ? a + 3 * (size_t)i : nullptr, sc, o, c);
  rec[i] = o;
  recc[i] = c;
}

// u = eta sum_i G(x_j - x_i) a_i: the forward Oseen tile with sources a / As.
__global__ __launch_bounds__(WG) void k_adj_oseen_tile(const RecC* __restrict__ rec, int n,
                                                       int jchunk, double* __restrict__ part) {
  const int i = blockIdx.x * WG + threadIdx.x;
  const int ii = i < n ? i : n - 1;
  const int s = blockIdx.y;
  const int j0 = s * jchunk;
  const int j1 = min(n, j0 + jchunk);
  const unsigned long long xi = ((unsigned long long)rec[ii].xh << 32) | rec[ii].xl;
  const unsigned long long yi = ((unsigned long long)rec[ii].yh << 32) | rec[ii].yl;
  const unsigned long long zi = ((unsigned long long)rec[ii].zh << 32) | rec[ii].zl;
  const dem::CorrGrid cg{};
  double m[3] = {0.0, 0.0, 0.0}, c[3] = {0.0, 0.0, 0.0};
  dem::conv_partial<false>(rec, xi, yi, zi, cg, j0, j1, kTile, m, c);
  if (i < n)
    for (int a = 0; a < 3; ++a) part[(size_t)(3 * s + a) * n + i] = m[a];
}

// c = b + u, the per-block parameter sums and max|c|.
__global__ __launch_bounds__(256) void k_adj_c(const double* __restrict__ part, int S, int n,
                                               const double* __restrict__ a,
                                               const double* __restrict__ F,
                                               const double* __restrict__ radii, double eta,
                                               double sf, const AdjScalars* __restrict__ sp,
                                               double* __restrict__ c,
                                               double* __restrict__ sums,
                                               unsigned long long* __restrict__ amax) {
  __shared__ double sh[3][256];
  const int i = blockIdx.x * 256 + threadIdx.x;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, cm = 0.0;
  if (i < n) {
    double m[3] = {0.0, 0.0, 0.0};
    for (int s = 0; s < S; ++s)
      for (int k = 0; k < 3; ++k) m[k] += part[(size_t)(3 * s + k) * n + i];
    DevScales sc = sp->sc;
    sc.Fs = sp->As;
    const dem::CorrScalars cs{};
    const double zero[3] = {0.0, 0.0, 0.0};
    double u[3];
    dem::conv_finish(m, zero, sc, eta, cs, zero, u);
    const double inv = 1.0 / (sf * radii[i]);
    for (int k = 0; k < 3; ++k) {
      const double b = a[3 * (size_t)i + k] * inv;
      const double f = F[3 * (size_t)i + k];
      const double ck = b + u[k];
      c[3 * (size_t)i + k] = ck;
      s0 += ck * f;
      s1 += f * u[k];
      s2 += b * f;
      cm = fmax(cm, fabs(ck));
    }
  }
  sh[0][threadIdx.x] = s0;
  sh[1][threadIdx.x] = s1;
  sh[2][threadIdx.x] = s2;
  for (int d = 32; d > 0; d >>= 1) cm = fmax(cm, __shfl_down(cm, d, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(&amax[1], (unsigned long long)__double_as_longlong(cm));
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 3; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) sums[3 * blockIdx.x + k] = sh[k][0];
}

__global__ void k_adj_sums(const double* __restrict__ part, int nblk, double* __restrict__ out) {
  if (threadIdx.x != 0) return;
  double s[3] = {0.0, 0.0, 0.0};
  for (int b = 0; b < nblk; ++b)
    for (int k = 0; k < 3; ++k) s[k] += part[3 * b + k];
  for (int k = 0; k < 3; ++k) out[k] = s[k];
}

__global__ __launch_bounds__(256) void k_adj_pack2(const Rec* __restrict__ rec,
                                                   const double* __restrict__ a,
                                                   const double* __restrict__ F,
                                                   const double* __restrict__ c, int n,
                                                   const AdjScalars* __restrict__ sp,
                                                   const unsigned long long* __restrict__ amax,
                                                   AdjRec* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double As = sp->As, Fs = sp->sc.Fs, Cs = pow2_at_least(amax[1]);
  const Rec r = rec[i];
  AdjRec o;
  o.x = r.x;
  o.y = r.y;
  o.z = r.z;
  o.r = r.r;
  o.r3 = r.r3;
  o.ax = (float)(a[3 * (size_t)i] / As);
  o.ay = (float)(a[3 * (size_t)i + 1] / As);
  o.az = (float)(a[3 * (size_t)i + 2] / As);
  o.fx = (float)(F[3 * (size_t)i] / Fs);
  o.fy = (float)(F[3 * (size_t)i + 1] / Fs);
  o.fz = (float)(F[3 * (size_t)i + 2] / Fs);
  o.cx = (float)(c[3 * (size_t)i] / Cs);
  o.cy = (float)(c[3 * (size_t)i + 1] / Cs);
  o.cz = (float)(c[3 * (size_t)i + 2] / Cs);
  o.pad[0] = o.pad[1] = 0.f;
  out[i] = o;
}

// Pair VJP: T1 (masked dipole Jacobian on c_j - c_i) and T2 (Oseen gradient)
// of target i over sources [j0, j1), in work units.
__global__ __launch_bounds__(WG) void k_adj_pair_tile(const AdjRec* __restrict__ rec, int n,
                                                      int jchunk, double* __restrict__ part) {
  const int i = blockIdx.x * WG + threadIdx.x;
  const int ii = i < n ? i : n - 1;
  const int s = blockIdx.y;
  const int j0 = s * jchunk;
  const int j1 = min(n, j0 + jchunk);
  const AdjRec me = rec[ii];
  double t1[3] = {0.0, 0.0, 0.0}, t2[3] = {0.0, 0.0, 0.0};
  for (int jt = j0; jt < j1; jt += kTile) {
    const int je = min(j1, jt + kTile);
    float p1x = 0.f, p1y = 0.f, p1z = 0.f, p2x = 0.f, p2y = 0.f, p2z = 0.f;
    for (int j = jt; j < je; ++j) {
      const AdjRec sj = rec[j];  // wave-uniform
      const float fxd = dem::diff32(sj.x, me.x), fyd = dem::diff32(sj.y, me.y),
                  fzd = dem::diff32(sj.z, me.z);
      const float r2 = dem::r2_work(fxd, fyd, fzd);
      const bool other = r2 > 0.f;
      const bool sep = dem::separated<false>(r2, me.r + sj.r);  // force mask
      const float dx = fxd * 0x1p-22f, dy = fyd * 0x1p-22f, dz = fzd * 0x1p-22f;
      const float ir = other ? __builtin_amdgcn_rsqf(r2) : 0.f;
      const float ir2 = ir * ir;
      // T1: |D|^-7 [Jg(D) dc - 7 g(D) (D.dc)/|D|^2], dc = c_j - c_i
      const float cx = sj.cx - me.cx, cy = sj.cy - me.cy, cz = sj.cz - me.cz;
      const float q = dx * dx + dy * dy;
      const float zz = dz * dz;
      const float A = 4.f * zz - q;
      const float B = 2.f * zz - 3.f * q;
      const float dd = dx * cx + dy * cy + dz * cz;
      const float jx = (A - 2.f * dx * dx) * cx - 2.f * dx * dy * cy + 8.f * dx * dz * cz;
      const float jy = -2.f * dx * dy * cx + (A - 2.f * dy * dy) * cy + 8.f * dy * dz * cz;
      const float jz = -6.f * dx * dz * cx - 6.f * dy * dz * cy + (6.f * zz - 3.f * q) * cz;
      const float ir7 = (ir2 * ir2) * (ir2 * ir);
      const float w1 = sep ? sj.r3 * ir7 : 0.f;
      const float k7 = 7.f * dd * ir2;
      p1x += w1 * (jx - k7 * dx * A);
      p1y += w1 * (jy - k7 * dy * A);
      p1z += w1 * (jz - k7 * dz * B);
      // T2: H(D; a_i, F_j) + H(D; a_j, F_i)
      const float adi = me.ax * dx + me.ay * dy + me.az * dz;
      const float adj = sj.ax * dx + sj.ay * dy + sj.az * dz;
      const float fdj = sj.fx * dx + sj.fy * dy + sj.fz * dz;
      const float fdi = me.fx * dx + me.fy * dy + me.fz * dz;
      const float af = me.ax * sj.fx + me.ay * sj.fy + me.az * sj.fz + sj.ax * me.fx +
                       sj.ay * me.fy + sj.az * me.fz;
      const float ir3 = ir2 * ir;
      const float k5 = 3.f * (adi * fdj + adj * fdi) * ir2;
      p2x += ir3 * (-af * dx + me.ax * fdj + sj.fx * adi + sj.ax * fdi + me.fx * adj - k5 * dx);
      p2y += ir3 * (-af * dy + me.ay * fdj + sj.fy * adi + sj.ay * fdi + me.fy * adj - k5 * dy);
      p2z += ir3 * (-af * dz + me.az * fdj + sj.fz * adi + sj.az * fdi + me.fz * adj - k5 * dz);
    }
    t1[0] += (double)p1x;
    t1[1] += (double)p1y;
    t1[2] += (double)p1z;
    t2[0] += (double)p2x;
    t2[1] += (double)p2y;
    t2[2] += (double)p2z;
  }
  if (i < n)
    for (int a = 0; a < 3; ++a) {
      part[(size_t)(6 * s + a) * n + i] = t1[a];
      part[(size_t)(6 * s + 3 + a) * n + i] = t2[a];
    }
}

// grad_x J_i = m lambda Cs rho_i^3 T1_i - eta As Fs / lambda^2 T2_i.
__global__ __launch_bounds__(256) void k_adj_reduce(const double* __restrict__ part, int S, int n,
                                                    const double* __restrict__ radii,
                                                    double m_const, double eta,
                                                    const AdjScalars* __restrict__ sp,
                                                    const unsigned long long* __restrict__ amax,
                                                    double* __restrict__ gx) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double t1[3] = {0.0, 0.0, 0.0}, t2[3] = {0.0, 0.0, 0.0};
  for (int s = 0; s < S; ++s)
    for (int a = 0; a < 3; ++a) {
      t1[a] += part[(size_t)(6 * s + a) * n + i];
      t2[a] += part[(size_t)(6 * s + 3 + a) * n + i];
    }
  const AdjScalars s = *sp;
  const double lam = s.sc.lambda;
  const double rho = radii[i] / lam;
  const double k1 = m_const * lam * pow2_at_least(amax[1]) * (rho * rho * rho);
  const double k2 = eta * s.As * s.sc.Fs / (lam * lam);
  for (int a = 0; a < 3; ++a) gx[3 * (size_t)i + a] = k1 * t1[a] - k2 * t2[a];
}

// ---------------------------------------------------------------- contact loss

// loss_i = 1/2 sum_{j != i} softplus(unit (factor (r_i + r_j) - |d_ij|)) and
// grad_i = unit sum_{j != i} sigmoid(s_ij) d_ij/|d_ij|, in work units.
__global__ __launch_bounds__(WG) void k_loss_tile(const Rec* __restrict__ rec, int n, int jchunk,
                                                  float factor, float k_unit,
                                                  double* __restrict__ part) {
  const int i = blockIdx.x * WG + threadIdx.x;
  const int ii = i < n ? i : n - 1;
  const int s = blockIdx.y;
  const int j0 = s * jchunk;
  const int j1 = min(n, j0 + jchunk);
  const Rec me = rec[ii];
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int jt = j0; jt < j1; jt += kTile) {
    const int je = min(j1, jt + kTile);
    float l = 0.f, gxs = 0.f, gys = 0.f, gzs = 0.f;
    for (int j = jt; j < je; ++j) {
      const Rec sj = rec[j];
      const float dx = dem::diff32(sj.x, me.x) * 0x1p-22f;
      const float dy = dem::diff32(sj.y, me.y) * 0x1p-22f;
      const float dz = dem::diff32(sj.z, me.z) * 0x1p-22f;
      const float r2 = dx * dx + dy * dy + dz * dz;
      const bool other = r2 > 0.f;
      const float ir = other ? __builtin_amdgcn_rsqf(r2) : 0.f;
      const float dist = r2 * ir;
      const float sarg = k_unit * (factor * (me.r + sj.r) - dist);
      const float e = __expf(-fabsf(sarg));
      const float sp = fmaxf(sarg, 0.f) + log1pf(e);
      const float sg = sarg >= 0.f ? 1.f / (1.f + e) : e / (1.f + e);
      l += other ? sp : 0.f;
      const float w = other ? sg * ir : 0.f;
      gxs += w * dx;
      gys += w * dy;
      gzs += w * dz;
    }
    acc[0] += (double)l;
    acc[1] += (double)gxs;
    acc[2] += (double)gys;
    acc[3] += (double)gzs;
  }
  if (i < n)
    for (int a = 0; a < 4; ++a) part[(size_t)(4 * s + a) * n + i] = acc[a];
}

__global__ __launch_bounds__(256) void k_loss_reduce(const double* __restrict__ part, int S, int n,
                                                     double unit, double* __restrict__ grad,
                                                     double* __restrict__ bsum) {
  __shared__ double sh[256];
  const int i = blockIdx.x * 256 + threadIdx.x;
  double l = 0.0;
  if (i < n) {
    double g[3] = {0.0, 0.0, 0.0};
    for (int s = 0; s < S; ++s) {
      l += part[(size_t)(4 * s) * n + i];
      for (int a = 0; a < 3; ++a) g[a] += part[(size_t)(4 * s + 1 + a) * n + i];
    }
    for (int a = 0; a < 3; ++a) grad[3 * (size_t)i + a] = unit * g[a];
  }
  sh[threadIdx.x] = 0.5 * l;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = sh[0];
}

__global__ void k_loss_total(const double* __restrict__ bsum, int nblk, double* __restrict__ out) {
  if (threadIdx.x != 0) return;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += bsum[b];
  *out = s;
}

int check_args(const double* pos, const double* radii, int n, const dem_phys* ph) {
  if (n < 0) return dem::set_error(DEM_ERR_INVALID, "n must be >= 0 (got %d)", n);
  if (!ph) return dem::set_error(DEM_ERR_INVALID, "phys is NULL");
  if (!ph->periodic)
    return dem::set_error(DEM_ERR_INVALID, "the adjoint supports periodic boxes only");
  if (!(ph->L > 0.0)) return dem::set_error(DEM_ERR_INVALID, "periodic box needs L > 0");
  if (n > 0 && (!pos || !radii)) return dem::set_error(DEM_ERR_INVALID, "null positions/radii");
  if ((n + 255) / 256 > kSumBlocks)
    return dem::set_error(DEM_ERR_CAPACITY, "n too large for the adjoint (%d > %d)", n,
                          256 * kSumBlocks);
  return DEM_OK;
}

}  // namespace

extern "C" {

size_t dem_adjoint_workspace_bytes(int n) { return carve(nullptr, n < 0 ? 0 : n).bytes; }

int dem_direct_step_vjp(const double* pos, const double* radii, const double* forces,
                        const double* adj, int n, const dem_phys* phys, double stokes_factor,
                        double* grad_x, double* sums, void* ws, size_t ws_bytes,
                        dem_stream_t stream) {
  int rc = check_args(pos, radii, n, phys);
  if (rc) return rc;
  if (n > 0 && (!forces || !adj || !grad_x)) return dem::set_error(DEM_ERR_INVALID, "null array");
  if (!sums) return dem::set_error(DEM_ERR_INVALID, "sums is NULL");
  if (!(stokes_factor > 0.0)) return dem::set_error(DEM_ERR_INVALID, "stokes_factor must be > 0");
  Layout l = carve(ws, n);
  if (!ws || ws_bytes < l.bytes)
    return dem::set_error(DEM_ERR_WORKSPACE, "workspace %zu < %zu bytes", ws_bytes, l.bytes);
  hipStream_t st = dem::as_stream(stream);
  if (n == 0) {
    hipLaunchKernelGGL(k_adj_sums, dim3(1), dim3(64), 0, st, (const double*)l.sums, 0, sums);
    DEM_CHECK_LAUNCH("k_adj_sums");
    return DEM_OK;
  }
  const int nb = (n + 255) / 256;
  const int S = split_for(n);
  const int jchunk = (n + S - 1) / S;
  const dim3 grid((n + WG - 1) / WG, S);
  hipLaunchKernelGGL(k_adj_zero, dim3(1), dim3(64), 0, st, l.amax);
  hipLaunchKernelGGL(k_adj_absmax, dim3(std::min(nb * 3, 1024)), dim3(256), 0, st, adj, 3 * n,
                     l.amax);
  hipLaunchKernelGGL(k_adj_setup, dim3(1), dim3(64), 0, st, phys->L, phys->m_const,
                     (const unsigned long long*)l.amax, l.s);
  hipLaunchKernelGGL(k_adj_pack, dim3(nb), dim3(256), 0, st, pos, radii, adj, n,
                     (const AdjScalars*)l.s, l.rec, l.recc);
  DEM_CHECK_LAUNCH("k_adj_pack");
  DEM_PROF_LAUNCH("k_adj_oseen_tile", st,
                  hipLaunchKernelGGL(k_adj_oseen_tile, grid, dim3(WG), 0, st, (const RecC*)l.recc,
                                     n, jchunk, l.part));
  DEM_CHECK_LAUNCH("k_adj_oseen_tile");
  const int cb = nb;  // <= kSumBlocks (check_args)
  hipLaunchKernelGGL(k_adj_c, dim3(cb), dim3(256), 0, st, (const double*)l.part, S, n, adj,
                     forces, radii, phys->eta_const, stokes_factor, (const AdjScalars*)l.s, l.c,
                     l.sums, l.amax);
  DEM_CHECK_LAUNCH("k_adj_c");
  hipLaunchKernelGGL(k_adj_sums, dim3(1), dim3(64), 0, st, (const double*)l.sums, cb, sums);
  hipLaunchKernelGGL(k_adj_pack2, dim3(nb), dim3(256), 0, st, (const Rec*)l.rec, adj, forces,
                     (const double*)l.c, n, (const AdjScalars*)l.s,
                     (const unsigned long long*)l.amax, l.arec);
  DEM_CHECK_LAUNCH("k_adj_pack2");
  DEM_PROF_LAUNCH("k_adj_pair_tile",
                  st, hipLaunchKernelGGL(k_adj_pair_tile, grid, dim3(WG), 0, st,
                                         (const AdjRec*)l.arec, n, jchunk, l.part));
  DEM_CHECK_LAUNCH("k_adj_pair_tile");
  hipLaunchKernelGGL(k_adj_reduce, dim3(nb), dim3(256), 0, st, (const double*)l.part, S, n, radii,
                     phys->m_const, phys->eta_const, (const AdjScalars*)l.s,
                     (const unsigned long long*)l.amax, grad_x);
  DEM_CHECK_LAUNCH("k_adj_reduce");
  return DEM_OK;
}

int dem_contact_loss(const double* pos, const double* radii, int n, const dem_phys* phys,
                     double factor, double unit, double* loss, double* grad, void* ws,
                     size_t ws_bytes, dem_stream_t stream) {
  int rc = check_args(pos, radii, n, phys);
  if (rc) return rc;
  if (!loss || (n > 0 && !grad)) return dem::set_error(DEM_ERR_INVALID, "null output");
  Layout l = carve(ws, n);
  if (!ws || ws_bytes < l.bytes)
    return dem::set_error(DEM_ERR_WORKSPACE, "workspace %zu < %zu bytes", ws_bytes, l.bytes);
  hipStream_t st = dem::as_stream(stream);
  const int nb = (n + 255) / 256;
  if (n == 0) {
    hipLaunchKernelGGL(k_loss_total, dim3(1), dim3(64), 0, st, (const double*)l.sums, 0, loss);
    DEM_CHECK_LAUNCH("k_loss_total");
    return DEM_OK;
  }
  const int S = split_for(n);
  const int jchunk = (n + S - 1) / S;
  hipLaunchKernelGGL(k_adj_zero, dim3(1), dim3(64), 0, st, l.amax);
  hipLaunchKernelGGL(k_adj_setup, dim3(1), dim3(64), 0, st, phys->L, 1.0,
                     (const unsigned long long*)l.amax, l.s);
  hipLaunchKernelGGL(k_adj_pack, dim3(nb), dim3(256), 0, st, pos, radii, (const double*)nullptr,
                     n, (const AdjScalars*)l.s, l.rec, l.recc);
  DEM_CHECK_LAUNCH("k_adj_pack");
  // k_unit converts work units to `unit` lengths: lambda = L / 2^10 metres
  const double lambda = phys->L / 1024.0;
  hipLaunchKernelGGL(k_loss_tile, dim3((n + WG - 1) / WG, S), dim3(WG), 0, st, (const Rec*)l.rec,
                     n, jchunk, (float)factor, (float)(unit * lambda), l.part);
  DEM_CHECK_LAUNCH("k_loss_tile");
  hipLaunchKernelGGL(k_loss_reduce, dim3(nb), dim3(256), 0, st, (const double*)l.part, S, n,
                     unit, grad, l.sums);
  DEM_CHECK_LAUNCH("k_loss_reduce");
  hipLaunchKernelGGL(k_loss_total, dim3(1), dim3(64), 0, st, (const double*)l.sums, nb, loss);
  DEM_CHECK_LAUNCH("k_loss_total");
  return DEM_OK;
}

}  // extern "C"
