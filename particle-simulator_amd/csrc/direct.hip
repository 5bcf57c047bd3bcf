// Direct O(N^2) dipole force + Oseen convection for gfx950.
//
// Reference semantics (dem/force_calculator/direct_droplet_force_calculator.py):
//   phase 1 (_calculate_forces, :202-253): for every target i, sum over j != i
//     with minimum-image d = x_j - x_i (:233-240) and contact mask
//     |d| > r_i + r_j (:244-249) of
//       F_i += m r_i^3 r_j^3 / |d|^7 * (dx(4dz^2-dx^2-dy^2), dy(..), dz(2dz^2-3dx^2-3dy^2))
//   phase 2 (_calculate_convection_velocities, :354-468): for every i, sum over
//     j != i (no contact mask) of the Oseen tensor applied to F_j,
//       v_i += eta (F_j/|x| + x (x.F_j)/|x|^3),  x = x_i - x_j (minimum image),
//     plus, when enabled, the COMSOL periodic-image correction (three permuted
//     trilinear lookups, :413-457) and the self term g_self F_i (:459-464).
//
// MI355X design:
//   * Positions are quantised to 32-bit fixed point.  Periodic boxes map [0,L)
//     onto the full uint32 ring, so (int32)(q_j - q_i) IS the minimum-image
//     difference (exact, one integer subtract); open boxes map the bounding box
//     onto [0, 2^31) so the same subtract never wraps.  Quantisation error is
//     L/2^33 (~1.6e-13 m for a 1.4 mm box).
//   * Arithmetic is FP32 VALU (the pair kernels are not contractions: no MFMA).
//     Work units are 2^22 fixed units (L/1024 for periodic boxes), which keeps
//     r^-7 and the cubic numerators inside the FP32 range for any droplet
//     size / box ratio up to ~1e5.
//   * One target per lane, 256 lanes per workgroup; the j (source) loop reads a
//     wave-uniform 32-byte record per source, which the compiler serves from
//     the scalar data cache (s_load) into SGPRs, so sources cost no VALU or LDS
//     traffic.  The j range is split over blockIdx.y to fill the 256 CUs; each
//     split writes a float64 partial that a deterministic reduction kernel sums.
//   * Inside a split the FP32 sum runs over `tile` sources and is then added to
//     a float64 accumulator (blocked summation: the FP32 rounding error grows
//     with the tile length, not with N).
#include <limits.h>
#include <math.h>

#include <algorithm>

#include "dem_common.h"
#include "direct_core.h"
#include "direct_pairs.h"
#include "direct_sym.h"
#include "image32.h"
#include "profile.h"

namespace {

using dem::DevScales;
using dem::ImgArgs;
using dem::img_correct;
using dem::img_probe;
using dem::rec_hi;
using dem::Rec;
using dem::RecC;

constexpr int WG = 256;

int g_split = 0;  // 0: heuristic
int g_tile = 0;   // 0: heuristic
int g_sym = 1;    // 0: one-sided only; 1: pair-once for kSymMin <= n <= kSymMax without an
                  // explicit split; 2: pair-once for every n, g_split giving the chunk count
int g_img32 = 1;  // mode 1: the pair-once Oseen sum on 32-bit separations (boundary pairs
                  // corrected); 0: on 64-bit separations as in mode 2

// Symmetric (pair-once) evaluation: blocks of kSymB droplets; block b pairs
// with the next sym_partners(b, B) blocks cyclically, which covers every
// unordered block pair exactly once.  Its float64 partials of the reversed
// pairs take B * B/2 * 256 * 3 doubles, so it runs for mid-size N only.
using dem::kSymB;
using dem::kSymTplConv;
using dem::kSymTplForce;
using dem::sym_partners;
constexpr int kSymMin = 10240;  // below, the one-sided kernels are faster
constexpr int kSymMax = 131072;

inline int sym_blocks(int n) { return (n + kSymB - 1) / kSymB; }
inline bool sym_fits(int n) { return n >= kSymMin && n <= kSymMax; }
// An explicit split (dem_direct_set_tuning) selects the one-sided kernels.
inline bool use_sym(int n) {
  if (g_sym == 2) return n >= 1;
  return g_sym == 1 && g_split == 0 && sym_fits(n);
}
// Workspaces hold the pair-once partials whenever the kernels may run.
inline bool sym_alloc(int n) { return (g_sym == 2 && n >= 1) || sym_fits(n); }
constexpr int kSymMaxChunks = 64;
// grid.y for `slots` co-resident workgroups: one round of B x S workgroups,
// at most one chunk per partner entry.
inline int sym_chunks(int n, int slots) {
  const int B = sym_blocks(n);
  int S = slots / B;
  const int E = 1 + B / 2;
  if (S > E) S = E;
  if (S > kSymMaxChunks) S = kSymMaxChunks;
  return S < 1 ? 1 : S;
}

inline int split_max(int n) {
  long long cap = (1ll << 21) / (n > 0 ? n : 1);
  long long s = (n + 63) / 64;
  if (cap < s) s = cap;
  if (s > 64) s = 64;
  if (s < 1) s = 1;
  return (int)s;
}

inline int choose_split(int n) {
  int iblocks = (n + WG - 1) / WG;
  int s = g_split > 0 ? g_split : (2048 + iblocks - 1) / iblocks;
  int smax = split_max(n);
  if (s > smax) s = smax;
  if (s < 1) s = 1;
  return s;
}

// FP32 partial-sum length before the float64 hand-off.  The Oseen sum is far
// worse conditioned than the force sum (1/r decay: sum|t|/|v| ~ 1e3-1e5 at
// N=65536), so it hands off four times more often.
inline int choose_tile_force() { return g_tile > 0 ? g_tile : 32; }
inline int choose_tile_conv() { return g_tile > 0 ? g_tile : 16; }

constexpr int kBBoxBlocks = 256;

struct Layout {
  DevScales* sc;
  double* bbox;  // [kBBoxBlocks][6]
  Rec* rec;      // [npad]: records beyond n are zero (no radius, no force)
  RecC* recc;
  double* part;  // [split][6][n]
  dem::PartJ* part_j;  // symmetric kernels: [B][B/2][3][kSymB]
  // boundary pairs of the 32-bit Oseen separations (pair-once, periodic)
  unsigned long long* img_slot;  // [3][1 << img_log2m] high-word hash tables
  int* img_partner;              // [n][3]: partner across half the ring, -1 none
  int* img_flag;                 // nonzero: a duplicated high word has a partner
  int img_log2m;
  int npad;
  size_t bytes;
};

// Slots per axis of the boundary-pair tables: a power of two >= 4 n.
inline int img_log2m(int n) {
  int m = 10;
  while (m < 30 && (1ll << m) < 4ll * n) ++m;
  return m;
}

Layout carve(void* ws, int n) {
  dem::Carver c(ws);
  Layout l;
  const int nn = n > 0 ? n : 1;
  l.npad = sym_alloc(n) ? sym_blocks(n) * kSymB : nn;
  l.sc = c.take<DevScales>(1);
  l.bbox = c.take<double>(kBBoxBlocks * 6);
  l.rec = c.take<Rec>((size_t)l.npad);
  l.recc = c.take<RecC>((size_t)l.npad);
  size_t parts = (size_t)split_max(n) * 6 * (size_t)nn;
  if (sym_alloc(n)) {
    const size_t ps = (size_t)kSymMaxChunks * 3 * (size_t)nn;
    if (ps > parts) parts = ps;
  }
  l.part = c.take<double>(parts);
  if (sym_alloc(n)) {
    const size_t B = (size_t)sym_blocks(n);
    l.part_j = c.take<dem::PartJ>(std::max<size_t>(B * (B / 2), 1) * 3 * kSymB);
  } else {
    l.part_j = nullptr;
  }
  l.img_log2m = img_log2m(nn);
  if (sym_fits(n)) {
    l.img_slot = c.take<unsigned long long>((size_t)3 << l.img_log2m);
    l.img_partner = c.take<int>((size_t)3 * nn);
    l.img_flag = c.take<int>(1);
  } else {
    l.img_slot = nullptr;
    l.img_partner = nullptr;
    l.img_flag = nullptr;
  }
  l.bytes = dem::align_up(c.used, 256);
  return l;
}

// ---------------------------------------------------------------- set-up

__global__ __launch_bounds__(256) void k_bbox_partial(const double* __restrict__ pos, int n,
                                                      double* __restrict__ part) {
  __shared__ double s_lo[3][256], s_hi[3][256];
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    for (int a = 0; a < 3; ++a) {
      double v = pos[3 * (size_t)i + a];
      lo[a] = fmin(lo[a], v);
      hi[a] = fmax(hi[a], v);
    }
  }
  for (int a = 0; a < 3; ++a) {
    s_lo[a][threadIdx.x] = lo[a];
    s_hi[a][threadIdx.x] = hi[a];
  }
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int a = 0; a < 3; ++a) {
        s_lo[a][threadIdx.x] = fmin(s_lo[a][threadIdx.x], s_lo[a][threadIdx.x + w]);
        s_hi[a][threadIdx.x] = fmax(s_hi[a][threadIdx.x], s_hi[a][threadIdx.x + w]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int a = 0; a < 3; ++a) {
      part[6 * blockIdx.x + a] = s_lo[a][0];
      part[6 * blockIdx.x + 3 + a] = s_hi[a][0];
    }
}

// Single thread: fold the bounding box and derive the per-call scales.
__global__ void k_setup(const double* __restrict__ bbox, int nblk, double L, int periodic,
                        double m_const, DevScales* __restrict__ sc) {
  if (threadIdx.x != 0) return;
  if (periodic) {
    *sc = dem::periodic_scales(L, m_const);
    return;
  }
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int b = 0; b < nblk; ++b)
    for (int a = 0; a < 3; ++a) {
      lo[a] = fmin(lo[a], bbox[6 * b + a]);
      hi[a] = fmax(hi[a], bbox[6 * b + 3 + a]);
    }
  double ext = 0.0;
  for (int a = 0; a < 3; ++a) ext = fmax(ext, hi[a] - lo[a]);
  *sc = dem::open_scales(lo, ext, m_const);
}

// Quantise positions, scale radii, and (optionally) load caller forces as
// sources; clear the boundary-pair tables (img_slot, img_words words) and flag
// when the call uses them.
__global__ __launch_bounds__(256) void k_pack(const double* __restrict__ pos,
                                              const double* __restrict__ radii,
                                              const double* __restrict__ forces, int n, int npad,
                                              const DevScales* __restrict__ scp,
                                              Rec* __restrict__ rec, RecC* __restrict__ recc,
                                              unsigned long long* __restrict__ img_slot,
                                              size_t img_words, int* __restrict__ img_flag) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (img_slot) {
    for (size_t u = i; u < img_words; u += (size_t)gridDim.x * 256) img_slot[u] = 0ull;
    if (i == 0) *img_flag = 0;
  }
  if (i >= n) {
    if (i < npad) {  // padding: no radius, no force, at the origin
      rec[i] = Rec{};
      recc[i] = RecC{};
    }
    return;
  }
  const DevScales sc = *scp;
  Rec o;
  RecC c;
  dem::pack_one(pos + 3 * (size_t)i, radii[i], forces ? forces + 3 * (size_t)i : nullptr, sc, o, c);
  rec[i] = o;
  recc[i] = c;
}

// ---------------------------------------------------------------- phase 1

__global__ __launch_bounds__(WG) void k_force_tile(const Rec* __restrict__ rec, int n, int jchunk,
                                                   int tile, double* __restrict__ part) {
  const int i = blockIdx.x * WG + threadIdx.x;
  const int ii = i < n ? i : n - 1;
  const int s = blockIdx.y;
  const int j0 = s * jchunk;
  const int j1 = min(n, j0 + jchunk);
  double ax = 0.0, ay = 0.0, az = 0.0;
  dem::force_partial(rec, rec[ii].x, rec[ii].y, rec[ii].z, rec[ii].r, j0, j1, tile, ax, ay, az);
  if (i < n) {
    part[(size_t)(6 * s + 0) * n + i] = ax;
    part[(size_t)(6 * s + 1) * n + i] = ay;
    part[(size_t)(6 * s + 2) * n + i] = az;
  }
}

__global__ __launch_bounds__(256) void k_force_reduce(const double* __restrict__ part, int S, int n,
                                                      const double* __restrict__ radii,
                                                      const DevScales* __restrict__ scp,
                                                      double m_const, double* __restrict__ F,
                                                      RecC* __restrict__ recc) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double sum[3] = {0.0, 0.0, 0.0};
  for (int s = 0; s < S; ++s)
    for (int a = 0; a < 3; ++a) sum[a] += part[(size_t)(6 * s + a) * n + i];
  const DevScales sc = *scp;
  double Fi[3];
  float fr[3];
  dem::force_finish(sum, radii[i], sc, m_const, Fi, fr);
  for (int a = 0; a < 3; ++a) F[3 * (size_t)i + a] = Fi[a];
  recc[i].fx = fr[0];
  recc[i].fy = fr[1];
  recc[i].fz = fr[2];
}

// ---------------------------------------------------------------- phase 1, pair-once

// Four lanes per target (the reversed partials are most of the bytes); with
// boundary-pair tables, lanes 1..3 then look up the target's partner across
// half the ring on axes 0..2.
__global__ __launch_bounds__(256) void k_force_sym_reduce(
    const double* __restrict__ part, int S, int n, const dem::PartJ* __restrict__ part_j, int B,
    const double* __restrict__ radii, const DevScales* __restrict__ scp, double m_const,
    double* __restrict__ F, RecC* __restrict__ recc, ImgArgs img) {
  const int i = blockIdx.x * 64 + (threadIdx.x >> 2);
  if (i >= n) return;  // the four lanes of a target leave together
  double sum[3] = {0.0, 0.0, 0.0};
  dem::sym_sum4(part, S, n, part_j, B, i, sum);
  const int k = threadIdx.x & 3;
  if (k) {
    if (img.slot) {
      const int a = k - 1;
      const int j = img_probe(img.slot + ((size_t)a << img.log2m), img.log2m,
                              rec_hi(recc[i], a));
      img.partner[3 * (size_t)i + a] = j < 0 ? -1 : j;
      if (j == -2) atomicOr(img.flag, 1);
    }
    return;
  }
  const DevScales sc = *scp;
  double Fi[3];
  float fr[3];
  dem::force_finish(sum, radii[i], sc, m_const, Fi, fr);
  for (int a = 0; a < 3; ++a) F[3 * (size_t)i + a] = Fi[a];
  recc[i].fx = fr[0];
  recc[i].fy = fr[1];
  recc[i].fz = fr[2];
}

// ---------------------------------------------------------------- phase 2

struct CorrArgs {
  const float4* grid;  // (u, v, w, 0) per node
  int n;
  float nmax;          // n - 1
  double gmin, inv_dx, L_ratio;
};

template <bool CORR>
__global__ __launch_bounds__(WG) void k_conv_tile(const RecC* __restrict__ rec, int n, int jchunk,
                                                  int tile, const DevScales* __restrict__ scp,
                                                  CorrArgs ca, double* __restrict__ part) {
  const int i = blockIdx.x * WG + threadIdx.x;
  const int ii = i < n ? i : n - 1;
  const int s = blockIdx.y;
  const int j0 = s * jchunk;
  const int j1 = min(n, j0 + jchunk);
  const unsigned long long xi = ((unsigned long long)rec[ii].xh << 32) | rec[ii].xl;
  const unsigned long long yi = ((unsigned long long)rec[ii].yh << 32) | rec[ii].yl;
  const unsigned long long zi = ((unsigned long long)rec[ii].zh << 32) | rec[ii].zl;
  dem::CorrGrid cg{};
  if (CORR) {
    cg.grid = ca.grid;
    cg.n = ca.n;
    cg.nmax = ca.nmax;
    cg.ga = (float)(-scp->U * ca.L_ratio * ca.inv_dx);
    cg.gb = (float)(-ca.gmin * ca.inv_dx);
  }
  double m[3] = {0.0, 0.0, 0.0}, c[3] = {0.0, 0.0, 0.0};
  dem::conv_partial<CORR>(rec, xi, yi, zi, cg, j0, j1, tile, m, c);
  if (i < n) {
    for (int a = 0; a < 3; ++a) part[(size_t)(6 * s + a) * n + i] = m[a];
    if (CORR)
      for (int a = 0; a < 3; ++a) part[(size_t)(6 * s + 3 + a) * n + i] = c[a];
  }
}

__global__ __launch_bounds__(256) void k_conv_reduce(const double* __restrict__ part, int S, int n,
                                                     const DevScales* __restrict__ scp,
                                                     double eta_const, dem::CorrScalars cs,
                                                     const double* __restrict__ F,
                                                     double* __restrict__ V) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double m[3] = {0.0, 0.0, 0.0}, c[3] = {0.0, 0.0, 0.0};
  for (int s = 0; s < S; ++s)
    for (int a = 0; a < 3; ++a) {
      m[a] += part[(size_t)(6 * s + a) * n + i];
      if (cs.enabled) c[a] += part[(size_t)(6 * s + 3 + a) * n + i];
    }
  const DevScales sc = *scp;
  double Vi[3];
  dem::conv_finish(m, c, sc, eta_const, cs, F + 3 * (size_t)i, Vi);
  for (int a = 0; a < 3; ++a) V[3 * (size_t)i + a] = Vi[a];
}

// Separations of the pair-once Oseen sum: 64-bit, 32-bit (open boxes, where
// they are exact), or 32-bit unless the boundary-pair lookup fell back.
enum ConvPos { kPos64 = 0, kPos32 = 1, kPos32Checked = 2 };

__global__ __launch_bounds__(256) void k_conv_sym_reduce(
    const double* __restrict__ part, int S, int n, const dem::PartJ* __restrict__ part_j, int B,
    const DevScales* __restrict__ scp, double eta_const, const double* __restrict__ F,
    double* __restrict__ V, const RecC* __restrict__ recc, int mode, ImgArgs img) {
  const int i = blockIdx.x * 64 + (threadIdx.x >> 2);
  if (i >= n) return;  // the four lanes of a target leave together
  double m[3] = {0.0, 0.0, 0.0};
  const double c[3] = {0.0, 0.0, 0.0};
  dem::sym_sum4(part, S, n, part_j, B, i, m);
  if (threadIdx.x & 3) return;
  if (mode == kPos32Checked && *img.flag == 0) img_correct(recc, n, B, i, img.partner, m);
  const DevScales sc = *scp;
  const dem::CorrScalars cs{};
  double Vi[3];
  dem::conv_finish(m, c, sc, eta_const, cs, F + 3 * (size_t)i, Vi);
  for (int a = 0; a < 3; ++a) V[3 * (size_t)i + a] = Vi[a];
}

// ---------------------------------------------------------------- FP64 oracle kernels

__device__ __forceinline__ double mic64(double d, double L, int periodic) {
  if (periodic && fabs(d) > L * 0.5) d -= copysign(L, d);
  return d;
}

__global__ __launch_bounds__(256) void k_f64_forces(const double* __restrict__ pos,
                                                    const double* __restrict__ r, int n, double m,
                                                    double L, int periodic,
                                                    const int* __restrict__ tgt, int nt,
                                                    double* __restrict__ F) {
  // targets: tgt[k] (k < nt), or k itself when tgt is NULL; output row k
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= nt) return;
  const int i = tgt ? tgt[k] : k;
  const double xi = pos[3 * (size_t)i], yi = pos[3 * (size_t)i + 1], zi = pos[3 * (size_t)i + 2];
  const double ri = r[i];
  double fx = 0.0, fy = 0.0, fz = 0.0;
  for (int j = 0; j < n; ++j) {
    if (j == i) continue;
    double dx = mic64(pos[3 * (size_t)j] - xi, L, periodic);
    double dy = mic64(pos[3 * (size_t)j + 1] - yi, L, periodic);
    double dz = mic64(pos[3 * (size_t)j + 2] - zi, L, periodic);
    double rmag = sqrt(dx * dx + dy * dy + dz * dz);
    double rj = r[j];
    if (rmag > ri + rj) {
      double r7 = rmag * rmag * rmag * rmag * rmag * rmag * rmag;
      double M = m * (ri * ri * ri) * (rj * rj * rj) / r7;
      double dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
      double fxy = M * (4.0 * dz2 - dx2 - dy2);
      fx += fxy * dx;
      fy += fxy * dy;
      fz += M * (2.0 * dz2 - 3.0 * dx2 - 3.0 * dy2) * dz;
    }
  }
  F[3 * (size_t)k] = fx;
  F[3 * (size_t)k + 1] = fy;
  F[3 * (size_t)k + 2] = fz;
}

struct Corr64Args {
  const double* grid;  // (u, v, w, 0) per node
  int n, enabled;
  double gmin, inv_dx, Fz_inv, L_ratio, eta_ratio, self_coeff;
};

__device__ void grid_idx64(double x, double gmin, double inv, int n, int& i, double& t) {
  double f = (x - gmin) * inv;
  f = fmax(0.0, fmin(f, (double)(n - 1)));
  i = min((int)floor(f), n - 2);
  t = f - (double)i;
}

__device__ double interp64(const double* g, int comp, int n, int ix, int iy, int iz, double tx,
                           double ty, double tz) {
  auto at = [&](int a, int b, int c) { return g[4 * (((size_t)(ix + a) * n + (iy + b)) * n + (iz + c)) + comp]; };
  return at(0, 0, 0) * (1 - tx) * (1 - ty) * (1 - tz) + at(0, 0, 1) * (1 - tx) * (1 - ty) * tz +
         at(0, 1, 0) * (1 - tx) * ty * (1 - tz) + at(0, 1, 1) * (1 - tx) * ty * tz +
         at(1, 0, 0) * tx * (1 - ty) * (1 - tz) + at(1, 0, 1) * tx * (1 - ty) * tz +
         at(1, 1, 0) * tx * ty * (1 - tz) + at(1, 1, 1) * tx * ty * tz;
}

__global__ __launch_bounds__(256) void k_f64_conv(const double* __restrict__ pos,
                                                  const double* __restrict__ F, int n, double eta,
                                                  double L, int periodic, Corr64Args c,
                                                  const int* __restrict__ tgt, int nt,
                                                  double* __restrict__ V) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= nt) return;
  const int i = tgt ? tgt[k] : k;
  const double xi = pos[3 * (size_t)i], yi = pos[3 * (size_t)i + 1], zi = pos[3 * (size_t)i + 2];
  double vx = 0.0, vy = 0.0, vz = 0.0;
  for (int j = 0; j < n; ++j) {
    if (j == i) continue;
    double x = mic64(xi - pos[3 * (size_t)j], L, periodic);
    double y = mic64(yi - pos[3 * (size_t)j + 1], L, periodic);
    double z = mic64(zi - pos[3 * (size_t)j + 2], L, periodic);
    double r = sqrt(x * x + y * y + z * z);
    double r3 = r * r * r;
    double fx = F[3 * (size_t)j], fy = F[3 * (size_t)j + 1], fz = F[3 * (size_t)j + 2];
    vx += eta * ((1.0 / r + (x * x) / r3) * fx + (x * y) / r3 * fy + (x * z) / r3 * fz);
    vy += eta * ((x * y) / r3 * fx + (1.0 / r + (y * y) / r3) * fy + (y * z) / r3 * fz);
    vz += eta * ((x * z) / r3 * fx + (y * z) / r3 * fy + (1.0 / r + (z * z) / r3) * fz);
    if (c.enabled) {
      const double cL = c.L_ratio;
      const double xc = x * cL, yc = y * cL, zc = z * cL;
      int i1x, i1y, i1z, i2x, i2y, i2z, i3x, i3y, i3z;
      double t1x, t1y, t1z, t2x, t2y, t2z, t3x, t3y, t3z;
      grid_idx64(xc, c.gmin, c.inv_dx, c.n, i1x, t1x);
      grid_idx64(yc, c.gmin, c.inv_dx, c.n, i1y, t1y);
      grid_idx64(zc, c.gmin, c.inv_dx, c.n, i1z, t1z);
      grid_idx64(zc, c.gmin, c.inv_dx, c.n, i2x, t2x);
      grid_idx64(yc, c.gmin, c.inv_dx, c.n, i2y, t2y);
      grid_idx64(xc, c.gmin, c.inv_dx, c.n, i2z, t2z);
      grid_idx64(xc, c.gmin, c.inv_dx, c.n, i3x, t3x);
      grid_idx64(zc, c.gmin, c.inv_dx, c.n, i3y, t3y);
      grid_idx64(yc, c.gmin, c.inv_dx, c.n, i3z, t3z);
      const int N = c.n;
      double gz_x = interp64(c.grid, 0, N, i1x, i1y, i1z, t1x, t1y, t1z);
      double gz_y = interp64(c.grid, 1, N, i1x, i1y, i1z, t1x, t1y, t1z);
      double gz_z = interp64(c.grid, 2, N, i1x, i1y, i1z, t1x, t1y, t1z);
      double gx_x = interp64(c.grid, 2, N, i2x, i2y, i2z, t2x, t2y, t2z);
      double gx_y = interp64(c.grid, 1, N, i2x, i2y, i2z, t2x, t2y, t2z);
      double gx_z = interp64(c.grid, 0, N, i2x, i2y, i2z, t2x, t2y, t2z);
      double gy_x = interp64(c.grid, 0, N, i3x, i3y, i3z, t3x, t3y, t3z);
      double gy_y = interp64(c.grid, 2, N, i3x, i3y, i3z, t3x, t3y, t3z);
      double gy_z = interp64(c.grid, 1, N, i3x, i3y, i3z, t3x, t3y, t3z);
      const double common = c.Fz_inv * c.eta_ratio * cL;
      const double sx = fx * common, sy = fy * common, sz = fz * common;
      vx += gx_x * sx + gy_x * sy + gz_x * sz;
      vy += gx_y * sx + gy_y * sy + gz_y * sz;
      vz += gx_z * sx + gy_z * sy + gz_z * sz;
    }
  }
  if (c.enabled) {
    vx += c.self_coeff * F[3 * (size_t)i];
    vy += c.self_coeff * F[3 * (size_t)i + 1];
    vz += c.self_coeff * F[3 * (size_t)i + 2];
  }
  V[3 * (size_t)k] = vx;
  V[3 * (size_t)k + 1] = vy;
  V[3 * (size_t)k + 2] = vz;
}

// ---------------------------------------------------------------- host drivers

int check_common(const double* pos, const double* radii, int n, const dem_phys* ph) {
  if (n < 0) return dem::set_error(DEM_ERR_INVALID, "n must be >= 0 (got %d)", n);
  if (!ph) return dem::set_error(DEM_ERR_INVALID, "phys is NULL");
  if (n > 0 && (!pos || !radii)) return dem::set_error(DEM_ERR_INVALID, "null positions/radii");
  if (ph->periodic && !(ph->L > 0.0))
    return dem::set_error(DEM_ERR_INVALID, "periodic box needs L > 0 (got %g)", ph->L);
  if (n > (1 << 28)) return dem::set_error(DEM_ERR_CAPACITY, "n too large (%d)", n);
  return DEM_OK;
}

int check_corr(const dem_correction* c) {
  if (!c || !c->enabled) return DEM_OK;
  if (!c->grid) return dem::set_error(DEM_ERR_INVALID, "correction enabled without a float grid");
  if (c->n < 2) return dem::set_error(DEM_ERR_INVALID, "correction grid needs n >= 2");
  return DEM_OK;
}

// Set-up + pack (shared by all FP32 entry points).
int prepare(const Layout& l, const double* pos, const double* radii, const double* forces, int n,
            const dem_phys* ph, hipStream_t st, bool img = false) {
  int nblk = 1;
  if (!ph->periodic) {
    nblk = (n + 255) / 256;
    if (nblk > kBBoxBlocks) nblk = kBBoxBlocks;
    if (nblk < 1) nblk = 1;
    hipLaunchKernelGGL(k_bbox_partial, dim3(nblk), dim3(256), 0, st, pos, n, l.bbox);
    DEM_CHECK_LAUNCH("k_bbox_partial");
  }
  hipLaunchKernelGGL(k_setup, dim3(1), dim3(64), 0, st, l.bbox, nblk, ph->L, ph->periodic,
                     ph->m_const, l.sc);
  DEM_CHECK_LAUNCH("k_setup");
  hipLaunchKernelGGL(k_pack, dim3((l.npad + 255) / 256), dim3(256), 0, st, pos, radii, forces, n,
                     l.npad, l.sc, l.rec, l.recc, img ? l.img_slot : nullptr,
                     img ? ((size_t)3 << l.img_log2m) : (size_t)0, l.img_flag);
  DEM_CHECK_LAUNCH("k_pack");
  return DEM_OK;
}

constexpr int kMaxDevices = 64;

// Chunks of the pair-once entry lists: the explicit split in mode 2, else
// one round of co-resident workgroups.
int sym_split(int n, int slots) {
  if (g_sym == 2 && g_split > 0) {
    int S = g_split;
    const int emax = dem::sym_entries_max(sym_blocks(n));
    if (S > emax) S = emax;
    if (S > kSymMaxChunks) S = kSymMaxChunks;
    return S < 1 ? 1 : S;
  }
  return sym_chunks(n, slots);
}

// Workgroups of `kernel` (`threads` each) the current device holds at once:
// compute units x resident workgroups per unit, cached per device.
int resident_slots(int* cache, const void* kernel, int threads) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) dev = 0;
  if (cache[dev] > 0) return cache[dev];
  int cus = 0, per = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus <= 0)
    cus = 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, 0) != hipSuccess ||
      per <= 0)
    per = 1;
  cache[dev] = cus * per;
  return cache[dev];
}

// The boundary-pair lookup runs with the pair-once kernels of mode 1 in a
// periodic box when one call evaluates both phases.
bool img_applies(const Layout& l, int n, const dem_phys* ph) {
  return g_sym == 1 && g_img32 && use_sym(n) && ph->periodic && l.img_slot != nullptr;
}

ImgArgs img_args(const Layout& l, bool on) {
  if (!on) return ImgArgs{nullptr, nullptr, nullptr, 0};
  return ImgArgs{l.img_slot, l.img_partner, l.img_flag, l.img_log2m};
}

// Separations of the pair-once Oseen kernel for this call.
int conv_mode(const dem_phys* ph, bool img) {
  if (g_sym != 1 || !g_img32) return kPos64;
  if (!ph->periodic) return kPos32;
  return img ? kPos32Checked : kPos64;
}

int run_forces(const Layout& l, const double* radii, int n, const dem_phys* ph, double* F,
               hipStream_t st, ImgArgs img = ImgArgs{nullptr, nullptr, nullptr, 0}) {
  if (use_sym(n)) {
    static int slots[kMaxDevices];
    const int B = sym_blocks(n);
    const int S = sym_split(n, resident_slots(slots, dem::force_sym_kernel(),
                                              kSymB / kSymTplForce));
    const int ec = (1 + B / 2 + S - 1) / S;
    DEM_PROF_LAUNCH("k_force_sym", st,
                    dem::launch_force_sym(dim3(B, S), st, (const Rec*)l.rec, (const RecC*)l.recc,
                                          n, B, ec, choose_tile_force(), l.part, l.part_j, img));
    DEM_CHECK_LAUNCH("k_force_sym");
    hipLaunchKernelGGL(k_force_sym_reduce, dim3((n + 63) / 64), dim3(256), 0, st,
                       (const double*)l.part, S, n, (const dem::PartJ*)l.part_j, B, radii,
                       (const DevScales*)l.sc, ph->m_const, F, l.recc, img);
    DEM_CHECK_LAUNCH("k_force_sym_reduce");
    return DEM_OK;
  }
  const int S = choose_split(n);
  const int jchunk = (n + S - 1) / S;
  dim3 grid((n + WG - 1) / WG, S);
  DEM_PROF_LAUNCH("k_force_tile", st,
                  hipLaunchKernelGGL(k_force_tile, grid, dim3(WG), 0, st, (const Rec*)l.rec, n,
                                     jchunk, choose_tile_force(), l.part));
  DEM_CHECK_LAUNCH("k_force_tile");
  hipLaunchKernelGGL(k_force_reduce, dim3((n + 255) / 256), dim3(256), 0, st, (const double*)l.part,
                     S, n, radii, (const DevScales*)l.sc, ph->m_const, F, l.recc);
  DEM_CHECK_LAUNCH("k_force_reduce");
  return DEM_OK;
}

int run_conv(const Layout& l, int n, const dem_phys* ph, const dem_correction* corr,
             const double* F, double* V, hipStream_t st, int mode = kPos64,
             ImgArgs img = ImgArgs{nullptr, nullptr, nullptr, 0}) {
  const int S = choose_split(n);
  const int jchunk = (n + S - 1) / S;
  dim3 grid((n + WG - 1) / WG, S);
  CorrArgs ca{};
  dem::CorrScalars cs{};
  const bool on = corr && corr->enabled;
  if (!on && use_sym(n)) {
    static int slots32[kMaxDevices], slots64[kMaxDevices];
    const int B = sym_blocks(n);
    const bool p32 = mode != kPos64;
    const int S = p32 ? sym_split(n, resident_slots(slots32, dem::conv_sym_kernel(true),
                                                    kSymB / kSymTplConv))
                      : sym_split(n, resident_slots(slots64, dem::conv_sym_kernel(false),
                                                    kSymB / kSymTplConv));
    const int ec = (1 + B / 2 + S - 1) / S;
    const int* flag = mode == kPos32Checked ? (const int*)img.flag : nullptr;
    if (p32)
      DEM_PROF_LAUNCH("k_conv_sym", st,
                      dem::launch_conv_sym(true, dim3(B, S), st, (const RecC*)l.recc, n, B, ec,
                                           choose_tile_conv(), l.part, l.part_j, flag, 0));
    if (!p32 || flag)
      DEM_PROF_LAUNCH(p32 ? "k_conv_sym_fallback" : "k_conv_sym", st,
                      dem::launch_conv_sym(false, dim3(B, S), st, (const RecC*)l.recc, n, B, ec,
                                           choose_tile_conv(), l.part, l.part_j, flag, 1));
    DEM_CHECK_LAUNCH("k_conv_sym");
    hipLaunchKernelGGL(k_conv_sym_reduce, dim3((n + 63) / 64), dim3(256), 0, st,
                       (const double*)l.part, S, n, (const dem::PartJ*)l.part_j, B,
                       (const DevScales*)l.sc, ph->eta_const, F, V, (const RecC*)l.recc, mode,
                       img);
    DEM_CHECK_LAUNCH("k_conv_sym_reduce");
    return DEM_OK;
  }
  if (on) {
    ca.grid = reinterpret_cast<const float4*>(corr->grid);
    ca.n = corr->n;
    ca.nmax = (float)(corr->n - 1);
    ca.gmin = corr->grid_min;
    ca.inv_dx = corr->inv_dx;
    ca.L_ratio = corr->L_ratio;
    cs.enabled = 1;
    cs.common = corr->Fz_inv * corr->eta_ratio * corr->L_ratio;
    cs.self_coeff = corr->self_coeff;
    DEM_PROF_LAUNCH("k_conv_tile_corr", st,
                    hipLaunchKernelGGL(k_conv_tile<true>, grid, dim3(WG), 0, st,
                                       (const RecC*)l.recc, n, jchunk, choose_tile_conv(),
                                       (const DevScales*)l.sc, ca, l.part));
  } else {
    DEM_PROF_LAUNCH("k_conv_tile", st,
                    hipLaunchKernelGGL(k_conv_tile<false>, grid, dim3(WG), 0, st,
                                       (const RecC*)l.recc, n, jchunk, choose_tile_conv(),
                                       (const DevScales*)l.sc, ca, l.part));
  }
  DEM_CHECK_LAUNCH("k_conv_tile");
  hipLaunchKernelGGL(k_conv_reduce, dim3((n + 255) / 256), dim3(256), 0, st, (const double*)l.part,
                     S, n, (const DevScales*)l.sc, ph->eta_const, cs, F, V);
  DEM_CHECK_LAUNCH("k_conv_reduce");
  return DEM_OK;
}

}  // namespace

extern "C" {

size_t dem_direct_workspace_bytes(int n) { return carve(nullptr, n < 0 ? 0 : n).bytes; }

int dem_direct_set_symmetric(int on) {
  const int prev = g_sym;
  g_sym = on < 0 ? 0 : (on > 2 ? 2 : on);
  return prev;
}

int dem_direct_set_image32(int on) {
  const int prev = g_img32;
  g_img32 = on ? 1 : 0;
  return prev;
}

int dem_direct_set_tuning(int split, int tile) {
  int prev = (g_split << 16) | (g_tile & 0xffff);
  g_split = split > 0 ? split : 0;
  g_tile = tile > 0 ? tile : 0;
  return prev;
}

int dem_direct_forces(const double* pos, const double* radii, int n, const dem_phys* ph, double* F,
                      void* ws, size_t ws_bytes, dem_stream_t stream) {
  int rc = check_common(pos, radii, n, ph);
  if (rc) return rc;
  if (n > 0 && !F) return dem::set_error(DEM_ERR_INVALID, "null forces");
  if (n == 0) return DEM_OK;
  Layout l = carve(ws, n);
  if (!ws || ws_bytes < l.bytes)
    return dem::set_error(DEM_ERR_WORKSPACE, "workspace %zu < %zu bytes", ws_bytes, l.bytes);
  hipStream_t st = dem::as_stream(stream);
  if ((rc = prepare(l, pos, radii, nullptr, n, ph, st))) return rc;
  return run_forces(l, radii, n, ph, F, st);
}

int dem_direct_convection(const double* pos, const double* radii, const double* F, int n,
                          const dem_phys* ph, const dem_correction* corr, double* V, void* ws,
                          size_t ws_bytes, dem_stream_t stream) {
  int rc = check_common(pos, radii, n, ph);
  if (rc) return rc;
  if ((rc = check_corr(corr))) return rc;
  if (n > 0 && (!F || !V)) return dem::set_error(DEM_ERR_INVALID, "null forces/velocities");
  if (n == 0) return DEM_OK;
  Layout l = carve(ws, n);
  if (!ws || ws_bytes < l.bytes)
    return dem::set_error(DEM_ERR_WORKSPACE, "workspace %zu < %zu bytes", ws_bytes, l.bytes);
  hipStream_t st = dem::as_stream(stream);
  if ((rc = prepare(l, pos, radii, F, n, ph, st))) return rc;
  return run_conv(l, n, ph, corr, F, V, st, conv_mode(ph, false));
}

int dem_direct_forces_and_convection(const double* pos, const double* radii, int n,
                                     const dem_phys* ph, const dem_correction* corr, double* F,
                                     double* V, void* ws, size_t ws_bytes, dem_stream_t stream) {
  int rc = check_common(pos, radii, n, ph);
  if (rc) return rc;
  if ((rc = check_corr(corr))) return rc;
  if (n > 0 && (!F || !V)) return dem::set_error(DEM_ERR_INVALID, "null forces/velocities");
  if (n == 0) return DEM_OK;
  Layout l = carve(ws, n);
  if (!ws || ws_bytes < l.bytes)
    return dem::set_error(DEM_ERR_WORKSPACE, "workspace %zu < %zu bytes", ws_bytes, l.bytes);
  hipStream_t st = dem::as_stream(stream);
  const bool img = !(corr && corr->enabled) && img_applies(l, n, ph);
  const ImgArgs ia = img_args(l, img);
  if ((rc = prepare(l, pos, radii, nullptr, n, ph, st, img))) return rc;
  if ((rc = run_forces(l, radii, n, ph, F, st, ia))) return rc;
  return run_conv(l, n, ph, corr, F, V, st, conv_mode(ph, img), ia);
}

int dem_direct_forces_and_convection_f64(const double* pos, const double* radii, int n,
                                         const dem_phys* ph, const dem_correction* corr,
                                         double* F, double* V, dem_stream_t stream) {
  int rc = check_common(pos, radii, n, ph);
  if (rc) return rc;
  if (n > 0 && (!F || !V)) return dem::set_error(DEM_ERR_INVALID, "null forces/velocities");
  Corr64Args c{};
  if (corr && corr->enabled) {
    if (!corr->grid64 || corr->n < 2)
      return dem::set_error(DEM_ERR_INVALID, "FP64 oracle needs the float64 correction grid");
    c.grid = corr->grid64;
    c.n = corr->n;
    c.enabled = 1;
    c.gmin = corr->grid_min;
    c.inv_dx = corr->inv_dx;
    c.Fz_inv = corr->Fz_inv;
    c.L_ratio = corr->L_ratio;
    c.eta_ratio = corr->eta_ratio;
    c.self_coeff = corr->self_coeff;
  }
  if (n == 0) return DEM_OK;
  hipStream_t st = dem::as_stream(stream);
  hipLaunchKernelGGL(k_f64_forces, dim3((n + 255) / 256), dim3(256), 0, st, pos, radii, n,
                     ph->m_const, ph->L, ph->periodic, (const int*)nullptr, n, F);
  DEM_CHECK_LAUNCH("k_f64_forces");
  hipLaunchKernelGGL(k_f64_conv, dim3((n + 255) / 256), dim3(256), 0, st, pos, (const double*)F, n,
                     ph->eta_const, ph->L, ph->periodic, c, (const int*)nullptr, n, V);
  DEM_CHECK_LAUNCH("k_f64_conv");
  return DEM_OK;
}

int dem_direct_f64_targets(const double* pos, const double* radii, int n, const dem_phys* ph,
                           const dem_correction* corr, const int32_t* targets, int nt,
                           const double* sources_forces, double* F, double* V,
                           dem_stream_t stream) {
  int rc = check_common(pos, radii, n, ph);
  if (rc) return rc;
  if (nt < 0 || (nt > 0 && (!targets || !F)))
    return dem::set_error(DEM_ERR_INVALID, "bad target list / output");
  if (V && !sources_forces)
    return dem::set_error(DEM_ERR_INVALID, "velocities need the forces of all sources");
  Corr64Args c{};
  if (corr && corr->enabled) {
    if (!corr->grid64 || corr->n < 2)
      return dem::set_error(DEM_ERR_INVALID, "FP64 oracle needs the float64 correction grid");
    c.grid = corr->grid64;
    c.n = corr->n;
    c.enabled = 1;
    c.gmin = corr->grid_min;
    c.inv_dx = corr->inv_dx;
    c.Fz_inv = corr->Fz_inv;
    c.L_ratio = corr->L_ratio;
    c.eta_ratio = corr->eta_ratio;
    c.self_coeff = corr->self_coeff;
  }
  if (nt == 0 || n == 0) return DEM_OK;
  hipStream_t st = dem::as_stream(stream);
  hipLaunchKernelGGL(k_f64_forces, dim3((nt + 255) / 256), dim3(256), 0, st, pos, radii, n,
                     ph->m_const, ph->L, ph->periodic, (const int*)targets, nt, F);
  DEM_CHECK_LAUNCH("k_f64_forces");
  if (V) {
    hipLaunchKernelGGL(k_f64_conv, dim3((nt + 255) / 256), dim3(256), 0, st, pos, sources_forces,
                       n, ph->eta_const, ph->L, ph->periodic, c, (const int*)targets, nt, V);
    DEM_CHECK_LAUNCH("k_f64_conv");
  }
  return DEM_OK;
}

}  // extern "C"
