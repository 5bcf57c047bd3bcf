// Pair-once (Newton's third law) direct sums shared by the single-system
// kernels (direct.hip) and the batched ones (batch.hip): the block-cyclic
// partner schedule, the LDS slot accumulators and the per-workgroup bodies.
#pragma once

#include <stdint.h>

#include <type_traits>

#include "direct_core.h"
#include "pair_math.h"

namespace dem {

// Reversed-pair partial sums as stored: each is a sum of FP32 slot
// accumulators, kept in FP32 (half the bytes of the round trip to the reduce).
using PartJ = float;

constexpr int kSymB = 256;           // droplets per block
constexpr int kSymLds = kSymB + 64;  // source slots lane + k, k < 256, without a wrap
constexpr int kSymTplForce = 2;      // targets per thread, force / Oseen bodies
constexpr int kSymTplConv = 2;

// Partners of block b among B blocks: b pairs with blocks (b + e) mod B,
// e = 1..P_b, which covers every unordered block pair exactly once.
__host__ __device__ inline int sym_partners(int b, int B) {
  if (B & 1) return (B - 1) / 2;
  return b < B / 2 ? B / 2 : B / 2 - 1;
}

// Entries (itself + partners) of the block with the most.
__host__ __device__ inline int sym_entries_max(int B) { return 1 + B / 2; }

// Block b (256 targets; 256 / TPL threads, thread t holding targets
// t + (256 / TPL) h, h < TPL) works through entries e of its partner list in
// [e0, e1): e = 0 is b itself (one-sided, every ordered pair of the block),
// e >= 1 is block c = (b + e) mod B, whose pairs with b are evaluated once
// and credited to both sides.  Sources sit in LDS as arrays; lane l at step
// k takes source slot l + k (the first 64 sources are repeated after 256, so
// no wrap) for each of its targets.  The reversed pairs' sums are added into
// the wave's own LDS accumulator of that slot by a plain read-add-write: one
// instruction's lanes touch distinct slots, and steps run in groups
// k = kk + 64u (u = 0..3) whose four slot windows [k, k + 63] are disjoint, so
// only consecutive groups share slots and wave_sync() orders them.  Per
// (b, c) the waves' accumulators are summed in fixed order (in float64) and
// stored as FP32 in part_j[b][e-1][3][256]; the reduce adds them in float64.
// The target side keeps float64 accumulators as the one-sided kernel does.

// Orders one wave's LDS accesses across its lanes: the compiler keeps program
// order across it, and the LDS executes a wave's instructions in order.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Source slot s (0..255) and, for s < 64, its repeat at s + 256.
template <typename T>
__device__ __forceinline__ void sym_stage(T* a, T v, int s) {
  a[s] = v;
  if (s < 64) a[s + kSymB] = v;
}

__device__ __forceinline__ void sym_zero(float (*acc)[kSymLds], int lane) {
  for (int q = lane; q < kSymLds; q += 64) {
    acc[0][q] = 0.f;
    acc[1][q] = 0.f;
    acc[2][q] = 0.f;
  }
}

// Reversed-pair sums of the partner block's sources t + T h, all W waves.
template <int T, int W>
__device__ __forceinline__ void sym_fold(const float (*acc)[3][kSymLds], int t,
                                         PartJ* __restrict__ dst) {
  for (int slot = t; slot < kSymB; slot += T) {
    for (int a = 0; a < 3; ++a) {
      double v = 0.0;
      for (int w = 0; w < W; ++w) {
        v += (double)acc[w][a][slot];
        if (slot < 64) v += (double)acc[w][a][slot + kSymB];
      }
      dst[a * kSymB + slot] = (PartJ)v;
    }
  }
}

// Dipole pair of one target with a source (radii rif, rjf in fixed units, rj3
// in work units): adds to the target's FP32 sums and returns g (the pair's
// force per r_i^3 r_j^3, odd in d) for the reversed side.  The contact mask
// |d| > r_i + r_j is h = |d|^2 - (r_i + r_j)^2 > 0 in fixed units, where a
// positive h is at least 2^20 (an ulp of |d|^2 there) and so exceeds 1/|d|
// in work units: med3(h, 0, 1/|d|) is 1/|d| for separated pairs and 0 for
// touching ones and for j == i.
__device__ __forceinline__ void force_pair_sym(float dx, float dy, float dz, float rif, float rjf,
                                               float rj3, float* f, float* g) {
  const float dz2 = dz * dz;
  const float dxy2 = dx * dx + dy * dy;
  const float r2f = dxy2 + dz2;
  const float rsf = rif + rjf;
  const float h = fmaf(-rsf, rsf, r2f);
  const float ir = __builtin_amdgcn_fmed3f(h, 0.f, __builtin_amdgcn_rsqf(r2f * dem::kSigma2f));
  const float ir2 = ir * ir;
  const float sm = (ir2 * ir2) * (ir2 * ir);
  const float A = 4.f * dz2 - dxy2;
  const float Bz = A - 2.f * r2f;  // 2 dz^2 - 3 (dx^2 + dy^2)
  const float cxy = sm * A, cz = sm * Bz;
  g[0] = cxy * dx;
  g[1] = cxy * dy;
  g[2] = cz * dz;
  f[0] += rj3 * g[0];
  f[1] += rj3 * g[1];
  f[2] += rj3 * g[2];
}

// One workgroup (kSymB / TPL threads) of the pair-once force sum of a system
// of n droplets in B blocks (records padded to 256 B, zero beyond n): target
// block b, chunk s, partner entries [e0, e1).  Writes the target sums of chunk
// s to part[(3 s + a) n + i] and, per partner entry e >= 1, the reversed sums
// to part_j[b][e-1][3][256].
template <int TPL>
__device__ __forceinline__ void force_sym_block(const Rec* __restrict__ rec, int n, int B, int b,
                                                int s, int e0, int e1, int tile,
                                                double* __restrict__ part,
                                                PartJ* __restrict__ part_j) {
  constexpr int T = kSymB / TPL, W = T / 64;
  __shared__ uint32_t sx[kSymLds], sy[kSymLds], sz[kSymLds];
  __shared__ float sr[kSymLds], sr3[kSymLds];  // radius in fixed units, r^3 in work units
  __shared__ float acc[W][3][kSymLds];
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  Rec m[TPL];
  float nr3[TPL], rf[TPL];
  double ai[TPL][3];
  for (int h = 0; h < TPL; ++h) {
    m[h] = rec[(size_t)b * kSymB + t + T * h];
    nr3[h] = -m[h].r3;
    rf[h] = m[h].r * 0x1p22f;
    ai[h][0] = ai[h][1] = ai[h][2] = 0.0;
  }
  const int gt = max(1, tile / 4);  // step groups per FP32 partial
  for (int e = e0; e < e1; ++e) {
    const int c = e == 0 ? b : (b + e) % B;
    __syncthreads();  // the previous entry's LDS reads are done
    for (int sl = t; sl < kSymB; sl += T) {
      const Rec rc = rec[(size_t)c * kSymB + sl];
      sym_stage(sx, rc.x, sl);
      sym_stage(sy, rc.y, sl);
      sym_stage(sz, rc.z, sl);
      sym_stage(sr, rc.r * 0x1p22f, sl);
      sym_stage(sr3, rc.r3, sl);
    }
    sym_zero(acc[w], lane);
    __syncthreads();
    if (e == 0) {
      for (int kt = 0; kt < kSymB; kt += tile) {
        const int ke = min(kt + tile, kSymB);
        float f[TPL][3] = {};
#pragma unroll 2
        for (int k = kt; k < ke; ++k) {
          const int q = lane + k;
          const uint32_t xj = sx[q], yj = sy[q], zj = sz[q];
          const float rj = sr[q], rj3 = sr3[q];
#pragma unroll
          for (int h = 0; h < TPL; ++h)
            dem::dipole_pair<false>(dem::diff32(xj, m[h].x), dem::diff32(yj, m[h].y),
                                    dem::diff32(zj, m[h].z), m[h].r, rj * 0x1p-22f, rj3, f[h][0],
                                    f[h][1], f[h][2]);
        }
        for (int h = 0; h < TPL; ++h)
          for (int a = 0; a < 3; ++a) ai[h][a] += (double)f[h][a];
      }
      continue;
    }
    for (int g0 = 0; g0 < 64; g0 += gt) {
      const int ge = min(g0 + gt, 64);
      float f[TPL][3] = {};
      for (int kk = g0; kk < ge; ++kk) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int q = lane + kk + 64 * u;
          const uint32_t xj = sx[q], yj = sy[q], zj = sz[q];
          const float rj = sr[q], rj3 = sr3[q];
          float o[3] = {acc[w][0][q], acc[w][1][q], acc[w][2][q]};
#pragma unroll
          for (int h = 0; h < TPL; ++h) {
            float g[3];
            force_pair_sym(dem::diff32(xj, m[h].x), dem::diff32(yj, m[h].y),
                           dem::diff32(zj, m[h].z), rf[h], rj, rj3, f[h], g);
            // g is odd in d: the reversed pair adds -r_i^3 g to source q
            for (int a = 0; a < 3; ++a) o[a] = fmaf(nr3[h], g[a], o[a]);
          }
          acc[w][0][q] = o[0];
          acc[w][1][q] = o[1];
          acc[w][2][q] = o[2];
        }
        wave_sync();
      }
      for (int h = 0; h < TPL; ++h)
        for (int a = 0; a < 3; ++a) ai[h][a] += (double)f[h][a];
    }
    __syncthreads();
    sym_fold<T, W>(acc, t, part_j + ((size_t)b * (B / 2) + (e - 1)) * 3 * kSymB);
  }
  for (int h = 0; h < TPL; ++h) {
    const int i = b * kSymB + t + T * h;
    if (i < n)
      for (int a = 0; a < 3; ++a) part[(size_t)(3 * s + a) * n + i] = ai[h][a];
  }
}

// Oseen sum, pair-once (no periodic correction): the Oseen tensor is even in
// the separation, so the pair's tensor serves both sides; same schedule, slot
// order and partial layout as k_force_sym.  Separations are the high words of
// the 64-bit ring differences, as in the one-sided kernel: a pair whose
// separation lies within one 32-bit unit of half the box along an axis still
// takes the image the exact positions give (32-bit positions could round it
// to the other image, which flips that component of the pair's term).

// Oseen pair of one target (force fi) with a source (force Fj): adds the
// source's term to the target's sums and returns the target's term for the
// source as ir * o.
__device__ __forceinline__ float oseen_pair_sym(float dx, float dy, float dz, float Fx, float Fy,
                                                float Fz, const float* fi, float* v, float* o) {
  const float r2 = dx * dx + dy * dy + dz * dz;
  // coincident droplets add nothing: a nonzero r2 is at least 1 >= 1/|d|
  const float ir = __builtin_amdgcn_fmed3f(r2, 0.f, __builtin_amdgcn_rsqf(r2));
  const float ir2 = ir * ir;
  const float tj = (dx * Fx + dy * Fy + dz * Fz) * ir2;
  v[0] += ir * (Fx + tj * dx);
  v[1] += ir * (Fy + tj * dy);
  v[2] += ir * (Fz + tj * dz);
  const float ti = (dx * fi[0] + dy * fi[1] + dz * fi[2]) * ir2;
  o[0] = fi[0] + ti * dx;
  o[1] = fi[1] + ti * dy;
  o[2] = fi[2] + ti * dz;
  return ir;
}

// LDS of one pair-once Oseen workgroup of W waves: source positions (64-bit
// ring words, or with POS32 their 32-bit high words), source forces and the
// waves' reversed-pair slot accumulators.
template <int W, bool POS32>
struct ConvSymLds {
  typename std::conditional<POS32, uint32_t, unsigned long long>::type pos[3][kSymLds];
  float f[3][kSymLds];
  float acc[W][3][kSymLds];
};

// One workgroup of the pair-once Oseen sum (same arguments and partial
// layout as force_sym_block) in the caller's LDS.  POS32 takes separations
// from the 32-bit high words of the positions (floor of the 64-bit ring
// value): c = h_j - h_i is the 64-bit difference's high word plus the low
// words' borrow, so it is exact to one unit except that c == 2^31 (exactly
// half the ring) may stand for +2^31 - 1, the other periodic image.  Callers
// that take POS32 in a periodic box correct those pairs afterwards; in an
// open box the high words stay below 2^31 and c never wraps.
template <int TPL, bool POS32>
__device__ __forceinline__ void conv_sym_block(ConvSymLds<kSymB / TPL / 64, POS32>& L,
                                               const RecC* __restrict__ rec, int n, int B, int b,
                                               int s, int e0, int e1, int tile,
                                               double* __restrict__ part,
                                               PartJ* __restrict__ part_j) {
  constexpr int T = kSymB / TPL, W = T / 64;
  using PosW = typename std::conditional<POS32, uint32_t, unsigned long long>::type;
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  unsigned long long nq[TPL][3];  // negated 64-bit target positions
  uint32_t hq[TPL][3];            // 32-bit target positions
  float fi[TPL][3];
  double ai[TPL][3];
  for (int h = 0; h < TPL; ++h) {
    const RecC m = rec[(size_t)b * kSymB + t + T * h];
    if (POS32) {
      hq[h][0] = m.xh;
      hq[h][1] = m.yh;
      hq[h][2] = m.zh;
    } else {
      nq[h][0] = 0ull - (((unsigned long long)m.xh << 32) | m.xl);
      nq[h][1] = 0ull - (((unsigned long long)m.yh << 32) | m.yl);
      nq[h][2] = 0ull - (((unsigned long long)m.zh << 32) | m.zl);
    }
    fi[h][0] = m.fx;
    fi[h][1] = m.fy;
    fi[h][2] = m.fz;
    ai[h][0] = ai[h][1] = ai[h][2] = 0.0;
  }
  // separation of target h from the source in LDS slot q, axis a
  auto sep = [&](int h, int a, int q) -> float {
    if (POS32) return dem::diff32((uint32_t)L.pos[a][q], hq[h][a]);
    return dem::diff64n((unsigned long long)L.pos[a][q], nq[h][a]);
  };
  const int gt = max(1, tile / 4);
  for (int e = e0; e < e1; ++e) {
    const int c = e == 0 ? b : (b + e) % B;
    __syncthreads();
    for (int sl = t; sl < kSymB; sl += T) {
      const RecC rc = rec[(size_t)c * kSymB + sl];
      if (POS32) {
        sym_stage(L.pos[0], (PosW)rc.xh, sl);
        sym_stage(L.pos[1], (PosW)rc.yh, sl);
        sym_stage(L.pos[2], (PosW)rc.zh, sl);
      } else {
        sym_stage(L.pos[0], (PosW)(((unsigned long long)rc.xh << 32) | rc.xl), sl);
        sym_stage(L.pos[1], (PosW)(((unsigned long long)rc.yh << 32) | rc.yl), sl);
        sym_stage(L.pos[2], (PosW)(((unsigned long long)rc.zh << 32) | rc.zl), sl);
      }
      sym_stage(L.f[0], rc.fx, sl);
      sym_stage(L.f[1], rc.fy, sl);
      sym_stage(L.f[2], rc.fz, sl);
    }
    sym_zero(L.acc[w], lane);
    __syncthreads();
    if (e == 0) {
      for (int kt = 0; kt < kSymB; kt += tile) {
        const int ke = min(kt + tile, kSymB);
        float v[TPL][3] = {};
#pragma unroll 2
        for (int k = kt; k < ke; ++k) {
          const int q = lane + k;
          const float Fx = L.f[0][q], Fy = L.f[1][q], Fz = L.f[2][q];
#pragma unroll
          for (int h = 0; h < TPL; ++h) {
            const float dx = sep(h, 0, q), dy = sep(h, 1, q), dz = sep(h, 2, q);
            dem::oseen_pair(dx, dy, dz, Fx, Fy, Fz, dx * dx + dy * dy + dz * dz > 0.f,  // j != i
                            v[h][0], v[h][1], v[h][2]);
          }
        }
        for (int h = 0; h < TPL; ++h)
          for (int a = 0; a < 3; ++a) ai[h][a] += (double)v[h][a];
      }
      continue;
    }
    for (int g0 = 0; g0 < 64; g0 += gt) {
      const int ge = min(g0 + gt, 64);
      float v[TPL][3] = {};
      for (int kk = g0; kk < ge; ++kk) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int q = lane + kk + 64 * u;
          const float Fx = L.f[0][q], Fy = L.f[1][q], Fz = L.f[2][q];
          float o[3] = {L.acc[w][0][q], L.acc[w][1][q], L.acc[w][2][q]};
#pragma unroll
          for (int h = 0; h < TPL; ++h) {
            float oh[3];
            const float ir = oseen_pair_sym(sep(h, 0, q), sep(h, 1, q), sep(h, 2, q), Fx, Fy, Fz,
                                            fi[h], v[h], oh);
            for (int a = 0; a < 3; ++a) o[a] = fmaf(ir, oh[a], o[a]);
          }
          L.acc[w][0][q] = o[0];
          L.acc[w][1][q] = o[1];
          L.acc[w][2][q] = o[2];
        }
        wave_sync();
      }
      for (int h = 0; h < TPL; ++h)
        for (int a = 0; a < 3; ++a) ai[h][a] += (double)v[h][a];
    }
    __syncthreads();
    sym_fold<T, W>(L.acc, t, part_j + ((size_t)b * (B / 2) + (e - 1)) * 3 * kSymB);
  }
  for (int h = 0; h < TPL; ++h) {
    const int i = b * kSymB + t + T * h;
    if (i < n)
      for (int a = 0; a < 3; ++a) part[(size_t)(3 * s + a) * n + i] = ai[h][a];
  }
}

// The 64-bit form in its own LDS (the batched kernels).
template <int TPL>
__device__ __forceinline__ void conv_sym_block(const RecC* __restrict__ rec, int n, int B, int b,
                                               int s, int e0, int e1, int tile,
                                               double* __restrict__ part,
                                               PartJ* __restrict__ part_j) {
  __shared__ ConvSymLds<kSymB / TPL / 64, false> L;
  conv_sym_block<TPL, false>(L, rec, n, B, b, s, e0, e1, tile, part, part_j);
}

// Reversed-pair sums of target i from the blocks b = bi - e (mod B) that
// have block bi as their e-th partner (e <= P_b), over e = k + 1, k + 5, ...
// (one of four interleaved streams, k = 0..3), in ascending e.
__device__ __forceinline__ void sym_reversed(const PartJ* __restrict__ part_j, int B, int i, int k,
                                             double* r) {
  const int bi = i / kSymB, il = i % kSymB, Pmax = B / 2;
  r[0] = r[1] = r[2] = 0.0;
#pragma unroll 2
  for (int e = k + 1; e <= Pmax; e += 4) {
    const int b = bi >= e ? bi - e : bi - e + B;
    if (e > sym_partners(b, B)) continue;
    const PartJ* q = part_j + ((size_t)b * Pmax + (e - 1)) * 3 * kSymB + il;
    r[0] += q[0];
    r[1] += q[kSymB];
    r[2] += q[2 * kSymB];
  }
}

// Sum of a target's pair-once partials: its own S chunks in order, plus the
// reversed pairs as ((r_0 + r_1) + (r_2 + r_3)) of the four streams of
// sym_reversed.  The four-lane reduce kernels form the same sums in the same
// order, so every caller gets identical bits.
__device__ __forceinline__ void sym_sum(const double* __restrict__ part, int S, int n,
                                        const PartJ* __restrict__ part_j, int B, int i,
                                        double* sum) {
  for (int s = 0; s < S; ++s)
    for (int a = 0; a < 3; ++a) sum[a] += part[(size_t)(3 * s + a) * n + i];
  double r[4][3];
  for (int k = 0; k < 4; ++k) sym_reversed(part_j, B, i, k, r[k]);
  for (int a = 0; a < 3; ++a) sum[a] += (r[0][a] + r[1][a]) + (r[2][a] + r[3][a]);
}

// sym_sum with the four streams on four consecutive lanes (lane k = lane & 3);
// the result is valid on lane k = 0.
__device__ __forceinline__ void sym_sum4(const double* __restrict__ part, int S, int n,
                                         const PartJ* __restrict__ part_j, int B, int i,
                                         double* sum) {
  const int k = threadIdx.x & 3;
  double r[3];
  sym_reversed(part_j, B, i, k, r);
  for (int a = 0; a < 3; ++a) {
    r[a] += __shfl_xor(r[a], 1);  // lanes 0, 1: r_0 + r_1; lanes 2, 3: r_2 + r_3
    r[a] += __shfl_xor(r[a], 2);  // lane 0: (r_0 + r_1) + (r_2 + r_3)
  }
  if (k != 0) return;
  for (int s = 0; s < S; ++s)
    for (int a = 0; a < 3; ++a) sum[a] += part[(size_t)(3 * s + a) * n + i];
  for (int a = 0; a < 3; ++a) sum[a] += r[a];
}

}  // namespace dem
