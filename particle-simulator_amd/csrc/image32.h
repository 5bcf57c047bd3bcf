// Boundary pairs of the pair-once Oseen sum on 32-bit separations: the
// high-word tables, their insert and lookup, and the correction the reduce
// applies.  Shared by the pair kernels (direct_pairs.hip) and the reduces
// (direct.hip).
#pragma once
#include <math.h>
#include <stdint.h>

#include <hip/hip_runtime.h>

#include "direct_core.h"
#include "direct_sym.h"

namespace dem {

// The pair-once Oseen kernel may take separations from the 32-bit high words
// h = q >> 32 of the 64-bit ring positions q.  Then c = h_j - h_i (mod 2^32)
// is hi(q_j - q_i) plus the borrow of the low words, and its sign -- the
// periodic image -- differs from the 64-bit difference's only when
// c == 2^31 exactly on some axis, where hi(q_j - q_i) may be 2^31 - 1 (the
// pair lies within one unit of half the box).  Those pairs are found by an
// exact-match join: per axis, an open-addressing table keyed by h (each slot
// (h << 32) | (i + 1), 0 empty, bit 31 marking a key that several droplets
// share) is probed for h ^ 2^31.  A droplet has at most one partner per axis
// unless a key repeats; a repeated key with a partner makes the call fall back
// to 64-bit separations (img_flag).  The reduce then replaces, for each
// partner, the term the kernel added with the 64-bit one.

constexpr unsigned long long kImgMulti = 1ull << 31;

__device__ __forceinline__ uint32_t img_hash(uint32_t key, int log2m) {
  return (key * 0x9E3779B1u) >> (32 - log2m);
}

__device__ __forceinline__ void img_insert(unsigned long long* __restrict__ tab, int log2m,
                                           uint32_t key, int i) {
  const uint32_t mask = (1u << log2m) - 1u;
  const unsigned long long val = ((unsigned long long)key << 32) | (unsigned long long)(i + 1);
  uint32_t slot = img_hash(key, log2m);
  for (uint32_t it = 0; it <= mask; ++it) {
    const unsigned long long old = atomicCAS(tab + slot, 0ull, val);
    if (old == 0ull) return;
    if ((uint32_t)(old >> 32) == key) {
      atomicOr(tab + slot, kImgMulti);
      return;
    }
    slot = (slot + 1u) & mask;
  }
}

// Droplet whose high word is key ^ 2^31: its index, -1 for none, -2 when
// several droplets share that word.
__device__ __forceinline__ int img_probe(const unsigned long long* __restrict__ tab, int log2m,
                                         uint32_t key) {
  const uint32_t mask = (1u << log2m) - 1u, want = key ^ 0x80000000u;
  uint32_t slot = img_hash(want, log2m);
  for (uint32_t it = 0; it <= mask; ++it) {
    const unsigned long long v = tab[slot];
    if (v == 0ull) return -1;
    if ((uint32_t)(v >> 32) == want) return (v & kImgMulti) ? -2 : (int)(v & 0x7fffffffull) - 1;
    slot = (slot + 1u) & mask;
  }
  return -1;
}

__device__ __forceinline__ uint32_t rec_hi(const RecC& r, int a) {
  return a == 0 ? r.xh : (a == 1 ? r.yh : r.zh);
}
__device__ __forceinline__ uint32_t rec_lo(const RecC& r, int a) {
  return a == 0 ? r.xl : (a == 1 ? r.yl : r.zl);
}

// Oseen term F/|d| + d (d.F)/|d|^3 in float64 (fixed units, as the pair
// kernels accumulate it); nothing for d == 0.
__device__ __forceinline__ void oseen_term64(const double* d, const double* f, double* out) {
#pragma clang fp contract(off)
  const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  if (!(r2 > 0.0)) {
    out[0] = out[1] = out[2] = 0.0;
    return;
  }
  const double ir = 1.0 / sqrt(r2);
  const double t = (d[0] * f[0] + d[1] * f[1] + d[2] * f[2]) * (ir * ir);
  for (int a = 0; a < 3; ++a) out[a] = ir * (f[a] + t * d[a]);
}

// Correction of droplet i's Oseen sum for its boundary partners (at most one
// per axis, taken in ascending index): the 64-bit term minus the term the
// 32-bit pair kernel added.  The kernel evaluates each pair from its target:
// droplet i is the target when j lies in its own block or in one of its
// block's partner blocks, else j is, and i received G(d_ji) F_j.
__device__ __forceinline__ void img_correct(const RecC* __restrict__ rec, int n, int B, int i,
                                            const int* __restrict__ partner, double* m) {
  // the three partners in ascending order, INT_MAX for none
  int js[3];
  for (int a = 0; a < 3; ++a) {
    const int j = partner[3 * (size_t)i + a];
    js[a] = (j < 0 || j >= n || j == i) ? INT_MAX : j;
  }
  auto order = [](int& u, int& v) {
    const int lo = min(u, v), hi = max(u, v);
    u = lo;
    v = hi;
  };
  order(js[0], js[1]);
  order(js[1], js[2]);
  order(js[0], js[1]);
  if (js[0] == INT_MAX) return;
  const RecC ri = rec[i];
  const int bi = i / kSymB;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int j = js[k];
    if (j == INT_MAX || (k > 0 && j == js[k - 1])) continue;  // none, or found on two axes
    const RecC rj = rec[j];
    const int bj = j / kSymB;
    const int e = (bj - bi + B) % B;
    const bool i_target = bi == bj || (e >= 1 && e <= sym_partners(bi, B));
    double d32[3], d64[3];
    for (int a = 0; a < 3; ++a) {
      const uint32_t hi = rec_hi(ri, a), hj = rec_hi(rj, a);
      d32[a] = (double)dem::diff32(i_target ? hj : hi, i_target ? hi : hj);
      const unsigned long long qi = ((unsigned long long)hi << 32) | rec_lo(ri, a);
      const unsigned long long qj = ((unsigned long long)hj << 32) | rec_lo(rj, a);
      d64[a] = (double)dem::diff64n(qj, 0ull - qi);
    }
    const double f[3] = {(double)rj.fx, (double)rj.fy, (double)rj.fz};
    double t64[3], t32[3];
    oseen_term64(d64, f, t64);
    oseen_term64(d32, f, t32);
    for (int a = 0; a < 3; ++a) m[a] += t64[a] - t32[a];
  }
}

// Boundary-pair tables of one call (null img_slot: not used).
struct ImgArgs {
  unsigned long long* slot;
  int* partner;
  int* flag;
  int log2m;
};

}  // namespace dem
