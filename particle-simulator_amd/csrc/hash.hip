// Spatial-hash contact detection for gfx950.
//
// Reference: dem/collision_detector/spatial_hash_collision_detector.py
//   detect() (:233-316): cell = 2*max(r) (:257-258; cell <= 0 -> no contacts),
//     g = max(1, ceil(L/cell)) (:262), table T = next_pow2(2*capacity) (:42-43);
//   _assign_cells (:78-112): ix = floor(x/cell), then Python-style ix % g
//     (periodic) or clamp to [0, g-1] (open); bucket = hash & (T-1) with
//     hash = ix*73856093 ^ iy*19349663 ^ iz*83492791 in wrapping int32 (:73-76);
//   _detect_collisions (:134-229): for the 27 neighbour cells (periodic wrap,
//     invalid when open) scan every particle j of the neighbour's bucket, keep
//     i < j with minimum-image distance sqrt(dx*dx+dy*dy+dz*dz) <= r_i + r_j;
//   host: np.unique(pairs, axis=0) (:305-314).
// The pair set is therefore {(i,j): i<j, contact(i,j), bucket(j) in B(i)} with
// B(i) the buckets of i's 27 neighbour cells — including bucket aliases and
// the periodic miss across a narrow last cell when L/cell is not an integer.
// This file reproduces exactly that set.
//
// MI355X design, seven launches and no host synchronisation:
//   prep     clear the bucket counters, partial maxima of the radii;
//   assign   cell and bucket of every droplet, bucket counts and each
//            droplet's rank in its bucket (atomics);
//   scan     per-tile exclusive scan of the bucket counts (one launch, tiles
//            of 256..4096 entries; consumers add the prefix of the tile sums
//            themselves);
//   scatter  bucket offsets, droplets in bucket order (offset + rank) with a
//            bucket-ordered copy of (x, y, z, r);
//   search   one one-wave workgroup per run of buckets: the droplets of a
//            bucket that share a cell share their neighbour buckets, so the
//            wave loads those buckets' candidates once (contiguous slots) and
//            tests them against every target of the cell; each target's
//            count and its first kStash partners (sorted by j) are kept;
//   scan     per-tile scan of the per-droplet pair counts;
//   emit     each droplet's pairs at its offset (its stash, or a rescan of its
//            neighbourhood when it has more partners than the stash holds);
//            the total goes to *count_out.
// Each target visits each distinct neighbour bucket once, so pairs are unique
// without a global sort; each target's pairs land in its own segment sorted by
// j, which makes the output lexicographic and deterministic whatever order
// the atomics give the slots.  The contact predicate is evaluated in float64
// with FMA contraction disabled so that it rounds like the reference's
// element-wise float64 arithmetic.
#include <math.h>

#include <mutex>

#include "dem_common.h"
#include "hash_core.h"

namespace {

using dem::cell_of;
using dem::contact_pair;
using dem::hash_cell;

struct HashParams {
  double cell;  // 2 * max radius
  int g;        // cells per axis
  int skip;     // 1: cell <= 0, no contacts
};

constexpr int kThreads = 256;
constexpr int kTileItems = 16;                // most entries per thread and scan tile
constexpr int kTile = kThreads * kTileItems;  // largest scan tile
constexpr int kMaxTiles = 8192;               // so table_size and n are at most 2^25
constexpr int kMaxRBlocks = 256;              // partial maxima of the radii
constexpr int kStash = 8;                     // partners kept per target by the search
constexpr int kBucketsPerWave = 4;
constexpr int kMapCap = 256;  // candidates per group mapped to their neighbour in LDS
constexpr int kCandPerLane = 2;  // candidate loads in flight per lane and pass
// One wave per search workgroup: a wave that finds its buckets empty frees
// its LDS at once instead of holding it until the block's slowest wave ends.
constexpr int kSearchThreads = 64;

// Scan tile of n entries: the smallest power-of-two multiple of kThreads, at
// most kTile, whose tile count fits kMaxTiles (as log2; small scans then
// spread over many workgroups).
inline int tile_log2(int n) {
  int lg = 8;  // kThreads entries
  while (lg < 12 && ((n + (1 << lg) - 1) >> lg) > kMaxTiles) ++lg;
  return lg;
}
inline int tiles(int n) { return (n + (1 << tile_log2(n)) - 1) >> tile_log2(n); }

struct Layout {
  HashParams* hp;
  int* oob;           // 1: a coordinate lies outside [0, L] (periodic boxes)
  double* rmax_part;  // [kMaxRBlocks]
  int* bucket;        // [n] bucket of each droplet
  int* rank;          // [n] its place among its bucket's droplets
  int* counts;        // [T] droplets per bucket
  int* offsets;       // [T] first slot of each bucket
  int* local;         // [T] per-tile scan of the counts; reused as the search's
  int* stash;         // stash (= local): kStash partners per droplet
  int* btile;         // [tiles(T)] tile sums of the bucket counts
  int* sorted;        // [n] droplet of each slot, bucket order
  double4* spos;      // [n] (x, y, z, r) of each slot
  int* pair_cnt;      // [n]
  int* pscan;         // [n] per-tile exclusive scan of pair_cnt
  int* ptile;         // [tiles(n)]
  size_t bytes;
};

Layout carve(void* ws, int n, int T) {
  dem::Carver c(ws);
  Layout l;
  const size_t nn = (size_t)(n > 0 ? n : 1);
  const size_t tt = (size_t)T;
  l.hp = c.take<HashParams>(1);
  l.oob = c.take<int>(1);
  l.rmax_part = c.take<double>(kMaxRBlocks);
  l.bucket = c.take<int>(nn);
  l.rank = c.take<int>(nn);
  l.counts = c.take<int>(tt);
  l.offsets = c.take<int>(tt);
  const size_t shared = tt > kStash * nn ? tt : kStash * nn;
  l.local = c.take<int>(shared);
  l.stash = l.local;
  l.btile = c.take<int>((size_t)tiles(T));
  l.sorted = c.take<int>(nn);
  l.spos = c.take<double4>(nn);
  l.pair_cnt = c.take<int>(nn);
  l.pscan = c.take<int>(nn);
  l.ptile = c.take<int>((size_t)tiles((int)nn));
  l.bytes = dem::align_up(c.used, 256);
  return l;
}

// ---------------------------------------------------------------- block helpers

__device__ __forceinline__ int wave_inclusive(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(v, o);
    if (lane >= o) v += u;
  }
  return v;
}

// Exclusive scan of one value per thread over the block (kThreads threads).
__device__ __forceinline__ int block_exclusive(int v, int* wsum, int* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int inc = wave_inclusive(v);
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  int before = 0, all = 0;
#pragma unroll
  for (int k = 0; k < kThreads / 64; ++k) {
    const int s = wsum[k];
    before += k < w ? s : 0;
    all += s;
  }
  __syncthreads();
  *total = all;
  return before + inc - v;
}

// pre[k] = sum of tile_sums[0..k) for k < nt (LDS), and the grand total.
__device__ __forceinline__ int tile_prefix(const int* __restrict__ tile_sums, int nt, int* pre) {
  __shared__ int wsum[kThreads / 64];
  const int per = (nt + kThreads - 1) / kThreads;
  const int b = threadIdx.x * per;
  int sum = 0;
#pragma unroll 4
  for (int k = 0; k < per; ++k)  // (unrolled: the loads go out together)
    if (b + k < nt) sum += tile_sums[b + k];
  int total;
  int run = block_exclusive(sum, wsum, &total);
  for (int k = 0; k < per; ++k)
    if (b + k < nt) {
      const int v = tile_sums[b + k];
      pre[b + k] = run;
      run += v;
    }
  __syncthreads();
  return total;
}

// ---------------------------------------------------------------- kernels

// Clears the bucket counters; blocks b < nrb also write the largest radius
// of droplets b*256 + k*nrb*256 (fmax: exact, order-free).
__global__ __launch_bounds__(kThreads) void k_hash_prep(const double* __restrict__ r, int n, int nrb,
                                                        int* __restrict__ counts, int T,
                                                        double* __restrict__ rmax_part, int* __restrict__ oob) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *oob = 0;
  for (int h = blockIdx.x * kThreads + threadIdx.x; h < T; h += gridDim.x * kThreads) counts[h] = 0;
  if ((int)blockIdx.x >= nrb) return;
  double m = -INFINITY;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += nrb * kThreads) m = fmax(m, r[i]);
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  __shared__ double wm[kThreads / 64];
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kThreads / 64; ++w) m = fmax(m, wm[w]);
    rmax_part[blockIdx.x] = fmax(m, wm[0]);
  }
}

// Cell size and grid from the largest radius (every block derives them; block
// 0 stores them), then the bucket of every droplet, the bucket counts and
// each droplet's rank in its bucket (the counter's value before its add).
__global__ __launch_bounds__(kThreads) void k_hash_assign(const double* __restrict__ pos, int n,
                                                          int periodic, int mask, double L,
                                                          const double* __restrict__ rmax_part,
                                                          int nrb, HashParams* __restrict__ hpo,
                                                          int* __restrict__ bucket,
                                                          int* __restrict__ rank,
                                                          int* __restrict__ counts,
                                                          int* __restrict__ oob) {
  static_assert(kMaxRBlocks <= kThreads, "one partial maximum per thread");
  __shared__ HashParams shp;
  __shared__ double wm[kThreads / 64];
  const int i = blockIdx.x * kThreads + threadIdx.x;
  // the droplet's position loads go out together with the partial maxima's
  double px = 0.0, py = 0.0, pz = 0.0;
  if (i < n) {
    px = pos[3 * (size_t)i];
    py = pos[3 * (size_t)i + 1];
    pz = pos[3 * (size_t)i + 2];
  }
  {
    double m = (int)threadIdx.x < nrb ? rmax_part[threadIdx.x] : -INFINITY;
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int w = 1; w < kThreads / 64; ++w) m = fmax(m, wm[w]);
      HashParams p;
      p.cell = 2.0 * m;
      p.skip = !(p.cell > 0.0);
      double q = p.skip ? 1.0 : ceil(L / p.cell);
      if (!(q >= 1.0)) q = 1.0;
      if (q > 1073741824.0) q = 1073741824.0;
      p.g = (int)q;
      shp = p;
      if (blockIdx.x == 0) *hpo = p;
    }
  }
  __syncthreads();
  const HashParams hp = shp;
  bool out = false;  // (NaN counts as outside)
  if (periodic && i < n)
    out = !(px >= 0.0 && px <= L && py >= 0.0 && py <= L && pz >= 0.0 && pz <= L);
  if (__syncthreads_or(out) && threadIdx.x == 0) atomicOr(oob, 1);
  if (i >= n || hp.skip) return;
  const int ix = cell_of(px, hp.cell, hp.g, periodic);
  const int iy = cell_of(py, hp.cell, hp.g, periodic);
  const int iz = cell_of(pz, hp.cell, hp.g, periodic);
  const int h = hash_cell(ix, iy, iz, mask);
  bucket[i] = h;
  rank[i] = atomicAdd(&counts[h], 1);
}

// Exclusive scan of in[0, n) within tiles of 2^lg entries (out) and the tile
// sums; the global prefix of entry k is out[k] + (sum of tile sums before k's
// tile).  2^lg = kThreads * items, items in {1, 2, ..., kTileItems}.
__global__ __launch_bounds__(kThreads) void k_tile_scan(const int* __restrict__ in,
                                                        int* __restrict__ out,
                                                        int* __restrict__ tile_sums, int n, int lg) {
  // padded LDS: entry e at e + e / 32 (a thread's consecutive entries then
  // fall in distinct banks for neighbouring threads)
  __shared__ int s[kTile + kTile / 32];
  __shared__ int wsum[kThreads / 64];
  const int items = 1 << (lg - 8);
  const int base = blockIdx.x << lg, t = threadIdx.x;
#pragma unroll
  for (int k = 0; k < kTileItems; ++k) {
    const int e = k * kThreads + t;
    if (k < items) s[e + (e >> 5)] = base + e < n ? in[base + e] : 0;
  }
  __syncthreads();
  int v[kTileItems];
  int sum = 0;
#pragma unroll
  for (int k = 0; k < kTileItems; ++k) {
    const int e = t * items + k;
    v[k] = k < items ? s[e + (e >> 5)] : 0;
    sum += v[k];
  }
  int total;
  int run = block_exclusive(sum, wsum, &total);
#pragma unroll
  for (int k = 0; k < kTileItems; ++k) {
    const int e = t * items + k;
    if (k < items) s[e + (e >> 5)] = run;
    run += v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kTileItems; ++k) {
    const int e = k * kThreads + t;
    if (k < items && base + e < n) out[base + e] = s[e + (e >> 5)];
  }
  if (t == 0) tile_sums[blockIdx.x] = total;
}

// Bucket offsets (all T buckets) and the droplets in bucket order (slot =
// bucket offset + rank).  A cleared grid (no buckets) leaves every droplet
// without pairs.
__global__ __launch_bounds__(kThreads) void k_hash_scatter(
    int n, int T, const HashParams* __restrict__ hpp, const int* __restrict__ bucket,
    const int* __restrict__ rank, const double* __restrict__ pos, const double* __restrict__ r,
    const int* __restrict__ local, const int* __restrict__ btile, int ntile, int lg,
    int* __restrict__ offsets, int* __restrict__ sorted, double4* __restrict__ spos,
    int* __restrict__ pair_cnt) {
  extern __shared__ int pre[];
  tile_prefix(btile, ntile, pre);
  const int stride = gridDim.x * kThreads;
  const int g0 = blockIdx.x * kThreads + threadIdx.x;
  for (int h = g0; h < T; h += stride) offsets[h] = local[h] + pre[h >> lg];
  const bool skip = hpp->skip;
  for (int i = g0; i < n; i += stride) {
    if (skip) {
      pair_cnt[i] = 0;
      continue;
    }
    const int h = bucket[i];
    const int slot = local[h] + pre[h >> lg] + rank[i];
    sorted[slot] = i;
    spos[slot] = make_double4(pos[3 * (size_t)i], pos[3 * (size_t)i + 1], pos[3 * (size_t)i + 2],
                              r[i]);
  }
}

__device__ __forceinline__ int readlane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

// Neighbour buckets of a cell: lane k < 27 gets neighbour k's bucket, or -1
// when it is outside an open box.  Two neighbours may share a bucket (a hash
// alias, or the same cell when g <= 2); the reference then visits the bucket
// twice and np.unique drops the repeated pairs.  The search scans every
// neighbour's bucket and counts a pair only in the first neighbour with that
// bucket (tested for candidates in contact only).
__device__ __forceinline__ int neighbour_bucket(int ix, int iy, int iz, int g, int periodic,
                                                int mask) {
  const int lane = threadIdx.x & 63;
  int nh = -1;
  if (lane < 27) {
    const int nx = dem::nb_coord(ix, lane / 9 - 1, g, periodic);
    const int ny = dem::nb_coord(iy, (lane / 3) % 3 - 1, g, periodic);
    const int nz = dem::nb_coord(iz, lane % 3 - 1, g, periodic);
    if (nx >= 0 && ny >= 0 && nz >= 0) nh = hash_cell(nx, ny, nz, mask);
  }
  return nh;
}

// Orders one wave's LDS accesses across its lanes.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct SearchLds {
  double x[64], y[64], z[64], r[64];  // the batch's targets
  int id[64];                         // their droplet indices
  int cnt[64];                        // partners found (i < j)
  int list[64][kStash];               // the first kStash of them
  int start[27];                      // neighbour bucket k: first slot
  int before[27];                     // candidates in neighbours < k
  int bucket[27];                     // neighbour k's bucket (-1: outside an open box)
  unsigned char nbr[kMapCap];         // neighbour of candidate m < kMapCap
};

// One wave per kBucketsPerWave consecutive buckets (their counts and offsets
// probed by one load per lane).  Targets of a bucket are taken 64 at a time
// and grouped by cell (aliased cells of one bucket have different
// neighbourhoods); a group's candidates are the slots of its neighbour
// buckets, kCandPerLane per lane with all loads in flight, and each lane
// tests its candidates against every target of the group (targets in the
// outer loop).  Cross-lane values move by readlane (scalar) where the source
// lane is uniform.
__global__ __launch_bounds__(kSearchThreads) void k_hash_search(
    int T, double L, int periodic, int mask, const HashParams* __restrict__ hpp, const int* __restrict__ oob,
    const int* __restrict__ counts, const int* __restrict__ offsets,
    const int* __restrict__ sorted, const double4* __restrict__ spos, int* __restrict__ pair_cnt,
    int* __restrict__ stash) {
  __shared__ SearchLds lds_all[kSearchThreads / 64];
  SearchLds& S = lds_all[threadIdx.x >> 6];
  const int lane = threadIdx.x & 63;
  const HashParams hp = *hpp;
  if (hp.skip) return;
  const int g = hp.g;
  const bool in_box = *oob == 0;
  const int b0 = (blockIdx.x * (kSearchThreads / 64) + (threadIdx.x >> 6)) * kBucketsPerWave;
  int my_c = 0, my_o = 0;
  if (lane < kBucketsPerWave && b0 + lane < T) {
    my_c = counts[b0 + lane];
    my_o = offsets[b0 + lane];
  }
  unsigned long long busy = __ballot(my_c > 0);
  while (busy) {
    const int kb = __ffsll((long long)busy) - 1;
    busy &= busy - 1;
    const int c = readlane(my_c, kb), o = readlane(my_o, kb);
    for (int t0 = 0; t0 < c; t0 += 64) {
      const int cb = min(64, c - t0);
      int kx = 0, ky = 0, kz = 0;
      if (lane < cb) {
        const double4 q = spos[o + t0 + lane];
        S.x[lane] = q.x;
        S.y[lane] = q.y;
        S.z[lane] = q.z;
        S.r[lane] = q.w;
        S.id[lane] = sorted[o + t0 + lane];
        S.cnt[lane] = 0;
        kx = cell_of(q.x, hp.cell, g, periodic);
        ky = cell_of(q.y, hp.cell, g, periodic);
        kz = cell_of(q.z, hp.cell, g, periodic);
      }
      unsigned long long remaining = __ballot(lane < cb);
      wave_sync();
      while (remaining) {
        const int lead = __ffsll((long long)remaining) - 1;
        const int gx = readlane(kx, lead), gy = readlane(ky, lead), gz = readlane(kz, lead);
        const bool mine = ((remaining >> lane) & 1ull) && kx == gx && ky == gy && kz == gz;
        const unsigned long long members = __ballot(mine);
        remaining &= ~members;
        // With every droplet in [0, L]^3, a group whose cell index lies in
        // [2, g - 4] on every axis (g >= 7) tests contacts without the
        // minimum image.  A droplet in contact with it is within r_i + r_j
        // <= cell (up to rounding) per axis after the image rule; the group's
        // droplets lie in [2 cell, (g - 3) cell] (up to rounding) and the box
        // edge L > (g - 1) cell, so that droplet is on the same side of the
        // seam, at |d_a| <= cell < L/2, where the rule changes nothing.  Any
        // other candidate is at least as far apart unwrapped as wrapped, so
        // out of contact either way.  Same decisions, fewer instructions.
        const bool inner = periodic && in_box && g >= 7 && gx >= 2 && gx <= g - 4 &&
                           gy >= 2 && gy <= g - 4 && gz >= 2 && gz <= g - 4;
        const int per_g = inner ? 0 : periodic;
        // candidates: the slots of the group's neighbour buckets, candidate
        // m in neighbour k for before[k] <= m < before[k] + len[k]
        const int nh = neighbour_bucket(gx, gy, gz, g, periodic, mask);
        const int len = nh >= 0 ? counts[nh] : 0;
        const int first = nh >= 0 ? offsets[nh] : 0;
        const int incl = wave_inclusive(len);  // lanes >= 27 add nothing
        const int before = incl - len;
        const int M = readlane(incl, 63);
        if (lane < 27) {
          S.start[lane] = first;
          S.before[lane] = before;
          S.bucket[lane] = nh;
          for (int t = before; t < before + len && t < kMapCap; ++t) S.nbr[t] = (unsigned char)lane;
        }
        wave_sync();
        // kCandPerLane candidates per lane and pass, all loads in flight
        // before the tests
        for (int m0 = 0; m0 < M; m0 += 64 * kCandPerLane) {
          int slot[kCandPerLane];
#pragma unroll
          for (int u = 0; u < kCandPerLane; ++u) {
            const int m = m0 + 64 * u + lane;
            int k = 0;
            if (m < kMapCap) {
              k = S.nbr[m < M ? m : 0];
            } else {  // the last neighbour whose range starts at or before m
              int hi = 26;
              while (k < hi) {
                const int mid = (k + hi + 1) >> 1;
                if (S.before[mid] <= m) k = mid;
                else hi = mid - 1;
              }
            }
            slot[u] = m < M ? S.start[k] + (m - S.before[k]) : -1;
          }
          double4 cand[kCandPerLane];
#pragma unroll
          for (int u = 0; u < kCandPerLane; ++u)
            if (slot[u] >= 0) cand[u] = spos[slot[u]];
          // targets outer, candidates inner: a target's values are read once
          // and the lane's candidate tests are independent of each other.
          // Contacts are rare: their bookkeeping (the candidate's index, its
          // neighbour and the first-neighbour rule) is recomputed per contact,
          // which keeps the loop at six waves per SIMD.
          unsigned long long mm = members;
          while (mm) {
            const int t = __ffsll((long long)mm) - 1;
            mm &= mm - 1;
            const double tx = S.x[t], ty = S.y[t], tz = S.z[t], tr = S.r[t];
#pragma unroll
            for (int u = 0; u < kCandPerLane; ++u) {
              if (slot[u] < 0) continue;
              const double4 q = cand[u];
              if (!contact_pair(tx, ty, tz, tr, q.x, q.y, q.z, q.w, L, per_g)) continue;
              const int j = sorted[slot[u]];
              if (S.id[t] >= j) continue;
              // counted only in the first neighbour holding this bucket
              const int m = m0 + 64 * u + lane;
              int k = 0;
              if (m < kMapCap) {
                k = S.nbr[m];
              } else {
                int hi = 26;
                while (k < hi) {
                  const int mid = (k + hi + 1) >> 1;
                  if (S.before[mid] <= m) k = mid;
                  else hi = mid - 1;
                }
              }
              const int h = S.bucket[k];
              bool first = true;
              for (int qk = 0; qk < k; ++qk)
                if (S.bucket[qk] == h) first = false;
              if (!first) continue;
              const int p = atomicAdd(&S.cnt[t], 1);
              if (p < kStash) S.list[t][p] = j;
            }
          }
        }
        wave_sync();
      }
      if (lane < cb) {
        const int i = S.id[lane];
        const int cnt = S.cnt[lane];
        pair_cnt[i] = cnt;
        const int kept = min(cnt, kStash);
        int* lst = S.list[lane];
        for (int a = 1; a < kept; ++a) {  // insertion sort by j
          const int v = lst[a];
          int b = a - 1;
          while (b >= 0 && lst[b] > v) {
            lst[b + 1] = lst[b];
            --b;
          }
          lst[b + 1] = v;
        }
        for (int a = 0; a < kept; ++a) stash[(size_t)kStash * i + a] = lst[a];
      }
      wave_sync();
    }
  }
}

// Pairs of droplet i with every candidate j > i of its neighbourhood, written
// to pairs[off ..] (within cap) and sorted by j: the search for a droplet
// with more partners than the stash holds.
__device__ void rescan_droplet(int i, int off, int cnt, const double* __restrict__ pos,
                               const double* __restrict__ r, double L, int periodic, int mask,
                               const HashParams& hp, const int* __restrict__ counts,
                               const int* __restrict__ offsets, const int* __restrict__ sorted,
                               const double4* __restrict__ spos, int* __restrict__ pairs,
                               int cap) {
  const double xi = pos[3 * (size_t)i], yi = pos[3 * (size_t)i + 1], zi = pos[3 * (size_t)i + 2];
  const double ri = r[i];
  const int g = hp.g;
  const int ix = cell_of(xi, hp.cell, g, periodic);
  const int iy = cell_of(yi, hp.cell, g, periodic);
  const int iz = cell_of(zi, hp.cell, g, periodic);
  int nb[27];
  int nseen = 0;
  for (int k = 0; k < 27; ++k) {
    const int nx = dem::nb_coord(ix, k / 9 - 1, g, periodic);
    const int ny = dem::nb_coord(iy, (k / 3) % 3 - 1, g, periodic);
    const int nz = dem::nb_coord(iz, k % 3 - 1, g, periodic);
    if (nx < 0 || ny < 0 || nz < 0) continue;
    const int nh = hash_cell(nx, ny, nz, mask);
    bool dup = false;
    for (int q = 0; q < nseen; ++q) dup |= nb[q] == nh;
    if (!dup) nb[nseen++] = nh;
  }
  int w = 0;
  for (int k = 0; k < nseen; ++k) {
    const int s1 = offsets[nb[k]] + counts[nb[k]];
    for (int slot = offsets[nb[k]]; slot < s1; ++slot) {
      const double4 q = spos[slot];
      if (!contact_pair(xi, yi, zi, ri, q.x, q.y, q.z, q.w, L, periodic)) continue;
      const int j = sorted[slot];
      if (j <= i) continue;
      const int at = off + w++;
      if (at < cap) {
        pairs[2 * (size_t)at] = i;
        pairs[2 * (size_t)at + 1] = j;
      }
    }
  }
  // in-segment insertion sort by j
  const int lo = off, hi = min(off + cnt, cap);
  for (int a = lo + 1; a < hi; ++a) {
    const int v = pairs[2 * (size_t)a + 1];
    int b = a - 1;
    while (b >= lo && pairs[2 * (size_t)b + 1] > v) {
      pairs[2 * (size_t)(b + 1) + 1] = pairs[2 * (size_t)b + 1];
      --b;
    }
    pairs[2 * (size_t)(b + 1) + 1] = v;
  }
}

// Every droplet's pairs at its offset in the lexicographic list; the total to
// *count_out (block 0).
__global__ __launch_bounds__(kThreads) void k_hash_emit(
    int n, const int* __restrict__ pair_cnt, const int* __restrict__ pscan,
    const int* __restrict__ ptile, int ntile, int lg, const int* __restrict__ stash,
    const double* __restrict__ pos, const double* __restrict__ r, double L, int periodic,
    int mask, const HashParams* __restrict__ hpp, const int* __restrict__ counts,
    const int* __restrict__ offsets, const int* __restrict__ sorted,
    const double4* __restrict__ spos, int* __restrict__ pairs, int cap,
    int* __restrict__ count_out) {
  extern __shared__ int pre[];
  const int total = tile_prefix(ptile, ntile, pre);
  // The total is final once the pair counts are scanned: workgroup 0 stores
  // it at once at system scope (to pinned host memory it goes straight out),
  // so a host polling the word learns it while the pairs are still being
  // written — readers of the pairs come after this kernel in stream order.
  if (blockIdx.x == 0 && threadIdx.x == 0)
    __hip_atomic_store(count_out, total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    const int cnt = pair_cnt[i];
    if (cnt == 0) continue;
    const int off = pscan[i] + pre[i >> lg];
    if (cnt <= kStash) {
      for (int a = 0; a < cnt; ++a) {
        const int at = off + a;
        if (at < cap) {
          pairs[2 * (size_t)at] = i;
          pairs[2 * (size_t)at + 1] = stash[(size_t)kStash * i + a];
        }
      }
    } else {
      rescan_droplet(i, off, cnt, pos, r, L, periodic, mask, *hpp, counts, offsets, sorted, spos,
                     pairs, cap);
    }
  }
}

__global__ void k_zero_count(int* p) {
  if (threadIdx.x == 0) __hip_atomic_store(p, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

int check_hash_args(const double* pos, const double* radii, int n, int table_size,
                    const int32_t* pairs, int pair_capacity, const int32_t* count_out) {
  if (n < 0) return dem::set_error(DEM_ERR_INVALID, "n must be >= 0");
  if (table_size < 1 || (table_size & (table_size - 1)))
    return dem::set_error(DEM_ERR_INVALID, "table_size must be a power of two (got %d)", table_size);
  if (table_size > kMaxTiles * kTile || n > kMaxTiles * kTile)
    return dem::set_error(DEM_ERR_INVALID, "n and table_size must be at most %d (got %d, %d)",
                          kMaxTiles * kTile, n, table_size);
  if (!count_out) return dem::set_error(DEM_ERR_INVALID, "count_out is NULL");
  if (pair_capacity < 0 || (pair_capacity > 0 && !pairs))
    return dem::set_error(DEM_ERR_INVALID, "bad pair buffer");
  if (n > 0 && (!pos || !radii)) return dem::set_error(DEM_ERR_INVALID, "null positions/radii");
  return DEM_OK;
}

int run_hash(const Layout& l, const double* pos, const double* radii, int n, double L, int periodic,
             int table_size, int32_t* pairs, int pair_capacity, int32_t* count_out,
             hipStream_t st) {
  if (n == 0) {
    hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(64), 0, st, count_out);
    DEM_CHECK_LAUNCH("k_zero_count");
    return DEM_OK;
  }
  const int T = table_size, mask = T - 1;
  const int per = periodic ? 1 : 0;
  const int nb = (n + kThreads - 1) / kThreads;
  const int nrb = nb < kMaxRBlocks ? nb : kMaxRBlocks;
  int prep_blocks = (T + kThreads - 1) / kThreads;
  if (prep_blocks > 2048) prep_blocks = 2048;
  if (prep_blocks < nrb) prep_blocks = nrb;
  hipLaunchKernelGGL(k_hash_prep, dim3(prep_blocks), dim3(kThreads), 0, st, radii, n, nrb,
                     l.counts, T, l.rmax_part, l.oob);
  DEM_CHECK_LAUNCH("k_hash_prep");
  hipLaunchKernelGGL(k_hash_assign, dim3(nb), dim3(kThreads), 0, st, pos, n, per, mask, L,
                     (const double*)l.rmax_part, nrb, l.hp, l.bucket, l.rank, l.counts, l.oob);
  DEM_CHECK_LAUNCH("k_hash_assign");
  const int lgT = tile_log2(T), ntT = tiles(T);
  hipLaunchKernelGGL(k_tile_scan, dim3(ntT), dim3(kThreads), 0, st, (const int*)l.counts, l.local,
                     l.btile, T, lgT);
  DEM_CHECK_LAUNCH("k_tile_scan<buckets>");
  int sc_blocks = ((n > T ? n : T) + kThreads - 1) / kThreads;
  if (sc_blocks > 4096) sc_blocks = 4096;
  hipLaunchKernelGGL(k_hash_scatter, dim3(sc_blocks), dim3(kThreads), ntT * sizeof(int), st, n, T,
                     (const HashParams*)l.hp, (const int*)l.bucket, (const int*)l.rank, pos,
                     radii, (const int*)l.local, (const int*)l.btile, ntT, lgT, l.offsets,
                     l.sorted, l.spos, l.pair_cnt);
  DEM_CHECK_LAUNCH("k_hash_scatter");
  const int waves = (T + kBucketsPerWave - 1) / kBucketsPerWave;
  const int se_blocks = (waves + kSearchThreads / 64 - 1) / (kSearchThreads / 64);
  hipLaunchKernelGGL(k_hash_search, dim3(se_blocks), dim3(kSearchThreads), 0, st, T, L, per, mask,
                     (const HashParams*)l.hp, (const int*)l.oob, (const int*)l.counts,
                     (const int*)l.offsets,
                     (const int*)l.sorted, (const double4*)l.spos, l.pair_cnt, l.stash);
  DEM_CHECK_LAUNCH("k_hash_search");
  const int lgn = tile_log2(n), ntn = tiles(n);
  hipLaunchKernelGGL(k_tile_scan, dim3(ntn), dim3(kThreads), 0, st, (const int*)l.pair_cnt,
                     l.pscan, l.ptile, n, lgn);
  DEM_CHECK_LAUNCH("k_tile_scan<pairs>");
  int em_blocks = nb < 2048 ? nb : 2048;
  hipLaunchKernelGGL(k_hash_emit, dim3(em_blocks), dim3(kThreads), ntn * sizeof(int), st, n,
                     (const int*)l.pair_cnt, (const int*)l.pscan, (const int*)l.ptile, ntn, lgn,
                     (const int*)l.stash, pos, radii, L, per, mask, (const HashParams*)l.hp,
                     (const int*)l.counts, (const int*)l.offsets, (const int*)l.sorted,
                     (const double4*)l.spos, pairs, pair_capacity, count_out);
  DEM_CHECK_LAUNCH("k_hash_emit");
  return DEM_OK;
}

// Scratch of the allocation-free entry point: one array per GPU, held until
// the process ends.  Host code never names the array in device work: its
// address on the current device comes from hipGetSymbolAddress, once per
// device, and is kept in g_scratch_base.  Calls that overlap on one GPU share
// the array and race; every call rewrites what it reads (bucket counters are
// cleared, offsets, slots, pair counts and stashes are rewritten) before
// reading it.
static_assert(DEM_HASH_NOALLOC_SCRATCH_BYTES % 256 == 0, "scratch size");
static __device__ __align__(256) unsigned char scratch[DEM_HASH_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_base[kMaxDevices] = {};

int scratch_of_current_device(void** out) {
  int dev = -1;
  DEM_CHECK_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return dem::set_error(DEM_ERR_INVALID, "device %d outside the scratch table", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_base[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return dem::set_error(DEM_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                            hipGetErrorString(e));
    g_scratch_base[dev] = p;
  }
  *out = g_scratch_base[dev];
  return DEM_OK;
}

}  // namespace

extern "C" {

size_t dem_hash_workspace_bytes(int n, int table_size) {
  if (n < 0) n = 0;
  if (table_size < 1) table_size = 1;
  return carve(nullptr, n, table_size).bytes;
}

int dem_hash_pairs(const double* pos, const double* radii, int n, double L, int periodic,
                   int table_size, int32_t* pairs, int pair_capacity, int32_t* count_out, void* ws,
                   size_t ws_bytes, dem_stream_t stream) {
  int rc = check_hash_args(pos, radii, n, table_size, pairs, pair_capacity, count_out);
  if (rc) return rc;
  Layout l = carve(ws, n, table_size);
  if (!ws || ws_bytes < l.bytes)
    return dem::set_error(DEM_ERR_WORKSPACE, "workspace %zu < %zu bytes", ws_bytes, l.bytes);
  return run_hash(l, pos, radii, n, L, periodic, table_size, pairs, pair_capacity, count_out,
                  dem::as_stream(stream));
}

int dem_hash_pairs_noalloc(const double* pos, const double* radii, int n, double L, int periodic,
                           int table_size, int32_t* pairs, int pair_capacity, int32_t* count_out,
                           dem_stream_t stream) {
  int rc = check_hash_args(pos, radii, n, table_size, pairs, pair_capacity, count_out);
  if (rc) return rc;
  const size_t need = carve(nullptr, n, table_size).bytes;
  if (need > sizeof(scratch))
    return dem::set_error(DEM_ERR_CAPACITY,
                          "dem_hash_pairs_noalloc: n = %d with table_size = %d needs %zu bytes of "
                          "scratch; the built-in scratch holds %zu (DEM_HASH_NOALLOC_SCRATCH_BYTES)",
                          n, table_size, need, sizeof(scratch));
  void* base = nullptr;
  rc = scratch_of_current_device(&base);
  if (rc) return rc;
  return run_hash(carve(base, n, table_size), pos, radii, n, L, periodic, table_size, pairs,
                  pair_capacity, count_out, dem::as_stream(stream));
}

}  // extern "C"
