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
// MI355X design: counting sort of particles by bucket (atomic counts, device
// scan, atomic scatter with a bucket-ordered copy of (x, y, z, r)), then two
// passes over targets in bucket order: count pairs per i, scan, emit.  Each target visits each distinct bucket once, so pairs are
// unique without a global sort; each target's pairs land in its own segment,
// sorted by j in-segment, which makes the output lexicographic and
// deterministic.  The contact predicate is evaluated in float64 with FMA
// contraction disabled so that it rounds like the reference's element-wise
// float64 arithmetic.
#include <math.h>

#include <mutex>

#include "dem_common.h"
#include "hash_core.h"
#include "scan.h"

namespace {

using dem::cell_of;
using dem::contact_pair;
using dem::hash_cell;

struct HashParams {
  double cell;  // 2 * max radius
  int g;        // cells per axis
  int skip;     // 1: cell <= 0, no contacts
};

struct Layout {
  HashParams* hp;
  double* rmax_part;  // per-block max radius
  int* bucket;        // [n]
  int* counts;        // [T]
  int* offsets;       // [T]
  int* cursor;        // [T]
  int* sorted;        // [n] droplet of each slot, bucket order
  double4* spos;      // [n] (x, y, z, r) of each slot
  int* pair_cnt;      // [n]
  int* pair_off;      // [n]
  int* scan_ws;
  size_t bytes;
};

constexpr int kMaxBlocks = 256;
constexpr int kCountSplit = 4;  // lanes per droplet in the count pass

Layout carve(void* ws, int n, int T) {
  dem::Carver c(ws);
  Layout l;
  const size_t nn = (size_t)(n > 0 ? n : 1);
  const int big = T > n ? T : n;
  l.hp = c.take<HashParams>(1);
  l.rmax_part = c.take<double>(kMaxBlocks);
  l.bucket = c.take<int>(nn);
  l.counts = c.take<int>((size_t)T);
  l.offsets = c.take<int>((size_t)T);
  l.cursor = c.take<int>((size_t)T);
  l.sorted = c.take<int>(nn);
  l.spos = c.take<double4>(nn);
  l.pair_cnt = c.take<int>(nn);
  l.pair_off = c.take<int>(nn);
  l.scan_ws = c.take<int>(dem::scan_workspace_ints(big));
  l.bytes = dem::align_up(c.used, 256);
  return l;
}

__global__ __launch_bounds__(256) void k_rmax_partial(const double* __restrict__ r, int n,
                                                      double* __restrict__ part) {
  __shared__ double sh[256];
  double m = -INFINITY;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) m = fmax(m, r[i]);
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// One wave: the largest radius over the partial maxima (fmax is exact and
// order-free), then the cell size and grid.
__global__ void k_hash_setup(const double* __restrict__ part, int nblk, double L,
                             HashParams* __restrict__ hp) {
  double m = -INFINITY;
  for (int b = threadIdx.x; b < nblk; b += 64) m = fmax(m, part[b]);
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  if (threadIdx.x != 0) return;
  HashParams p;
  p.cell = 2.0 * m;
  p.skip = !(p.cell > 0.0);
  double q = p.skip ? 1.0 : ceil(L / p.cell);
  if (!(q >= 1.0)) q = 1.0;
  if (q > 1073741824.0) q = 1073741824.0;
  p.g = (int)q;
  *hp = p;
}

__global__ __launch_bounds__(256) void k_assign(const double* __restrict__ pos, int n, int periodic,
                                                int mask, const HashParams* __restrict__ hpp,
                                                int* __restrict__ bucket, int* __restrict__ counts) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const HashParams hp = *hpp;
  if (hp.skip) return;
  int ix = cell_of(pos[3 * (size_t)i], hp.cell, hp.g, periodic);
  int iy = cell_of(pos[3 * (size_t)i + 1], hp.cell, hp.g, periodic);
  int iz = cell_of(pos[3 * (size_t)i + 2], hp.cell, hp.g, periodic);
  int h = hash_cell(ix, iy, iz, mask);
  bucket[i] = h;
  atomicAdd(&counts[h], 1);
}

__global__ __launch_bounds__(256) void k_scatter(int n, const HashParams* __restrict__ hpp,
                                                 const int* __restrict__ bucket,
                                                 const double* __restrict__ pos,
                                                 const double* __restrict__ r,
                                                 int* __restrict__ cursor, int* __restrict__ sorted,
                                                 double4* __restrict__ spos) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n || hpp->skip) return;
  int slot = atomicAdd(&cursor[bucket[i]], 1);
  sorted[slot] = i;
  spos[slot] = make_double4(pos[3 * (size_t)i], pos[3 * (size_t)i + 1], pos[3 * (size_t)i + 2],
                            r[i]);
}

// EMIT = false: count pairs of each target; EMIT = true: write them.  The
// droplet in bucket slot t is searched by SPLIT consecutive lanes (neighbouring
// slots then search the same buckets), lane `sub` taking the distinct
// neighbour buckets k = sub, sub + SPLIT, ...; their counts are summed across
// the lanes.  The buckets' ranges are loaded up front and candidates come from
// the bucket-ordered (x, y, z, r) copy; the droplet index of a slot is read
// only for a contact.  The emit pass (SPLIT = 1) skips droplets the count
// pass found without pairs.
template <bool EMIT, int SPLIT>
__global__ __launch_bounds__(256) void k_pairs(int n, double L, int periodic, int mask,
                                               const HashParams* __restrict__ hpp,
                                               const int* __restrict__ counts,
                                               const int* __restrict__ offsets,
                                               const int* __restrict__ sorted,
                                               const double4* __restrict__ spos,
                                               int* __restrict__ pair_cnt,
                                               const int* __restrict__ pair_off,
                                               int* __restrict__ pairs, int cap) {
  static_assert(!EMIT || SPLIT == 1, "the emit pass writes a target's pairs from one lane");
  static_assert(SPLIT == 1 || SPLIT == 2 || SPLIT == 4, "lanes per target");
  const int gid = blockIdx.x * 256 + threadIdx.x;
  const int t = gid / SPLIT, sub = gid % SPLIT;
  if (t >= n) return;  // the SPLIT lanes of a target leave together
  const HashParams hp = *hpp;
  if (hp.skip) {  // nothing was bucketed: slot t stands for droplet t
    if (!EMIT && sub == 0) pair_cnt[t] = 0;
    return;
  }
  const int g = hp.g;
  const int i = sorted[t];
  if (EMIT && pair_cnt[i] == 0) return;
  const double4 me = spos[t];
  const int ix = cell_of(me.x, hp.cell, g, periodic);
  const int iy = cell_of(me.y, hp.cell, g, periodic);
  const int iz = cell_of(me.z, hp.cell, g, periodic);
  int nb[27];
  int nseen = 0;
  int cx[3], cy[3], cz[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    cx[d] = dem::nb_coord(ix, d - 1, g, periodic);
    cy[d] = dem::nb_coord(iy, d - 1, g, periodic);
    cz[d] = dem::nb_coord(iz, d - 1, g, periodic);
  }
#pragma unroll
  for (int k = 0; k < 27; ++k) {
    const int nx = cx[k / 9], ny = cy[(k / 3) % 3], nz = cz[k % 3];
    const bool valid = nx >= 0 && ny >= 0 && nz >= 0;
    if (valid) {
      const int nh = hash_cell(nx, ny, nz, mask);
      bool dup = false;
#pragma unroll
      for (int q = 0; q < k; ++q) dup |= (q < nseen && nb[q] == nh);
      if (!dup) nb[nseen++] = nh;
    }
  }
  constexpr int KS = (27 + SPLIT - 1) / SPLIT;
  int s0[KS], s1[KS];
#pragma unroll
  for (int u = 0; u < KS; ++u) {
    const int k = sub + SPLIT * u;
    const bool ok = k < nseen;
    const int h = ok ? nb[k < 27 ? k : 0] : 0;
    s0[u] = offsets[h];
    s1[u] = ok ? s0[u] + counts[h] : s0[u];
  }
  int cnt = 0;
  const int base = EMIT ? pair_off[i] : 0;
#pragma unroll
  for (int u = 0; u < KS; ++u) {
    for (int slot = s0[u]; slot < s1[u]; ++slot) {
      const double4 q = spos[slot];
      if (contact_pair(me.x, me.y, me.z, me.w, q.x, q.y, q.z, q.w, L, periodic)) {
        const int j = sorted[slot];
        if (i < j) {
          if (EMIT) {
            const int at = base + cnt;
            if (at < cap) {
              pairs[2 * (size_t)at] = i;
              pairs[2 * (size_t)at + 1] = j;
            }
          }
          ++cnt;
        }
      }
    }
  }
  if (!EMIT) {
    for (int o = 1; o < SPLIT; o <<= 1) cnt += __shfl_xor(cnt, o);
    if (sub == 0) pair_cnt[i] = cnt;
  } else if (cnt > 1) {
    // In-segment insertion sort by j (segments hold a handful of pairs).
    const int lo = base, hi = min(base + cnt, cap);
    for (int a = lo + 1; a < hi; ++a) {
      int v = pairs[2 * (size_t)a + 1];
      int b = a - 1;
      while (b >= lo && pairs[2 * (size_t)b + 1] > v) {
        pairs[2 * (size_t)(b + 1) + 1] = pairs[2 * (size_t)b + 1];
        --b;
      }
      pairs[2 * (size_t)(b + 1) + 1] = v;
    }
  }
}

__global__ void k_copy(const int* __restrict__ a, int* __restrict__ b, int n) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) b[i] = a[i];
}

__global__ void k_fill(int* p, int n, int v) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

int check_hash_args(const double* pos, const double* radii, int n, int table_size,
                    const int32_t* pairs, int pair_capacity, const int32_t* count_out) {
  if (n < 0) return dem::set_error(DEM_ERR_INVALID, "n must be >= 0");
  if (table_size < 1 || (table_size & (table_size - 1)))
    return dem::set_error(DEM_ERR_INVALID, "table_size must be a power of two (got %d)", table_size);
  if (!count_out) return dem::set_error(DEM_ERR_INVALID, "count_out is NULL");
  if (pair_capacity < 0 || (pair_capacity > 0 && !pairs))
    return dem::set_error(DEM_ERR_INVALID, "bad pair buffer");
  if (n > 0 && (!pos || !radii)) return dem::set_error(DEM_ERR_INVALID, "null positions/radii");
  return DEM_OK;
}

int run_hash(const Layout& l, const double* pos, const double* radii, int n, double L, int periodic,
             int table_size, int32_t* pairs, int pair_capacity, int32_t* count_out,
             hipStream_t st);

// Scratch of the allocation-free entry point: one array per GPU, held until
// the process ends.  Host code never names the array in device work: its
// address on the current device comes from hipGetSymbolAddress, once per
// device, and is kept in g_scratch_base.  Calls that overlap on one GPU share
// the array and race; every call rewrites what it reads (bucket counts are
// cleared, cursors and pair counts are rewritten) before reading it.
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

namespace {

int run_hash(const Layout& l, const double* pos, const double* radii, int n, double L, int periodic,
             int table_size, int32_t* pairs, int pair_capacity, int32_t* count_out,
             hipStream_t st) {
  if (n == 0) {
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(256), 0, st, count_out, 1, 0);
    DEM_CHECK_LAUNCH("k_fill");
    return DEM_OK;
  }
  const int mask = table_size - 1;
  const int T = table_size;
  int nblk = (n + 255) / 256;
  if (nblk > kMaxBlocks) nblk = kMaxBlocks;
  hipLaunchKernelGGL(k_rmax_partial, dim3(nblk), dim3(256), 0, st, radii, n, l.rmax_part);
  DEM_CHECK_LAUNCH("k_rmax_partial");
  hipLaunchKernelGGL(k_hash_setup, dim3(1), dim3(64), 0, st, (const double*)l.rmax_part, nblk, L,
                     l.hp);
  DEM_CHECK_LAUNCH("k_hash_setup");
  hipLaunchKernelGGL(k_fill, dim3((T + 255) / 256), dim3(256), 0, st, l.counts, T, 0);
  DEM_CHECK_LAUNCH("k_fill");
  const int nb = (n + 255) / 256;
  hipLaunchKernelGGL(k_assign, dim3(nb), dim3(256), 0, st, pos, n, periodic ? 1 : 0, mask,
                     (const HashParams*)l.hp, l.bucket, l.counts);
  DEM_CHECK_LAUNCH("k_assign");
  int rc = dem::exclusive_scan(l.counts, l.offsets, T, l.scan_ws, nullptr, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_copy, dim3((T + 255) / 256), dim3(256), 0, st, (const int*)l.offsets,
                     l.cursor, T);
  DEM_CHECK_LAUNCH("k_copy");
  hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(256), 0, st, n, (const HashParams*)l.hp,
                     (const int*)l.bucket, pos, radii, l.cursor, l.sorted, l.spos);
  DEM_CHECK_LAUNCH("k_scatter");
  hipLaunchKernelGGL((k_pairs<false, kCountSplit>), dim3((int)(((long long)n * kCountSplit + 255) / 256)),
                     dim3(256), 0, st, n, L, periodic ? 1 : 0, mask,
                     (const HashParams*)l.hp, (const int*)l.counts, (const int*)l.offsets,
                     (const int*)l.sorted, (const double4*)l.spos, l.pair_cnt,
                     (const int*)l.pair_off, pairs, pair_capacity);
  DEM_CHECK_LAUNCH("k_pairs<count>");
  rc = dem::exclusive_scan(l.pair_cnt, l.pair_off, n, l.scan_ws, count_out, st);
  if (rc) return rc;
  hipLaunchKernelGGL((k_pairs<true, 1>), dim3(nb), dim3(256), 0, st, n, L, periodic ? 1 : 0, mask,
                     (const HashParams*)l.hp, (const int*)l.counts, (const int*)l.offsets,
                     (const int*)l.sorted, (const double4*)l.spos, l.pair_cnt,
                     (const int*)l.pair_off, pairs, pair_capacity);
  DEM_CHECK_LAUNCH("k_pairs<emit>");
  return DEM_OK;
}

}  // namespace
