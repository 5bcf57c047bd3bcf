// Batched ensembles: one global step of many independent periodic droplet
// systems in five launches, capturable as one HIP graph.
//
// Reference semantics, per system (dem/solver/euler_droplet_solver.py:43-159
// driven system after system by main_multi.py:43-71):
//   1. contact detection (SpatialHashCollisionDetector.detect,
//      spatial_hash_collision_detector.py:233-316); on contact the step is
//      abandoned, the caller merges (droplet_solution.py:208-287) and the
//      step is retried;
//   2. forces and Oseen velocities (direct_droplet_force_calculator.py:
//      470-527);
//   3. x += (F/(stokes_factor r) + v) dt, then x %= L (:113-131).
//
// MI355X design:
//   * All systems live in concatenated arrays (system e owns droplets
//     [offset_e, offset_e + capacity_e)).  Work is described by two tables
//     rebuilt on the host only when a merge changes a system's size: items
//     (system, first target, source split) for the O(N^2) sums, and rows
//     (system, first droplet) of 256 droplets for per-droplet work.  Launch
//     grids are sized to the capacities and read the live counts from device
//     memory, so one captured graph stays valid as systems shrink.
//   * Launches: k_b_contact (one 1024-thread workgroup per system: bucket
//     counts, block scan, scatter and the 27-cell search of the reference's
//     hash, reduced to "any contact"), k_b_force (items), k_b_force_reduce
//     (rows), k_b_conv (items), k_b_conv_euler (rows: Oseen reduction, Euler
//     step, wrap and re-quantisation for the next step).
//   * Items are ordered by decreasing cost so the long ones start first.
//   * The per-target sums use the single-system functions (direct_core.h)
//     with the same source split, so results equal the single-system path's.
#include <math.h>

#include <stddef.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <vector>

#include "dem_common.h"
#include "direct_core.h"
#include "euler_core.h"
#include "hash_core.h"
#include "profile.h"

namespace {

using dem::DevScales;
using dem::Rec;
using dem::RecC;

constexpr int kRow = 256;        // droplets per row
constexpr int kContactWG = 1024; // threads of a contact workgroup

inline int split_max(int n) {  // same cap as the single-system kernels
  long long cap = (1ll << 21) / (n > 0 ? n : 1);
  long long s = (n + 63) / 64;
  if (cap < s) s = cap;
  if (s > 64) s = 64;
  if (s < 1) s = 1;
  return (int)s;
}

inline int batch_split(int n, int jlen) {
  if (n <= 0 || jlen <= 0) return 1;
  long long s = ((long long)n + jlen - 1) / jlen;
  int smax = split_max(n);
  if (s > smax) s = smax;
  if (s < 1) s = 1;
  return (int)s;
}

// Largest split of any size n <= cap: S(n) <= min(64, ceil(cap / jlen)).
inline int split_bound(int cap, int jlen) {
  if (cap <= 0 || jlen <= 0) return 1;
  long long s = ((long long)cap + jlen - 1) / jlen;
  return (int)(s > 64 ? 64 : (s < 1 ? 1 : s));
}

inline int next_pow2(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

// Device description of one system.
struct BEns {
  DevScales sc;
  long long part_off;  // doubles into the partial-sum array
  long long hash_off;  // ints into the hash scratch
  int off, n, cap, S;
  int jchunk, steps, T, pad;
  double L, dt, sf, m_const, eta_const;
};

struct BHead {
  int n_items;
  int n_rows;
  int n_ens;
  int pad;
};

struct BLayout {
  BEns* ens;
  BHead* head;
  int4* items;  // (e, i0, s, 0)
  int2* rows;   // (e, i0)
  Rec* rec;
  RecC* recc;
  double* part;
  int* hash;    // per system: start[T+1], cnt[T], rank[cap], sorted[cap]
  int* contact;
  int* adv;
  int* steps_done;
  double* time;
  size_t bytes;
};

struct Plan {
  long long items_cap = 0, rows_cap = 0, part_total = 0, hash_total = 0, n_total = 0;
  std::vector<long long> part_off, hash_off;
};

Plan make_plan(const dem_batch_ensemble* ens, int n_ens, int block, int jlen) {
  Plan p;
  p.part_off.resize(n_ens);
  p.hash_off.resize(n_ens);
  for (int e = 0; e < n_ens; ++e) {
    const int cap = ens[e].capacity;
    // a system only shrinks: bound its items and partials over every n <= cap
    const int S = split_bound(cap, jlen);
    p.items_cap += (long long)((cap + block - 1) / block) * S;
    p.rows_cap += (cap + kRow - 1) / kRow;
    p.part_off[e] = p.part_total;
    p.part_total += (long long)S * 6 * cap;
    const int T = next_pow2(2 * (cap > 0 ? cap : 1));
    p.hash_off[e] = p.hash_total;
    p.hash_total += 2LL * T + 1 + 2LL * cap;
    long long end = (long long)ens[e].offset + cap;
    if (end > p.n_total) p.n_total = end;
  }
  return p;
}

BLayout carve(void* ws, int n_ens, const Plan& p) {
  dem::Carver c(ws);
  BLayout l;
  const size_t ne = (size_t)(n_ens > 0 ? n_ens : 1);
  l.ens = c.take<BEns>(ne);
  l.head = c.take<BHead>(1);
  l.items = c.take<int4>((size_t)std::max<long long>(p.items_cap, 1));
  l.rows = c.take<int2>((size_t)std::max<long long>(p.rows_cap, 1));
  l.rec = c.take<Rec>((size_t)std::max<long long>(p.n_total, 1));
  l.recc = c.take<RecC>((size_t)std::max<long long>(p.n_total, 1));
  l.part = c.take<double>((size_t)std::max<long long>(p.part_total, 1));
  l.hash = c.take<int>((size_t)std::max<long long>(p.hash_total, 1));
  l.contact = c.take<int>(ne);
  l.adv = c.take<int>(ne);
  l.steps_done = c.take<int>(ne);
  l.time = c.take<double>(ne);
  l.bytes = dem::align_up(c.used, 256);
  return l;
}

// ---------------------------------------------------------------- set-up kernels

__global__ void k_b_init(BEns* __restrict__ ens, int n_ens, int* __restrict__ contact,
                         int* __restrict__ adv, int* __restrict__ steps_done,
                         double* __restrict__ time) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_ens) return;
  ens[e].sc = dem::periodic_scales(ens[e].L, ens[e].m_const);
  contact[e] = 0;
  adv[e] = 0;
  steps_done[e] = 0;
  time[e] = 0.0;
}

// Quantise the droplets of every row (only_e < 0) or of system only_e.
__global__ __launch_bounds__(kRow) void k_b_pack(const BEns* __restrict__ ens,
                                                 const int2* __restrict__ rows,
                                                 const BHead* __restrict__ head, int only_e,
                                                 const double* __restrict__ pos,
                                                 const double* __restrict__ radii,
                                                 Rec* __restrict__ rec, RecC* __restrict__ recc) {
  if ((int)blockIdx.x >= head->n_rows) return;
  const int2 row = rows[blockIdx.x];
  if (only_e >= 0 && row.x != only_e) return;
  const BEns& E = ens[row.x];
  const int i = row.y + threadIdx.x;
  if (i >= E.n) return;
  const size_t g = (size_t)E.off + i;
  Rec o;
  RecC c;
  dem::pack_one(pos + 3 * g, radii[g], nullptr, E.sc, o, c);
  rec[g] = o;
  recc[g] = c;
}

// ---------------------------------------------------------------- 1. contacts

// Inclusive block scan of one int per thread (blockDim.x == kContactWG).
__device__ int block_inclusive_scan(int v, int* sh) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    int u = __shfl_up(v, d, 64);
    if (lane >= d) v += u;
  }
  if (lane == 63) sh[wave] = v;
  __syncthreads();
  if (wave == 0) {
    int w = lane < kContactWG / 64 ? sh[lane] : 0;
#pragma unroll
    for (int d = 1; d < kContactWG / 64; d <<= 1) {
      int u = __shfl_up(w, d, 64);
      if (lane >= d) w += u;
    }
    if (lane < kContactWG / 64) sh[lane] = w;
  }
  __syncthreads();
  const int add = wave > 0 ? sh[wave - 1] : 0;
  __syncthreads();
  return v + add;
}

// One workgroup per system: does the reference's hash pass find any pair?
// Decides whether the system advances this step (active and no contact).
__global__ __launch_bounds__(kContactWG) void k_b_contact(const BEns* __restrict__ ens,
                                                          const BHead* __restrict__ head,
                                                          const double* __restrict__ pos,
                                                          const double* __restrict__ radii,
                                                          int* __restrict__ hash,
                                                          int* __restrict__ contact,
                                                          int* __restrict__ adv,
                                                          const int* __restrict__ steps_done) {
  __shared__ double s_max[kContactWG / 64];
  __shared__ int s_scan[kContactWG / 64];
  const int e = blockIdx.x;
  if (e >= head->n_ens) return;
  const BEns& E = ens[e];
  const int n = E.n;
  const bool active = steps_done[e] < E.steps && n > 0;
  if (!active) {
    if (threadIdx.x == 0) {
      contact[e] = 0;
      adv[e] = 0;
    }
    return;
  }
  const double* p = pos + 3 * (size_t)E.off;
  const double* r = radii + E.off;
  // cell = 2 max r (spatial_hash...py:257-258), g = max(1, ceil(L/cell)) (:262)
  double m = -INFINITY;
  for (int i = threadIdx.x; i < n; i += kContactWG) m = fmax(m, r[i]);
  for (int d = 32; d > 0; d >>= 1) m = fmax(m, __shfl_down(m, d, 64));
  if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = m;
  __syncthreads();
  m = s_max[0];
  for (int w = 1; w < kContactWG / 64; ++w) m = fmax(m, s_max[w]);
  const double cell = 2.0 * m;
  if (!(cell > 0.0)) {  // no contacts (:259-260)
    if (threadIdx.x == 0) {
      contact[e] = 0;
      adv[e] = 1;
    }
    return;
  }
  double q = ceil(E.L / cell);
  if (!(q >= 1.0)) q = 1.0;
  if (q > 1073741824.0) q = 1073741824.0;
  const int g = (int)q;
  const int T = E.T, mask = T - 1;
  int* start = hash + E.hash_off;  // [T + 1]
  int* cnt = start + T + 1;        // [T]
  int* rank = cnt + T;             // [cap]
  int* sorted = rank + E.cap;      // [cap]
  for (int b = threadIdx.x; b < T; b += kContactWG) cnt[b] = 0;
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += kContactWG) {
    const int h = dem::hash_cell(dem::cell_of(p[3 * (size_t)i], cell, g, 1),
                                 dem::cell_of(p[3 * (size_t)i + 1], cell, g, 1),
                                 dem::cell_of(p[3 * (size_t)i + 2], cell, g, 1), mask);
    rank[i] = atomicAdd(&cnt[h], 1);
  }
  __syncthreads();
  // exclusive scan of cnt into start: contiguous chunks per thread
  const int chunk = (T + kContactWG - 1) / kContactWG;
  const int b0 = min(T, (int)threadIdx.x * chunk), b1 = min(T, b0 + chunk);
  int local = 0;
  for (int b = b0; b < b1; ++b) local += cnt[b];
  int run = block_inclusive_scan(local, s_scan) - local;
  for (int b = b0; b < b1; ++b) {
    start[b] = run;
    run += cnt[b];
  }
  if (threadIdx.x == kContactWG - 1) start[T] = run;
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += kContactWG) {
    const int h = dem::hash_cell(dem::cell_of(p[3 * (size_t)i], cell, g, 1),
                                 dem::cell_of(p[3 * (size_t)i + 1], cell, g, 1),
                                 dem::cell_of(p[3 * (size_t)i + 2], cell, g, 1), mask);
    sorted[start[h] + rank[i]] = i;
  }
  __syncthreads();
  // 27-cell search over distinct neighbour buckets (:134-229), i < j
  int found = 0;
  for (int i = threadIdx.x; i < n && !found; i += kContactWG) {
    const int ix = dem::cell_of(p[3 * (size_t)i], cell, g, 1);
    const int iy = dem::cell_of(p[3 * (size_t)i + 1], cell, g, 1);
    const int iz = dem::cell_of(p[3 * (size_t)i + 2], cell, g, 1);
    int seen[27];
    int nseen = 0;
    for (int di = -1; di <= 1 && !found; ++di)
      for (int dj = -1; dj <= 1 && !found; ++dj)
        for (int dk = -1; dk <= 1 && !found; ++dk) {
          const int nh = dem::hash_cell(dem::pymod(ix + di, g), dem::pymod(iy + dj, g),
                                        dem::pymod(iz + dk, g), mask);
          bool dup = false;
          for (int k = 0; k < nseen; ++k) dup |= (seen[k] == nh);
          if (dup) continue;
          seen[nseen++] = nh;
          for (int slot = start[nh]; slot < start[nh + 1]; ++slot) {
            const int j = sorted[slot];
            if (i < j && dem::contact(p, r, i, j, E.L, 1)) {
              found = 1;
              break;
            }
          }
        }
  }
  const int any = __syncthreads_or(found);
  if (threadIdx.x == 0) {
    contact[e] = any ? 1 : 0;
    adv[e] = any ? 0 : 1;
  }
}

// ---------------------------------------------------------------- 2-3. forces

template <int BW>
__global__ __launch_bounds__(BW) void k_b_force(const BEns* __restrict__ ens,
                                                const int4* __restrict__ items,
                                                const BHead* __restrict__ head,
                                                const int* __restrict__ adv,
                                                const Rec* __restrict__ rec, int tile,
                                                double* __restrict__ part) {
  if ((int)blockIdx.x >= head->n_items) return;
  const int4 it = items[blockIdx.x];
  if (!adv[it.x]) return;
  const BEns& E = ens[it.x];
  const int n = E.n;
  const Rec* r = rec + E.off;
  const int i = it.y + threadIdx.x;
  const int ii = i < n ? i : n - 1;
  const int s = it.z;
  const int j0 = s * E.jchunk;
  const int j1 = min(n, j0 + E.jchunk);
  double ax = 0.0, ay = 0.0, az = 0.0;
  dem::force_partial(r, r[ii].x, r[ii].y, r[ii].z, r[ii].r, j0, j1, tile, ax, ay, az);
  if (i < n) {
    double* pp = part + E.part_off;
    pp[(size_t)(6 * s + 0) * n + i] = ax;
    pp[(size_t)(6 * s + 1) * n + i] = ay;
    pp[(size_t)(6 * s + 2) * n + i] = az;
  }
}

__global__ __launch_bounds__(kRow) void k_b_force_reduce(const BEns* __restrict__ ens,
                                                         const int2* __restrict__ rows,
                                                         const BHead* __restrict__ head,
                                                         const int* __restrict__ adv,
                                                         const double* __restrict__ part,
                                                         const double* __restrict__ radii,
                                                         double* __restrict__ F,
                                                         RecC* __restrict__ recc) {
  if ((int)blockIdx.x >= head->n_rows) return;
  const int2 row = rows[blockIdx.x];
  if (!adv[row.x]) return;
  const BEns& E = ens[row.x];
  const int n = E.n;
  const int i = row.y + threadIdx.x;
  if (i >= n) return;
  const double* pp = part + E.part_off;
  double sum[3] = {0.0, 0.0, 0.0};
  for (int s = 0; s < E.S; ++s)
    for (int a = 0; a < 3; ++a) sum[a] += pp[(size_t)(6 * s + a) * n + i];
  const size_t g = (size_t)E.off + i;
  double Fi[3];
  float fr[3];
  dem::force_finish(sum, radii[g], E.sc, E.m_const, Fi, fr);
  for (int a = 0; a < 3; ++a) F[3 * g + a] = Fi[a];
  recc[g].fx = fr[0];
  recc[g].fy = fr[1];
  recc[g].fz = fr[2];
}

// ---------------------------------------------------------------- 4-5. Oseen + Euler

template <int BW>
__global__ __launch_bounds__(BW) void k_b_conv(const BEns* __restrict__ ens,
                                               const int4* __restrict__ items,
                                               const BHead* __restrict__ head,
                                               const int* __restrict__ adv,
                                               const RecC* __restrict__ recc, int tile,
                                               double* __restrict__ part) {
  if ((int)blockIdx.x >= head->n_items) return;
  const int4 it = items[blockIdx.x];
  if (!adv[it.x]) return;
  const BEns& E = ens[it.x];
  const int n = E.n;
  const RecC* r = recc + E.off;
  const int i = it.y + threadIdx.x;
  const int ii = i < n ? i : n - 1;
  const int s = it.z;
  const int j0 = s * E.jchunk;
  const int j1 = min(n, j0 + E.jchunk);
  const unsigned long long xi = ((unsigned long long)r[ii].xh << 32) | r[ii].xl;
  const unsigned long long yi = ((unsigned long long)r[ii].yh << 32) | r[ii].yl;
  const unsigned long long zi = ((unsigned long long)r[ii].zh << 32) | r[ii].zl;
  const dem::CorrGrid cg{};
  double m[3] = {0.0, 0.0, 0.0}, c[3] = {0.0, 0.0, 0.0};
  dem::conv_partial<false>(r, xi, yi, zi, cg, j0, j1, tile, m, c);
  if (i < n) {
    double* pp = part + E.part_off;
    for (int a = 0; a < 3; ++a) pp[(size_t)(6 * s + a) * n + i] = m[a];
  }
}

__global__ __launch_bounds__(kRow) void k_b_conv_euler(const BEns* __restrict__ ens,
                                                       const int2* __restrict__ rows,
                                                       const BHead* __restrict__ head,
                                                       const int* __restrict__ adv,
                                                       const double* __restrict__ part,
                                                       const double* __restrict__ radii,
                                                       const double* __restrict__ F,
                                                       double* __restrict__ V,
                                                       double* __restrict__ pos,
                                                       Rec* __restrict__ rec,
                                                       RecC* __restrict__ recc,
                                                       int* __restrict__ steps_done,
                                                       double* __restrict__ time) {
  if ((int)blockIdx.x >= head->n_rows) return;
  const int2 row = rows[blockIdx.x];
  if (!adv[row.x]) return;
  const BEns& E = ens[row.x];
  const int n = E.n;
  const int i = row.y + threadIdx.x;
  if (row.y == 0 && threadIdx.x == 0) {
#pragma clang fp contract(off)
    steps_done[row.x] += 1;
    time[row.x] = time[row.x] + E.dt;
  }
  if (i >= n) return;
  const double* pp = part + E.part_off;
  double m[3] = {0.0, 0.0, 0.0};
  const double c[3] = {0.0, 0.0, 0.0};
  for (int s = 0; s < E.S; ++s)
    for (int a = 0; a < 3; ++a) m[a] += pp[(size_t)(6 * s + a) * n + i];
  const size_t g = (size_t)E.off + i;
  dem::CorrScalars cs{};
  double Vi[3];
  dem::conv_finish(m, c, E.sc, E.eta_const, cs, F + 3 * g, Vi);
  const double ri = radii[g];
  double x[3];
  for (int a = 0; a < 3; ++a) {
    V[3 * g + a] = Vi[a];
    x[a] = dem::euler_coord(pos[3 * g + a], F[3 * g + a], Vi[a], true, ri, E.sf, E.dt, E.L, 1);
    pos[3 * g + a] = x[a];
  }
  Rec o;
  RecC cc;
  dem::pack_one(x, ri, nullptr, E.sc, o, cc);
  rec[g] = o;
  recc[g] = cc;
}

}  // namespace

// ---------------------------------------------------------------- host handle

struct dem_batch {
  int n_ens = 0;
  dem_batch_tuning tune{};
  std::vector<dem_batch_ensemble> desc;
  Plan plan;
  BLayout l{};
  double *pos = nullptr, *radii = nullptr, *F = nullptr, *V = nullptr;
  std::vector<int4> items_h;
  std::vector<int2> rows_h;
  hipStream_t cap_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

int check_tuning(const dem_batch_tuning* t) {
  if (!t) return dem::set_error(DEM_ERR_INVALID, "tuning is NULL");
  if (t->block != 64 && t->block != 128 && t->block != 256)
    return dem::set_error(DEM_ERR_INVALID, "tuning.block must be 64, 128 or 256 (got %d)", t->block);
  if (t->force_tile < 0 || t->conv_tile < 0)
    return dem::set_error(DEM_ERR_INVALID, "tiles must be >= 0");
  return DEM_OK;
}

int check_ensembles(const dem_batch_ensemble* ens, int n_ens) {
  if (n_ens <= 0 || !ens) return dem::set_error(DEM_ERR_INVALID, "need at least one ensemble");
  for (int e = 0; e < n_ens; ++e) {
    const dem_batch_ensemble& d = ens[e];
    if (d.offset < 0 || d.capacity < 0 || d.n < 0 || d.n > d.capacity)
      return dem::set_error(DEM_ERR_INVALID, "ensemble %d: bad offset/n/capacity", e);
    if (d.capacity > (1 << 24))
      return dem::set_error(DEM_ERR_CAPACITY, "ensemble %d: capacity %d too large", e, d.capacity);
    if (!(d.L > 0.0)) return dem::set_error(DEM_ERR_INVALID, "ensemble %d: L must be > 0", e);
    if (d.steps < 0) return dem::set_error(DEM_ERR_INVALID, "ensemble %d: steps < 0", e);
    for (int f = 0; f < e; ++f) {  // reserved ranges must not overlap
      const dem_batch_ensemble& o = ens[f];
      const long long a0 = d.offset, a1 = (long long)d.offset + d.capacity;
      const long long b0 = o.offset, b1 = (long long)o.offset + o.capacity;
      if (a0 < b1 && b0 < a1 && d.capacity > 0 && o.capacity > 0)
        return dem::set_error(DEM_ERR_INVALID, "ensembles %d and %d overlap", f, e);
    }
  }
  return DEM_OK;
}

// Rebuilds items / rows / device system table from the host descriptions.
int upload_tables(dem_batch* b, hipStream_t st, bool full) {
  const int bw = b->tune.block;
  std::vector<BEns> eh(b->n_ens);
  struct Cost {
    long long c;
    int4 it;
  };
  std::vector<Cost> items;
  b->rows_h.clear();
  for (int e = 0; e < b->n_ens; ++e) {
    const dem_batch_ensemble& d = b->desc[e];
    BEns& E = eh[e];
    E = BEns{};
    E.part_off = b->plan.part_off[e];
    E.hash_off = b->plan.hash_off[e];
    E.off = d.offset;
    E.n = d.n;
    E.cap = d.capacity;
    E.S = batch_split(d.n, b->tune.jlen);
    E.jchunk = d.n > 0 ? (d.n + E.S - 1) / E.S : 0;
    E.steps = d.steps;
    E.T = next_pow2(2 * (d.capacity > 0 ? d.capacity : 1));
    E.L = d.L;
    E.dt = d.dt;
    E.sf = d.stokes_factor;
    E.m_const = d.m_const;
    E.eta_const = d.eta_const;
    for (int i0 = 0; i0 < d.n; i0 += bw)
      for (int s = 0; s < E.S; ++s) {
        const int j0 = s * E.jchunk, j1 = std::min(d.n, j0 + E.jchunk);
        if (j1 <= j0) continue;
        const long long cost = (long long)std::min(bw, d.n - i0) * (j1 - j0);
        items.push_back({cost, make_int4(e, i0, s, 0)});
      }
    for (int i0 = 0; i0 < d.n; i0 += kRow) b->rows_h.push_back(make_int2(e, i0));
  }
  std::stable_sort(items.begin(), items.end(),
                   [](const Cost& a, const Cost& c) { return a.c > c.c; });
  b->items_h.resize(items.size());
  for (size_t k = 0; k < items.size(); ++k) b->items_h[k] = items[k].it;
  if ((long long)b->items_h.size() > b->plan.items_cap || (long long)b->rows_h.size() > b->plan.rows_cap)
    return dem::set_error(DEM_ERR_CAPACITY, "work tables exceed their capacity");
  BHead h{(int)b->items_h.size(), (int)b->rows_h.size(), b->n_ens, 0};
  if (full) {
    DEM_CHECK_HIP(hipMemcpyAsync(b->l.ens, eh.data(), sizeof(BEns) * eh.size(),
                                 hipMemcpyHostToDevice, st));
  } else {
    // keep the device-computed scales: copy everything after them
    for (int e = 0; e < b->n_ens; ++e) {
      const size_t skip = offsetof(BEns, part_off);
      DEM_CHECK_HIP(hipMemcpyAsync(reinterpret_cast<char*>(b->l.ens + e) + skip,
                                   reinterpret_cast<const char*>(&eh[e]) + skip,
                                   sizeof(BEns) - skip, hipMemcpyHostToDevice, st));
    }
  }
  if (!b->items_h.empty())
    DEM_CHECK_HIP(hipMemcpyAsync(b->l.items, b->items_h.data(), sizeof(int4) * b->items_h.size(),
                                 hipMemcpyHostToDevice, st));
  if (!b->rows_h.empty())
    DEM_CHECK_HIP(hipMemcpyAsync(b->l.rows, b->rows_h.data(), sizeof(int2) * b->rows_h.size(),
                                 hipMemcpyHostToDevice, st));
  DEM_CHECK_HIP(hipMemcpyAsync(b->l.head, &h, sizeof(h), hipMemcpyHostToDevice, st));
  // host tables are reused by the next update: wait for the copies
  DEM_CHECK_HIP(hipStreamSynchronize(st));
  return DEM_OK;
}

int tile_force(const dem_batch* b) { return b->tune.force_tile > 0 ? b->tune.force_tile : 32; }
int tile_conv(const dem_batch* b) { return b->tune.conv_tile > 0 ? b->tune.conv_tile : 16; }

template <int BW>
int launch_items(dem_batch* b, hipStream_t st, bool prof) {
  const BLayout& l = b->l;
  const int gi = (int)std::max<long long>(b->plan.items_cap, 1);
  const int gr = (int)std::max<long long>(b->plan.rows_cap, 1);
  if (prof) {
    DEM_PROF_LAUNCH("k_b_force", st,
                    hipLaunchKernelGGL(k_b_force<BW>, dim3(gi), dim3(BW), 0, st,
                                       (const BEns*)l.ens, (const int4*)l.items,
                                       (const BHead*)l.head, (const int*)l.adv,
                                       (const Rec*)l.rec, tile_force(b), l.part));
  } else {
    hipLaunchKernelGGL(k_b_force<BW>, dim3(gi), dim3(BW), 0, st, (const BEns*)l.ens,
                       (const int4*)l.items, (const BHead*)l.head, (const int*)l.adv,
                       (const Rec*)l.rec, tile_force(b), l.part);
  }
  DEM_CHECK_LAUNCH("k_b_force");
  hipLaunchKernelGGL(k_b_force_reduce, dim3(gr), dim3(kRow), 0, st, (const BEns*)l.ens,
                     (const int2*)l.rows, (const BHead*)l.head, (const int*)l.adv,
                     (const double*)l.part, (const double*)b->radii, b->F, l.recc);
  DEM_CHECK_LAUNCH("k_b_force_reduce");
  if (prof) {
    DEM_PROF_LAUNCH("k_b_conv", st,
                    hipLaunchKernelGGL(k_b_conv<BW>, dim3(gi), dim3(BW), 0, st,
                                       (const BEns*)l.ens, (const int4*)l.items,
                                       (const BHead*)l.head, (const int*)l.adv,
                                       (const RecC*)l.recc, tile_conv(b), l.part));
  } else {
    hipLaunchKernelGGL(k_b_conv<BW>, dim3(gi), dim3(BW), 0, st, (const BEns*)l.ens,
                       (const int4*)l.items, (const BHead*)l.head, (const int*)l.adv,
                       (const RecC*)l.recc, tile_conv(b), l.part);
  }
  DEM_CHECK_LAUNCH("k_b_conv");
  hipLaunchKernelGGL(k_b_conv_euler, dim3(gr), dim3(kRow), 0, st, (const BEns*)l.ens,
                     (const int2*)l.rows, (const BHead*)l.head, (const int*)l.adv,
                     (const double*)l.part, (const double*)b->radii, (const double*)b->F, b->V,
                     b->pos, l.rec, l.recc, l.steps_done, l.time);
  DEM_CHECK_LAUNCH("k_b_conv_euler");
  return DEM_OK;
}

int enqueue_step(dem_batch* b, hipStream_t st, bool prof) {
  const BLayout& l = b->l;
  if (prof) {
    DEM_PROF_LAUNCH("k_b_contact", st,
                    hipLaunchKernelGGL(k_b_contact, dim3(b->n_ens), dim3(kContactWG), 0, st,
                                       (const BEns*)l.ens, (const BHead*)l.head,
                                       (const double*)b->pos, (const double*)b->radii, l.hash,
                                       l.contact, l.adv, (const int*)l.steps_done));
  } else {
    hipLaunchKernelGGL(k_b_contact, dim3(b->n_ens), dim3(kContactWG), 0, st, (const BEns*)l.ens,
                       (const BHead*)l.head, (const double*)b->pos, (const double*)b->radii,
                       l.hash, l.contact, l.adv, (const int*)l.steps_done);
  }
  DEM_CHECK_LAUNCH("k_b_contact");
  switch (b->tune.block) {
    case 64: return launch_items<64>(b, st, prof);
    case 128: return launch_items<128>(b, st, prof);
    default: return launch_items<256>(b, st, prof);
  }
}

void drop_graph(dem_batch* b) {
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

}  // namespace

extern "C" {

int dem_batch_split(int n, int jlen) { return batch_split(n, jlen); }

size_t dem_batch_workspace_bytes(const dem_batch_ensemble* ens, int n_ens,
                                 const dem_batch_tuning* tuning) {
  if (check_ensembles(ens, n_ens) || check_tuning(tuning)) return 0;
  Plan p = make_plan(ens, n_ens, tuning->block, tuning->jlen);
  return carve(nullptr, n_ens, p).bytes;
}

int dem_batch_create(const dem_batch_ensemble* ens, int n_ens, const dem_batch_tuning* tuning,
                     double* pos, double* radii, double* forces, double* velocities, void* ws,
                     size_t ws_bytes, dem_stream_t stream, dem_batch_t* out) {
  if (!out) return dem::set_error(DEM_ERR_INVALID, "out is NULL");
  *out = nullptr;
  int rc = check_ensembles(ens, n_ens);
  if (rc) return rc;
  if ((rc = check_tuning(tuning))) return rc;
  if (!pos || !radii || !forces || !velocities)
    return dem::set_error(DEM_ERR_INVALID, "null state array");
  Plan p = make_plan(ens, n_ens, tuning->block, tuning->jlen);
  if (p.items_cap > 0x7fffffffLL || p.rows_cap > 0x7fffffffLL)
    return dem::set_error(DEM_ERR_CAPACITY, "too many work items");
  BLayout l = carve(ws, n_ens, p);
  if (!ws || ws_bytes < l.bytes)
    return dem::set_error(DEM_ERR_WORKSPACE, "workspace %zu < %zu bytes", ws_bytes, l.bytes);
  dem_batch* b = new (std::nothrow) dem_batch();
  if (!b) return dem::set_error(DEM_ERR_CAPACITY, "out of host memory");
  b->n_ens = n_ens;
  b->tune = *tuning;
  b->desc.assign(ens, ens + n_ens);
  b->plan = p;
  b->l = l;
  b->pos = pos;
  b->radii = radii;
  b->F = forces;
  b->V = velocities;
  hipStream_t st = dem::as_stream(stream);
  rc = upload_tables(b, st, true);
  if (!rc) {
    hipLaunchKernelGGL(k_b_init, dim3((n_ens + 255) / 256), dim3(256), 0, st, l.ens, n_ens,
                       l.contact, l.adv, l.steps_done, l.time);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = dem::set_error(DEM_ERR_HIP, "k_b_init: %s", hipGetErrorString(e));
  }
  if (!rc) {
    hipLaunchKernelGGL(k_b_pack, dim3((unsigned)std::max<long long>(p.rows_cap, 1)), dim3(kRow), 0,
                       st, (const BEns*)l.ens, (const int2*)l.rows, (const BHead*)l.head, -1,
                       (const double*)pos, (const double*)radii, l.rec, l.recc);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = dem::set_error(DEM_ERR_HIP, "k_b_pack: %s", hipGetErrorString(e));
  }
  if (rc) {
    delete b;
    return rc;
  }
  *out = b;
  return DEM_OK;
}

int dem_batch_update(dem_batch_t b, int e, int n, dem_stream_t stream) {
  if (!b) return dem::set_error(DEM_ERR_INVALID, "null handle");
  if (e < 0 || e >= b->n_ens) return dem::set_error(DEM_ERR_INVALID, "no ensemble %d", e);
  if (n < 0 || n > b->desc[e].capacity)
    return dem::set_error(DEM_ERR_CAPACITY, "ensemble %d: n=%d outside [0, %d]", e, n,
                          b->desc[e].capacity);
  hipStream_t st = dem::as_stream(stream);
  DEM_CHECK_HIP(hipStreamSynchronize(st));
  b->desc[e].n = n;
  int rc = upload_tables(b, st, false);
  if (rc) return rc;
  hipLaunchKernelGGL(k_b_pack, dim3((unsigned)std::max<long long>(b->plan.rows_cap, 1)),
                     dim3(kRow), 0, st, (const BEns*)b->l.ens, (const int2*)b->l.rows,
                     (const BHead*)b->l.head, e, (const double*)b->pos, (const double*)b->radii,
                     b->l.rec, b->l.recc);
  DEM_CHECK_LAUNCH("k_b_pack");
  return DEM_OK;
}

int dem_batch_step(dem_batch_t b, dem_stream_t stream) {
  if (!b) return dem::set_error(DEM_ERR_INVALID, "null handle");
  return enqueue_step(b, dem::as_stream(stream), true);
}

int dem_batch_capture(dem_batch_t b) {
  if (!b) return dem::set_error(DEM_ERR_INVALID, "null handle");
  drop_graph(b);
  if (!b->cap_stream) DEM_CHECK_HIP(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
  DEM_CHECK_HIP(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeRelaxed));
  int rc = enqueue_step(b, b->cap_stream, false);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
  if (rc) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return dem::set_error(DEM_ERR_HIP, "capture: %s", hipGetErrorString(e));
  b->graph = g;
  e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    drop_graph(b);
    return dem::set_error(DEM_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
  }
  return DEM_OK;
}

int dem_batch_run(dem_batch_t b, int steps, dem_stream_t stream) {
  if (!b) return dem::set_error(DEM_ERR_INVALID, "null handle");
  if (steps < 0) return dem::set_error(DEM_ERR_INVALID, "steps < 0");
  hipStream_t st = dem::as_stream(stream);
  for (int k = 0; k < steps; ++k) {
    if (b->exec) {
      DEM_CHECK_HIP(hipGraphLaunch(b->exec, st));
    } else {
      int rc = enqueue_step(b, st, true);
      if (rc) return rc;
    }
  }
  return DEM_OK;
}

int dem_batch_field(dem_batch_t b, const char* name, void** ptr, size_t* bytes) {
  if (!b || !name || !ptr || !bytes) return dem::set_error(DEM_ERR_INVALID, "null argument");
  const size_t ne = (size_t)b->n_ens;
  struct F {
    const char* n;
    void* p;
    size_t sz;
  } fields[] = {{"contact", b->l.contact, ne * sizeof(int)},
                {"advanced", b->l.adv, ne * sizeof(int)},
                {"steps_done", b->l.steps_done, ne * sizeof(int)},
                {"time", b->l.time, ne * sizeof(double)}};
  for (const F& f : fields)
    if (strcmp(f.n, name) == 0) {
      *ptr = f.p;
      *bytes = f.sz;
      return DEM_OK;
    }
  return dem::set_error(DEM_ERR_INVALID, "no batch field '%s'", name);
}

int dem_batch_destroy(dem_batch_t b) {
  if (!b) return DEM_OK;
  drop_graph(b);
  if (b->cap_stream) hipStreamDestroy(b->cap_stream);
  delete b;
  return DEM_OK;
}

}  // extern "C"
