// Batched ensembles: many independent periodic droplet systems stepped
// together on one GPU, in two captured HIP graphs per global step.
//
// Reference semantics, per system (dem/solver/euler_droplet_solver.py:43-159,
// driven system after system by main_multi.py:43-71):
//   1. contact detection (SpatialHashCollisionDetector.detect,
//      spatial_hash_collision_detector.py:233-316); on contact the caller
//      merges (droplet_solution.py:208-287) and checks again, without
//      advancing time;
//   2. forces and Oseen velocities (direct_droplet_force_calculator.py:
//      470-527);
//   3. x += (F/(stokes_factor r) + v) dt, then x %= L (:113-131).
//
// MI355X design:
//   * All systems live in concatenated arrays (system e owns droplets
//     [offset_e, offset_e + capacity_e)).  Work is described by two tables
//     rebuilt on the host only when merges change sizes: items (system, first
//     target, source split) for the O(N^2) sums and rows (system, first
//     droplet) of 256 droplets for per-droplet work.  Grids are sized to the
//     capacities and read live counts from device memory, so the captured
//     graphs stay valid as systems shrink.
//   * "check" graphs: the reference's hash pass for every active system at
//     once — bucket counts (row-parallel atomics), one bucket scan per
//     system, scatter (with a bucket-ordered copy of positions), the 27-cell
//     search of every target (row-parallel), and for systems with contacts
//     the lexicographically ordered pair list, packed system after system
//     with both droplets' (x, y, z, r).  The first check of a global step
//     clears the per-system "stepped" flags and covers every system with
//     steps left; a re-check covers only the systems just merged.
//   * "advance" graph: go = steps left, no contact in the last check, not yet
//     stepped (fixed at the start of the launch); force items, force
//     reduction, Oseen items, Oseen reduction + Euler step + wrap +
//     re-quantisation for go systems.
//   * dem_batch_merge applies host-computed merges on the device: survivors
//     compacted in order, merged droplets appended.
//   A global step is therefore: check; advance the clean systems while the
//   host plans merges of the others; merge, check again ... ; advance the
//   rest — each system goes through exactly the reference loop's states and
//   advances once.
//   * The per-target sums call the single-system functions (direct_core.h)
//     with the same source split, so every system's result equals the
//     single-system path's bit for bit.
#include <math.h>
#include <stddef.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <numeric>
#include <vector>

#include "dem_common.h"
#include "direct_core.h"
#include "direct_sym.h"
#include "euler_core.h"
#include "hash_core.h"
#include "profile.h"

namespace {

using dem::DevScales;
using dem::Rec;
using dem::RecC;

constexpr int kRow = 256;       // droplets per row
constexpr int kSearchSplit = 4;  // lanes per droplet in the contact count pass
constexpr int kCandUnroll = 2;   // candidate loads in flight per lane in the count pass
constexpr int kScanWG = 1024;   // threads of a per-system scan workgroup
constexpr int kZeroBlocks = 1024;
constexpr int kIncCap = 1024;   // pairs an incremental re-check holds per system
constexpr int kIncFresh = 64;   // merged droplets an incremental re-check takes per system
constexpr int kIncOld = 256;    // survivors merged since the last bucket table, per system

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

// Pair-once mode: chunks S of each block's partner list for about jlen / 256
// partner blocks per chunk (the split to pass to dem_direct_set_tuning in
// pair-once mode 2 for bit-identical single-system sums); entries per chunk
// ceil(E / S), E = 1 + B / 2.  Never decreases with n.
inline int pair_once_split(int n, int jlen) {
  const int B = (n + dem::kSymB - 1) / dem::kSymB;
  const int emax = dem::sym_entries_max(B > 0 ? B : 1);
  const int ect = std::max(1, jlen / dem::kSymB);
  int S = (emax + ect - 1) / ect;
  S = std::min(S, std::min(emax, 64));
  return S < 1 ? 1 : S;
}

inline int next_pow2(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

inline int pair_capacity(int cap) { return cap > 64 ? cap : 64; }

// Host-owned description of one system (uploaded whole).
struct BEns {
  long long part_off;   // doubles into the partial-sum array
  long long cnt_off;    // ints into the bucket counts (T per system)
  long long start_off;  // ints into the bucket starts (T + 1 per system)
  long long partj_off;  // entries into the reversed-pair partials (pair-once mode)
  int off;              // first droplet in the caller's arrays
  int roff;             // first record (regions padded to 256 records)
  int n, cap, S;        // S: source splits, or chunks of a block's partner list
  int jchunk;           // sources per split, or partner entries per chunk
  int steps, T;
  int fresh;  // droplets appended by the caller's last merge (a complete greedy
              // plan of the last check's pairs), 0 when unknown
  int pad_;
  double L, dt, sf, m_const, eta_const;
};

// Device-owned per-system values.
struct BDev {
  DevScales sc;
  double cell;  // 2 max r
  int g;        // cells per axis
  int skip;     // cell <= 0: no contacts
  double cell_checked;  // cell of the last check
};

struct BHead {
  int n_items;
  int n_rows;
  int n_ens;
  int contacts;  // 0: every active system advances (no contact pass)
  int pair_cap;  // pairs the packed pair buffer holds
  int pad[3];
};

struct BLayout {
  BEns* ens;
  BDev* dev;
  BHead* head;
  int4* items;  // (e, i0, s, 0)
  int2* rows;   // (e, i0)
  Rec* rec;
  RecC* recc;
  double* part;
  dem::PartJ* part_j;  // pair-once reversed-pair partials [sum B (B/2) 3 256]
  int* cnt;      // [sum T]
  int* start;    // [sum T + 1]
  int* bucket;   // [sum cap]
  int* rank;     // [sum cap]
  int* sorted;   // [sum cap] local droplet indices in bucket order
  double4* spos; // [sum cap] (x, y, z, r) in bucket order
  int* pc;       // [sum cap] pairs found for each target
  int* poff;     // [sum cap] exclusive scan of pc within a system
  int2* pairs;   // [pair_cap] packed system after system
  double4* pvals;// [2 pair_cap] (x, y, z, r) of i and of j
  int* pair_base;// [n_ens + 1] exclusive scan of pair_count
  int* inc;      // [n_ens] 1: this re-check searches only around the merged droplets
  int* oob;      // [n_ens] 1: a droplet lies outside [0, L]^3
  int* merged;   // [n_ens] 1: merged on the device since the last check
  int* t0n;      // [n_ens] droplets in the last bucket table (built by a full check)
  int* t0alive;  // [n_ens] of which still present (their indices are 0 .. t0alive - 1)
  double4* cscr; // [sum cap] survivors of a merge (compaction scratch)
  int2* inc_pairs;  // [n_ens kIncCap] sorted pairs of the incremental re-checks
  int* stage_i;  // merge lists uploaded by dem_batch_merge
  double* stage_d;
  BEns* ens_stage;  // [n_ens] updated descriptors of the systems just merged
  int* active;
  int* stepped;
  int* go;
  int* pair_count;
  int* steps_done;
  double* time;
  size_t bytes;
};

struct Plan {
  long long items_cap = 0, rows_cap = 0, part_total = 0, cnt_total = 0, start_total = 0;
  long long pair_total = 0, n_total = 0, cap_total = 0, rec_total = 0, partj_total = 0;
  std::vector<long long> part_off, cnt_off, start_off, partj_off;
  std::vector<int> rec_off;
};

Plan make_plan(const dem_batch_ensemble* ens, int n_ens, int block, int jlen, int pair_once) {
  Plan p;
  p.part_off.resize(n_ens);
  p.cnt_off.resize(n_ens);
  p.start_off.resize(n_ens);
  p.partj_off.resize(n_ens);
  p.rec_off.resize(n_ens);
  for (int e = 0; e < n_ens; ++e) {
    const int cap = ens[e].capacity;
    const int Bcap = std::max(1, (cap + dem::kSymB - 1) / dem::kSymB);
    // a system only shrinks: bound its items and partials over every n <= cap
    p.part_off[e] = p.part_total;
    p.partj_off[e] = p.partj_total;
    if (pair_once) {
      const int S = pair_once_split(cap, jlen);
      p.items_cap += (long long)Bcap * S;
      p.part_total += (long long)S * 3 * cap;
      p.partj_total += (long long)std::max(1, Bcap * (Bcap / 2)) * 3 * dem::kSymB;
    } else {
      const int S = split_bound(cap, jlen);
      p.items_cap += (long long)((cap + block - 1) / block) * S;
      p.part_total += (long long)S * 6 * cap;
    }
    p.rows_cap += (cap + kRow - 1) / kRow;
    p.rec_off[e] = (int)p.rec_total;
    p.rec_total += (long long)Bcap * dem::kSymB;
    const int T = next_pow2(2 * (cap > 0 ? cap : 1));
    p.cnt_off[e] = p.cnt_total;
    p.cnt_total += T;
    p.start_off[e] = p.start_total;
    p.start_total += T + 1;
    p.pair_total += pair_capacity(cap);
    p.cap_total += cap;
    long long end = (long long)ens[e].offset + cap;
    if (end > p.n_total) p.n_total = end;
  }
  if (p.pair_total > 0x3fffffffLL) p.pair_total = 0x3fffffffLL;  // int offsets
  return p;
}

BLayout carve(void* ws, int n_ens, const Plan& p) {
  dem::Carver c(ws);
  BLayout l;
  const size_t ne = (size_t)(n_ens > 0 ? n_ens : 1);
  auto at_least1 = [](long long v) { return (size_t)std::max<long long>(v, 1); };
  l.ens = c.take<BEns>(ne);
  l.dev = c.take<BDev>(ne);
  l.head = c.take<BHead>(1);
  l.items = c.take<int4>(at_least1(p.items_cap));
  l.rows = c.take<int2>(at_least1(p.rows_cap));
  l.rec = c.take<Rec>(at_least1(p.rec_total));
  l.recc = c.take<RecC>(at_least1(p.rec_total));
  l.part = c.take<double>(at_least1(p.part_total));
  l.part_j = c.take<dem::PartJ>(at_least1(p.partj_total));
  l.cnt = c.take<int>(at_least1(p.cnt_total));
  l.start = c.take<int>(at_least1(p.start_total));
  l.bucket = c.take<int>(at_least1(p.n_total));
  l.rank = c.take<int>(at_least1(p.n_total));
  l.sorted = c.take<int>(at_least1(p.n_total));
  l.spos = c.take<double4>(at_least1(p.n_total));
  l.pc = c.take<int>(at_least1(p.n_total));
  l.poff = c.take<int>(at_least1(p.n_total));
  l.pairs = c.take<int2>(at_least1(p.pair_total));
  l.pvals = c.take<double4>(2 * at_least1(p.pair_total));
  l.pair_base = c.take<int>(ne + 1);
  l.inc = c.take<int>(ne);
  l.oob = c.take<int>(ne);
  l.merged = c.take<int>(ne);
  l.t0n = c.take<int>(ne);
  l.t0alive = c.take<int>(ne);
  l.cscr = c.take<double4>(at_least1(p.n_total));
  l.inc_pairs = c.take<int2>(ne * kIncCap);
  l.stage_i = c.take<int>(at_least1(p.cap_total) + 3 * ne + 3);
  l.stage_d = c.take<double>(4 * at_least1(p.cap_total));
  l.ens_stage = c.take<BEns>(ne);
  l.active = c.take<int>(ne);
  l.stepped = c.take<int>(ne);
  l.go = c.take<int>(ne);
  l.pair_count = c.take<int>(ne);
  l.steps_done = c.take<int>(ne);
  l.time = c.take<double>(ne);
  l.bytes = dem::align_up(c.used, 256);
  return l;
}

// ---------------------------------------------------------------- set-up kernels

__global__ void k_b_init(const BEns* __restrict__ ens, BDev* __restrict__ dev, int n_ens,
                         int* __restrict__ active, int* __restrict__ stepped,
                         int* __restrict__ go, int* __restrict__ pair_count,
                         int* __restrict__ steps_done, double* __restrict__ time,
                         int* __restrict__ inc, int* __restrict__ oob,
                         int* __restrict__ merged, int* __restrict__ t0n,
                         int* __restrict__ t0alive) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_ens) return;
  dev[e].sc = dem::periodic_scales(ens[e].L, ens[e].m_const);
  dev[e].cell_checked = -1.0;
  inc[e] = 0;
  oob[e] = 0;
  merged[e] = 0;
  t0n[e] = 0;
  t0alive[e] = 0;
  active[e] = 0;
  stepped[e] = 0;
  go[e] = 0;
  pair_count[e] = 0;
  steps_done[e] = 0;
  time[e] = 0.0;
}

// Hash cell of every system: cell = 2 max r (spatial_hash...py:257-258; cell
// <= 0 means no contacts, :259-260), g = max(1, ceil(L / cell)) (:262).
__device__ void cells_of(const BEns* __restrict__ ens, const double* __restrict__ radii,
                         BDev* __restrict__ dev, int e) {
  __shared__ double s_max[4];
  const BEns& E = ens[e];
  double m = -INFINITY;
  for (int i = threadIdx.x; i < E.n; i += 256) m = fmax(m, radii[(size_t)E.off + i]);
  for (int d = 32; d > 0; d >>= 1) m = fmax(m, __shfl_down(m, d, 64));
  if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x != 0) return;
  m = fmax(fmax(s_max[0], s_max[1]), fmax(s_max[2], s_max[3]));
  const double cell = 2.0 * m;
  const int skip = !(cell > 0.0);
  double q = skip ? 1.0 : ceil(E.L / cell);
  if (!(q >= 1.0)) q = 1.0;
  if (q > 1073741824.0) q = 1073741824.0;
  dev[e].cell = cell;
  dev[e].g = (int)q;
  dev[e].skip = skip;
}

__global__ __launch_bounds__(256) void k_b_cells(const BEns* __restrict__ ens,
                                                 const BHead* __restrict__ head,
                                                 const double* __restrict__ radii,
                                                 BDev* __restrict__ dev) {
  if ((int)blockIdx.x >= head->n_ens) return;
  cells_of(ens, radii, dev, blockIdx.x);
}

// The same for the `count` listed systems only (grid = count).
__global__ __launch_bounds__(256) void k_b_cells_list(const BEns* __restrict__ ens, int count,
                                                      const int* __restrict__ systems,
                                                      const double* __restrict__ radii,
                                                      BDev* __restrict__ dev) {
  if ((int)blockIdx.x >= count) return;
  cells_of(ens, radii, dev, systems[blockIdx.x]);
}

// Quantise droplet i of system e into its record (zero padding at i >= n).
__device__ __forceinline__ void pack_droplet(const BEns* __restrict__ ens,
                                             const BDev* __restrict__ dev,
                                             const double* __restrict__ pos,
                                             const double* __restrict__ radii,
                                             Rec* __restrict__ rec, RecC* __restrict__ recc, int e,
                                             int i) {
  const BEns& E = ens[e];
  const size_t q = (size_t)E.roff + i;
  if (i >= E.n) {  // the last block's padding: no radius, no force
    rec[q] = Rec{};
    recc[q] = RecC{};
    return;
  }
  const size_t g = (size_t)E.off + i;
  Rec o;
  RecC c;
  dem::pack_one(pos + 3 * g, radii[g], nullptr, dev[e].sc, o, c);
  rec[q] = o;
  recc[q] = c;
}

// Quantise the droplets of every row.
__global__ __launch_bounds__(kRow) void k_b_pack(const BEns* __restrict__ ens,
                                                 const BDev* __restrict__ dev,
                                                 const int2* __restrict__ rows,
                                                 const BHead* __restrict__ head,
                                                 const double* __restrict__ pos,
                                                 const double* __restrict__ radii,
                                                 Rec* __restrict__ rec, RecC* __restrict__ recc) {
  if ((int)blockIdx.x >= head->n_rows) return;
  const int2 row = rows[blockIdx.x];
  pack_droplet(ens, dev, pos, radii, rec, recc, row.x, row.y + threadIdx.x);
}

// The same for the rows of the `count` listed systems (grid = count x
// rows_max, rows_max = rows of the largest capacity): droplets below n and
// the padding of the last 256-record block.
__global__ __launch_bounds__(kRow) void k_b_pack_list(const BEns* __restrict__ ens,
                                                      const BDev* __restrict__ dev, int count,
                                                      int rows_max,
                                                      const int* __restrict__ systems,
                                                      const double* __restrict__ pos,
                                                      const double* __restrict__ radii,
                                                      Rec* __restrict__ rec,
                                                      RecC* __restrict__ recc) {
  const int k = (int)blockIdx.x / rows_max, r = (int)blockIdx.x % rows_max;
  if (k >= count) return;
  const int e = systems[k];
  const int i = r * kRow + threadIdx.x;
  const int n = ens[e].n;
  if (i >= ((n + dem::kSymB - 1) / dem::kSymB) * dem::kSymB) return;
  pack_droplet(ens, dev, pos, radii, rec, recc, e, i);
}

// Descriptors of the systems just merged (dem_batch_merge_async).
__global__ void k_b_ens_scatter(int count, const int* __restrict__ systems,
                                const BEns* __restrict__ src, BEns* __restrict__ dst) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < count) dst[systems[k]] = src[k];
}

// Merges computed by the caller.  Stage layout: ints [0,c) systems,
// [c, 2c+1) removed offsets, [2c+1, 3c+2) merged offsets, then the removed
// local indices (ascending per system); doubles: (x, y, z, r) of every merged
// droplet.  Survivors keep their order and the merged droplets follow them
// (droplet_solution.py:263-268).  Two passes over a grid of c x rows_max
// workgroups of 256 droplets (ens still holds the sizes before the merge):
// the survivors are gathered into `scratch` (a survivor's new index is i minus
// the removed indices below it), then copied back and the merged droplets
// appended.
// Removed indices below i in the ascending list R[0, m).
__device__ __forceinline__ int removed_below(const int* __restrict__ R, int m, int i) {
  int lo = 0, hi = m;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (R[mid] < i) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// The gather also carries the system's last bucket table through the merge:
// each table slot's droplet index becomes its index after the merge, or -1
// when it leaves, and t0alive counts the table's droplets still present
// (survivors keep their order, so they are the first t0alive).
__global__ __launch_bounds__(256) void k_b_compact_gather(
    const BEns* __restrict__ ens, int c, int rows_max, const int* __restrict__ stage_i,
    const double* __restrict__ pos, const double* __restrict__ radii,
    double4* __restrict__ scratch, int* __restrict__ sorted, const int* __restrict__ t0n,
    int* __restrict__ t0alive) {
  const int k = (int)blockIdx.x / rows_max, r = (int)blockIdx.x % rows_max;
  if (k >= c) return;
  const int e = stage_i[k];
  const int* roff = stage_i + c;
  const int* R = stage_i + 3 * c + 2 + roff[k];
  const int m = roff[k + 1] - roff[k];
  const BEns& E = ens[e];
  const int i = r * 256 + threadIdx.x;
  if (r == 0 && threadIdx.x == 0) t0alive[e] -= removed_below(R, m, t0alive[e]);
  if (i < t0n[e]) {
    int* so = sorted + E.off;
    const int a = so[i];
    if (a >= 0) {
      const int lo = removed_below(R, m, a);
      so[i] = (lo < m && R[lo] == a) ? -1 : a - lo;
    }
  }
  if (i >= E.n) return;
  const int lo = removed_below(R, m, i);
  if (lo < m && R[lo] == i) return;
  const size_t off = (size_t)E.off, g = off + i;
  scratch[off + (i - lo)] = make_double4(pos[3 * g], pos[3 * g + 1], pos[3 * g + 2], radii[g]);
}

__global__ __launch_bounds__(256) void k_b_compact_scatter(
    const BEns* __restrict__ ens, int c, int rows_max, const int* __restrict__ stage_i,
    const double* __restrict__ stage_d, const double4* __restrict__ scratch,
    double* __restrict__ pos, double* __restrict__ radii, int* __restrict__ merged) {
  const int k = (int)blockIdx.x / rows_max, r = (int)blockIdx.x % rows_max;
  if (k >= c) return;
  const int e = stage_i[k];
  if (r == 0 && threadIdx.x == 0) merged[e] = 1;
  const int* roff = stage_i + c;
  const int* moff = stage_i + 2 * c + 1;
  const int m = roff[k + 1] - roff[k];
  const int m0 = moff[k], mc = moff[k + 1] - moff[k];
  const BEns& E = ens[e];
  const int ns = E.n - m;
  const int q = r * 256 + threadIdx.x;
  if (q >= ns + mc) return;
  const size_t off = (size_t)E.off, g = off + q;
  if (q < ns) {
    const double4 v = scratch[off + q];
    pos[3 * g] = v.x;
    pos[3 * g + 1] = v.y;
    pos[3 * g + 2] = v.z;
    radii[g] = v.w;
  } else {
    const double* v = stage_d + 4 * (size_t)(m0 + q - ns);
    pos[3 * g] = v[0];
    pos[3 * g + 1] = v[1];
    pos[3 * g + 2] = v[2];
    radii[g] = v[3];
  }
}

// ---------------------------------------------------------------- check graph

__device__ __forceinline__ bool eligible(const BEns& E, const int* steps_done,
                                         const int* stepped, int e) {
  return steps_done[e] < E.steps && E.n > 0 && !stepped[e];
}

// Which systems this check hashes ("active").  new_step (the first check of
// a global step) clears "stepped" and checks every system with steps left; a
// re-check only covers the systems the previous check found in contact (the
// caller has merged them since); the others keep pair_count = 0.
//
// A re-check may be incremental (inc): after a greedy merge of the complete
// pair list every pair of that list lost a droplet (a pair whose droplets
// were both free when it was visited merged them), and survivors keep their
// positions, radii and relative order.  If the hash cell is unchanged too
// (so are g and every bucket), the survivors' mutual pairs are those of the
// last check without a removed droplet: none.  The re-check's pairs then all
// involve a merged droplet, which the caller appended (the last `fresh`
// indices), and k_b_inc_search finds exactly those.  (`merged`, set by the
// compaction and cleared here, says that the merge came after the last check.)  Otherwise, and when the
// last cell is too narrow for k_b_inc_search's neighbourhood bound, the
// re-check is a full one.
__global__ void k_b_begin(const BEns* __restrict__ ens, BDev* __restrict__ dev,
                          const BHead* __restrict__ head, int new_step,
                          const int* __restrict__ steps_done, int* __restrict__ stepped,
                          int* __restrict__ active, int* __restrict__ pair_count,
                          int* __restrict__ inc, int* __restrict__ oob,
                          int* __restrict__ merged, int* __restrict__ t0n,
                          int* __restrict__ t0alive) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= head->n_ens) return;
  const bool was_merged = merged[e] != 0;
  merged[e] = 0;
  if (new_step) stepped[e] = 0;
  const BEns& E = ens[e];
  const bool want = new_step ? true : pair_count[e] > 0;
  const bool act = want && eligible(E, steps_done, stepped, e);
  active[e] = act ? 1 : 0;
  pair_count[e] = 0;
  bool incremental = false;
  if (act) {
    BDev& D = dev[e];
    const double w = E.L / D.cell - (double)(D.g - 1);  // width of the last cell, in cells
    incremental = !new_step && was_merged && E.fresh > 0 && E.fresh <= kIncFresh &&
                  E.fresh <= E.n && !oob[e] && t0n[e] > 0 &&
                  !D.skip && D.cell == D.cell_checked && (D.g <= 5 || w >= 1e-3);
    D.cell_checked = D.cell;
    if (!incremental) {  // this check builds a new bucket table
      oob[e] = 0;
      t0n[e] = E.n;
      t0alive[e] = E.n;
    }
  }
  inc[e] = incremental ? 1 : 0;
}

__global__ void k_b_zero(int* __restrict__ p, long long count) {
  for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < count;
       k += (long long)gridDim.x * blockDim.x)
    p[k] = 0;
}

__device__ __forceinline__ bool hashed(const BHead* head, const int* active, const BDev* dev,
                                       int e) {
  return head->contacts && active[e] && !dev[e].skip;
}

// Systems the full search covers: hashed and not re-checked incrementally.
__device__ __forceinline__ bool searched(const BHead* head, const int* active, const BDev* dev,
                                         const int* inc, int e) {
  return hashed(head, active, dev, e) && !inc[e];
}

__device__ __forceinline__ int bucket_of(const double* p, const BDev& D, int mask) {
  return dem::hash_cell(dem::cell_of(p[0], D.cell, D.g, 1), dem::cell_of(p[1], D.cell, D.g, 1),
                        dem::cell_of(p[2], D.cell, D.g, 1), mask);
}

// Bucket of every droplet and its rank among the bucket's droplets.
__global__ __launch_bounds__(kRow) void k_b_assign(const BEns* __restrict__ ens,
                                                   const BDev* __restrict__ dev,
                                                   const int2* __restrict__ rows,
                                                   const BHead* __restrict__ head,
                                                   const int* __restrict__ active,
                                                   const int* __restrict__ inc,
                                                   const double* __restrict__ pos,
                                                   int* __restrict__ cnt, int* __restrict__ bucket,
                                                   int* __restrict__ rank, int* __restrict__ oob) {
  if ((int)blockIdx.x >= head->n_rows) return;
  const int2 row = rows[blockIdx.x];
  if (!searched(head, active, dev, inc, row.x)) return;
  const BEns& E = ens[row.x];
  const int i = row.y + threadIdx.x;
  if (i >= E.n) return;
  const size_t g = (size_t)E.off + i;
  const double* p = pos + 3 * g;
  // (NaN fails the test too)
  if (!(p[0] >= 0.0 && p[0] <= E.L && p[1] >= 0.0 && p[1] <= E.L && p[2] >= 0.0 && p[2] <= E.L))
    oob[row.x] = 1;
  const int h = bucket_of(p, dev[row.x], E.T - 1);
  bucket[g] = h;
  rank[g] = atomicAdd(&cnt[E.cnt_off + h], 1);
}

// Exclusive scan of in[0, len) into out[0, len) (and out[len] = total when
// `total`) by one kScanWG-thread workgroup, in tiles of 4 kScanWG values:
// thread t takes values 4t .. 4t + 3 of a tile (consecutive threads read
// consecutive 16-byte runs), a wave scan and a scan of the wave totals give
// its offset, and the tile total carries into the next tile.
// With `clear`, every input value is set to 0 once read (in and clear are
// then the same array, hence no __restrict__ on either).
__device__ void block_scan(const int* in, int* __restrict__ out, int len, bool total,
                           int* clear = nullptr) {
  constexpr int kW = kScanWG / 64;
  __shared__ int s_w[kW];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int carry = 0;
  for (int base = 0; base < len; base += 4 * kScanWG) {
    const int b = base + 4 * (int)threadIdx.x;
    int x[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) x[k] = b + k < len ? in[b + k] : 0;
    if (clear) {
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (b + k < len) clear[b + k] = 0;
    }
    const int local = x[0] + x[1] + x[2] + x[3];
    int v = local;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int u = __shfl_up(v, d, 64);
      if (lane >= d) v += u;
    }
    if (lane == 63) s_w[wave] = v;
    __syncthreads();
    if (wave == 0) {
      int w = lane < kW ? s_w[lane] : 0;
#pragma unroll
      for (int d = 1; d < kW; d <<= 1) {
        const int u = __shfl_up(w, d, 64);
        if (lane >= d) w += u;
      }
      if (lane < kW) s_w[lane] = w;
    }
    __syncthreads();
    int run = carry + v - local + (wave > 0 ? s_w[wave - 1] : 0);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (b + k < len) out[b + k] = run;
      run += x[k];
    }
    carry += s_w[kW - 1];
    __syncthreads();  // s_w is rewritten by the next tile
  }
  if (total && threadIdx.x == 0) out[len] = carry;
}

__global__ __launch_bounds__(kScanWG) void k_b_bucket_scan(const BEns* __restrict__ ens,
                                                           const BDev* __restrict__ dev,
                                                           const BHead* __restrict__ head,
                                                           const int* __restrict__ active,
                                                           const int* __restrict__ inc,
                                                           int* cnt, int* __restrict__ start) {
  const int e = blockIdx.x;
  if (e >= head->n_ens || !searched(head, active, dev, inc, e)) return;
  const BEns& E = ens[e];
  // T + 1 slots per system; the counts are left at 0 for the next check's atomics
  block_scan(cnt + E.cnt_off, start + E.start_off, E.T, true, cnt + E.cnt_off);
}

__global__ __launch_bounds__(kRow) void k_b_scatter(const BEns* __restrict__ ens,
                                                    const BDev* __restrict__ dev,
                                                    const int2* __restrict__ rows,
                                                    const BHead* __restrict__ head,
                                                    const int* __restrict__ active,
                                                    const int* __restrict__ inc,
                                                    const int* __restrict__ start,
                                                    const int* __restrict__ bucket,
                                                    const int* __restrict__ rank,
                                                    const double* __restrict__ pos,
                                                    const double* __restrict__ radii,
                                                    int* __restrict__ sorted,
                                                    double4* __restrict__ spos) {
  if ((int)blockIdx.x >= head->n_rows) return;
  const int2 row = rows[blockIdx.x];
  if (!searched(head, active, dev, inc, row.x)) return;
  const BEns& E = ens[row.x];
  const int i = row.y + threadIdx.x;
  if (i >= E.n) return;
  const size_t g = (size_t)E.off + i;
  const size_t slot = (size_t)E.off + start[E.start_off + bucket[g]] + rank[g];
  sorted[slot] = i;
  spos[slot] = make_double4(pos[3 * g], pos[3 * g + 1], pos[3 * g + 2], radii[g]);
}

// The 27 neighbour cells of cell (ix, iy, iz) (periodic), neighbour k =
// 9 a + 3 b + c at offsets (a - 1, b - 1, c - 1).  The reference visits the
// bucket of every neighbour, so a bucket that two neighbours share (a hash
// alias, or the same cell when g <= 2) yields duplicate pairs, which its
// np.unique removes: its pair set is {(i, j): i < j, contact, bucket(j) in
// B(i)}.  Here a candidate is counted in neighbour k only when no neighbour
// q < k has the same bucket; that test runs only for candidates in contact,
// so the neighbour buckets need no deduplication up front.
struct Nbhd {
  int cx[3], cy[3], cz[3];
  int mask;
  __device__ __forceinline__ Nbhd(int ix, int iy, int iz, int g, int m) : mask(m) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      cx[d] = dem::nb_coord(ix, d - 1, g, 1);
      cy[d] = dem::nb_coord(iy, d - 1, g, 1);
      cz[d] = dem::nb_coord(iz, d - 1, g, 1);
    }
  }
  __device__ __forceinline__ int bucket(int k) const {
    return dem::hash_cell(cx[k / 9], cy[(k / 3) % 3], cz[k % 3], mask);
  }
  // whether neighbour k is the first with bucket h
  __device__ __forceinline__ bool first(int k, int h) const {
    for (int q = 0; q < k; ++q)
      if (bucket(q) == h) return false;
    return true;
  }
};

// Pairs (i, j), i < j, of target i over its neighbour buckets
// (spatial_hash...py:134-229); EMIT writes them to out[0, room) sorted by j.
// The 27 bucket ranges are loaded up front and candidates are read from the
// bucket-ordered copy (sorted index and (x, y, z, r) of the same slot), so a
// candidate costs one memory round trip.
template <bool EMIT>
__device__ int search_target(const BEns& E, const BDev& D, const double* __restrict__ p,
                             const double* __restrict__ r, const int* __restrict__ start,
                             const int* __restrict__ sorted, const double4* __restrict__ spos,
                             int i, int2* __restrict__ out, double4* __restrict__ vals,
                             int room) {
  const int g = D.g;
  const double xi = p[3 * (size_t)i], yi = p[3 * (size_t)i + 1], zi = p[3 * (size_t)i + 2];
  const double ri = r[i];
  const Nbhd nb(dem::cell_of(xi, D.cell, g, 1), dem::cell_of(yi, D.cell, g, 1),
                dem::cell_of(zi, D.cell, g, 1), g, E.T - 1);
  int hk[27], s0[27], s1[27];
#pragma unroll
  for (int k = 0; k < 27; ++k) {
    hk[k] = nb.bucket(k);
    s0[k] = start[hk[k]];
    s1[k] = start[hk[k] + 1];
  }
  int cnt = 0;
#pragma unroll
  for (int k = 0; k < 27; ++k) {
    for (int slot = s0[k]; slot < s1[k]; ++slot) {
      const int j = sorted[slot];
      const double4 q = spos[slot];
      if (i < j && dem::contact_pair(xi, yi, zi, ri, q.x, q.y, q.z, q.w, E.L, 1) &&
          nb.first(k, hk[k])) {
        if (EMIT && cnt < room) {
          out[cnt] = make_int2(i, j);
          vals[2 * cnt] = make_double4(xi, yi, zi, ri);
          vals[2 * cnt + 1] = q;
        }
        ++cnt;
      }
    }
  }
  if (EMIT) {  // in-segment insertion sort by j (a handful of pairs)
    const int hi = min(cnt, room);
    for (int a = 1; a < hi; ++a) {
      const int2 v = out[a];
      const double4 vj = vals[2 * a + 1];
      int b = a - 1;
      while (b >= 0 && out[b].y > v.y) {
        out[b + 1] = out[b];
        vals[2 * (b + 1) + 1] = vals[2 * b + 1];
        --b;
      }
      out[b + 1] = v;
      vals[2 * (b + 1) + 1] = vj;
    }
  }
  return cnt;
}

// Count pass of the droplet in bucket slot `slot` (neighbouring slots share
// neighbour buckets): lane `sub` of SPLIT takes the neighbours k = sub,
// sub + SPLIT, ...; the candidate's droplet index is read only for a contact.
// Returns this lane's share of the droplet's pairs (i, j), i < j.
template <int SPLIT>
__device__ int count_slot(const BEns& E, const BDev& D, const int* __restrict__ start,
                          const int* __restrict__ sorted, const double4* __restrict__ spos,
                          int slot, int sub, int* i_out) {
  const int g = D.g;
  const double4 me = spos[slot];
  const int i = sorted[slot];
  *i_out = i;
  const Nbhd nb(dem::cell_of(me.x, D.cell, g, 1), dem::cell_of(me.y, D.cell, g, 1),
                dem::cell_of(me.z, D.cell, g, 1), g, E.T - 1);
  constexpr int KS = (27 + SPLIT - 1) / SPLIT;
  int hk[KS], s0[KS], s1[KS];
#pragma unroll
  for (int u = 0; u < KS; ++u) {
    const int k = sub + SPLIT * u;
    const bool ok = k < 27;
    hk[u] = nb.bucket(ok ? k : 0);
    s0[u] = start[hk[u]];
    s1[u] = ok ? start[hk[u] + 1] : s0[u];
  }
  int cnt = 0;
#pragma unroll
  for (int u = 0; u < KS; ++u) {
    // kCandUnroll candidates per round, their loads in flight together
    for (int q = s0[u]; q < s1[u]; q += kCandUnroll) {
      double4 c[kCandUnroll];
#pragma unroll
      for (int t = 0; t < kCandUnroll; ++t)
        if (q + t < s1[u]) c[t] = spos[q + t];
#pragma unroll
      for (int t = 0; t < kCandUnroll; ++t)
        if (q + t < s1[u] &&
            dem::contact_pair(me.x, me.y, me.z, me.w, c[t].x, c[t].y, c[t].z, c[t].w, E.L, 1) &&
            i < sorted[q + t] && nb.first(sub + SPLIT * u, hk[u]))
          ++cnt;
    }
  }
  return cnt;
}

// Per-droplet pair counts pc (original order) and per-system totals.  A row
// of kRow droplets is searched by kSearchSplit workgroups, kSearchSplit lanes
// per droplet, in bucket-slot order.
__global__ __launch_bounds__(kRow) void k_b_search(const BEns* __restrict__ ens,
                                                   const BDev* __restrict__ dev,
                                                   const int2* __restrict__ rows,
                                                   const BHead* __restrict__ head,
                                                   const int* __restrict__ active,
                                                   const int* __restrict__ inc,
                                                   const int* __restrict__ start,
                                                   const int* __restrict__ sorted,
                                                   const double4* __restrict__ spos,
                                                   int* __restrict__ pc,
                                                   int* __restrict__ pair_count) {
  __shared__ int s_sum[kRow / 64];
  const int rb = (int)blockIdx.x / kSearchSplit, part = (int)blockIdx.x % kSearchSplit;
  if (rb >= head->n_rows) return;
  const int2 row = rows[rb];
  if (!searched(head, active, dev, inc, row.x)) return;  // uniform over the block
  const BEns& E = ens[row.x];
  const int sub = threadIdx.x % kSearchSplit;
  const int slot = row.y + part * (kRow / kSearchSplit) + (int)threadIdx.x / kSearchSplit;
  int c = 0;
  if (slot < E.n) {  // the kSearchSplit lanes of a droplet are all here or all not
    int i = 0;
    c = count_slot<kSearchSplit>(E, dev[row.x], start + E.start_off, sorted + E.off,
                                 spos + E.off, slot, sub, &i);
    int tot = c;
    for (int o = 1; o < kSearchSplit; o <<= 1) tot += __shfl_xor(tot, o);
    if (sub == 0) pc[(size_t)E.off + i] = tot;
  }
  for (int d = 32; d > 0; d >>= 1) c += __shfl_down(c, d, 64);
  if ((threadIdx.x & 63) == 0) s_sum[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int tot = s_sum[0] + s_sum[1] + s_sum[2] + s_sum[3];
    if (tot) atomicAdd(&pair_count[row.x], tot);
  }
}

// poff of a system: exclusive scan of its per-target pair counts (no total:
// poff[off + n] may be the next system's first slot).
__global__ __launch_bounds__(kScanWG) void k_b_pair_scan(const BEns* __restrict__ ens,
                                                         const BHead* __restrict__ head,
                                                         const int* __restrict__ pair_count,
                                                         const int* __restrict__ inc,
                                                         const int* __restrict__ pc,
                                                         int* __restrict__ poff) {
  const int e = blockIdx.x;
  if (e >= head->n_ens || pair_count[e] == 0 || inc[e]) return;
  const BEns& E = ens[e];
  block_scan(pc + E.off, poff + E.off, E.n, false);
}

// ---------------------------------------------------------------- incremental re-check

__device__ __forceinline__ int b_wave_inclusive(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(v, o, 64);
    if (lane >= o) v += u;
  }
  return v;
}

// Whether bucket h is one of the buckets of the 27 cells around cell
// (cx, cy, cz) (periodic): h in B(i) for a droplet i of that cell.
__device__ __forceinline__ bool in_neighbourhood(int cx, int cy, int cz, int g, int mask, int h) {
  bool hit = false;
  for (int k = 0; k < 27; ++k) {
    const int nx = dem::nb_coord(cx, k / 9 - 1, g, 1);
    const int ny = dem::nb_coord(cy, (k / 3) % 3 - 1, g, 1);
    const int nz = dem::nb_coord(cz, k % 3 - 1, g, 1);
    hit |= dem::hash_cell(nx, ny, nz, mask) == h;
  }
  return hit;
}

struct IncLds {
  unsigned long long key[kIncCap];  // (i << 32) | j of each pair found
  int nb[4][128];                   // per wave: buckets of the 125 cells around m
  int pre[4][128];                  // per wave: candidates in those buckets before each
  int first[4][128];                // per wave: first slot of each bucket
  int count;
};

// Pairs of an incremental re-check (one workgroup per system; k_b_begin has
// set inc).  The reference's pair set is {(i, j): i < j, contact(i, j),
// bucket(j) in B(i)}; with no pair among the n_s = n - fresh survivors left,
// it is
//   (a) (a, m), a < n_s <= m: survivor a, merged m, contact and bucket(m) in
//       B(a).  Contact needs |d| <= r_a + r_m <= cell per axis (up to
//       rounding), so along each axis a's cell lies within two cells of m's on
//       the periodic ring (floor(fl(x/cell)) of the two positions differ by at
//       most 2 directly, and across the seam by g - 2 .. g + 1 when the last
//       cell is at least 1e-3 cells wide, k_b_begin's condition; with g <= 5
//       the five offsets cover the ring).  Survivors from the system's last
//       bucket table (built by its last full check, carried through the merges
//       since by k_b_compact_gather: positions unchanged, indices remapped,
//       -1 once removed) are found in the buckets of those 125 cells, each
//       bucket once; the few survivors merged after that table was built
//       (indices t0alive .. n_s - 1) are tested one by one.
//   (b) (m1, m2), both merged: every one of the fresh^2 / 2 pairs is tested.
// contact_pair is symmetric in its two droplets (negated differences, the same
// squares and sum of radii), so a pair is tested as in the full search.  The
// pairs are sorted lexicographically into inc_pairs[e].  More than kIncCap of
// them, more than kIncOld later survivors, or a merged droplet outside
// [0, L]^3 (the ring bound assumes wrapped positions) clears inc: the full
// check (new table, full search) then runs for the system.
__global__ __launch_bounds__(256) void k_b_inc_search(
    const BEns* __restrict__ ens, const BDev* __restrict__ dev, const BHead* __restrict__ head,
    const int* __restrict__ active, int* __restrict__ inc, int* __restrict__ oob,
    int* __restrict__ t0n, int* __restrict__ t0alive, const double* __restrict__ pos,
    const double* __restrict__ radii, const int* __restrict__ start,
    const int* __restrict__ sorted, const double4* __restrict__ spos,
    int* __restrict__ pair_count, int2* __restrict__ inc_pairs) {
  __shared__ IncLds S;
  const int e = blockIdx.x;
  if (e >= head->n_ens || !inc[e]) return;  // uniform over the block
  const BEns& E = ens[e];
  const BDev D = dev[e];
  const int n = E.n, k = E.fresh, ns = n - k, g = D.g, mask = E.T - 1;
  const int old0 = t0alive[e];
  const double* P = pos + 3 * (size_t)E.off;
  const double* R = radii + E.off;
  bool bad = false;
  for (int m = ns + (int)threadIdx.x; m < n; m += 256) {
    const double* x = P + 3 * (size_t)m;
    bad |= !(x[0] >= 0.0 && x[0] <= E.L && x[1] >= 0.0 && x[1] <= E.L && x[2] >= 0.0 &&
             x[2] <= E.L);
  }
  bad = __syncthreads_or(bad) != 0;
  if (!hashed(head, active, dev, e) || bad || old0 > ns || ns - old0 > kIncOld) {
    if (threadIdx.x == 0) {  // a full check: new table for the current droplets
      inc[e] = 0;
      oob[e] = 0;
      t0n[e] = n;
      t0alive[e] = n;
    }
    return;
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int* st = start + E.start_off;
  const int* so = sorted + E.off;
  const double4* sp = spos + E.off;
  if (threadIdx.x == 0) S.count = 0;
  __syncthreads();
  // (a), table survivors: wave w takes merged droplets ns + w, ns + w + 4, ...
  // (a uniform loop)
  for (int m = ns + w; m < n; m += 4) {
    const double xm = P[3 * (size_t)m], ym = P[3 * (size_t)m + 1], zm = P[3 * (size_t)m + 2];
    const double rm = R[m];
    const int mx = dem::cell_of(xm, D.cell, g, 1);
    const int my = dem::cell_of(ym, D.cell, g, 1);
    const int mz = dem::cell_of(zm, D.cell, g, 1);
    const int bm = dem::hash_cell(mx, my, mz, mask);
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const int o = lane + 64 * p;
      int h = -1;
      if (o < 125)
        h = dem::hash_cell(dem::pymod(mx + o / 25 - 2, g), dem::pymod(my + (o / 5) % 5 - 2, g),
                           dem::pymod(mz + o % 5 - 2, g), mask);
      S.nb[w][o] = h;
    }
    dem::wave_sync();
    int len[2], fst[2];
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const int o = lane + 64 * p;
      const int h = S.nb[w][o];
      bool dup = h < 0;
      for (int q = 0; q < o; ++q) dup |= S.nb[w][q] == h;
      fst[p] = dup ? 0 : st[h];
      len[p] = dup ? 0 : st[h + 1] - fst[p];
    }
    const int inc0 = b_wave_inclusive(len[0]);
    const int inc1 = b_wave_inclusive(len[1]) + __shfl(inc0, 63, 64);
    const int M = __shfl(inc1, 63, 64);
    S.pre[w][lane] = inc0 - len[0];
    S.pre[w][lane + 64] = inc1 - len[1];
    S.first[w][lane] = fst[0];
    S.first[w][lane + 64] = fst[1];
    dem::wave_sync();
    for (int c0 = 0; c0 < M; c0 += 64) {
      const int c = c0 + lane;
      if (c >= M) continue;
      int lo = 0, hi = 127;  // the last bucket whose candidates start at or before c
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (S.pre[w][mid] <= c) lo = mid;
        else hi = mid - 1;
      }
      const int slot = S.first[w][lo] + (c - S.pre[w][lo]);
      const int a = so[slot];
      if (a < 0 || a >= ns) continue;
      const double4 q = sp[slot];
      if (!dem::contact_pair(q.x, q.y, q.z, q.w, xm, ym, zm, rm, E.L, 1)) continue;
      if (!in_neighbourhood(dem::cell_of(q.x, D.cell, g, 1), dem::cell_of(q.y, D.cell, g, 1),
                            dem::cell_of(q.z, D.cell, g, 1), g, mask, bm))
        continue;
      const int at = atomicAdd(&S.count, 1);
      if (at < kIncCap) S.key[at] = ((unsigned long long)a << 32) | (unsigned)m;
    }
    dem::wave_sync();  // nb, pre and first are rewritten for the next droplet
  }
  // (a), later survivors, and (b): target t < m, merged m
  const int nt = n - old0;  // targets old0 .. n - 1
  for (int idx = threadIdx.x; idx < nt * k; idx += 256) {
    const int t = old0 + idx / k, m = ns + idx % k;
    if (t >= m) continue;
    const double xt = P[3 * (size_t)t], yt = P[3 * (size_t)t + 1], zt = P[3 * (size_t)t + 2];
    const double xm = P[3 * (size_t)m], ym = P[3 * (size_t)m + 1], zm = P[3 * (size_t)m + 2];
    if (!dem::contact_pair(xt, yt, zt, R[t], xm, ym, zm, R[m], E.L, 1)) continue;
    const int hm = dem::hash_cell(dem::cell_of(xm, D.cell, g, 1), dem::cell_of(ym, D.cell, g, 1),
                                  dem::cell_of(zm, D.cell, g, 1), mask);
    if (!in_neighbourhood(dem::cell_of(xt, D.cell, g, 1), dem::cell_of(yt, D.cell, g, 1),
                          dem::cell_of(zt, D.cell, g, 1), g, mask, hm))
      continue;
    const int at = atomicAdd(&S.count, 1);
    if (at < kIncCap) S.key[at] = ((unsigned long long)t << 32) | (unsigned)m;
  }
  __syncthreads();
  const int cnt = S.count;
  if (cnt > kIncCap) {
    if (threadIdx.x == 0) {
      inc[e] = 0;
      oob[e] = 0;
      t0n[e] = n;
      t0alive[e] = n;
    }
    return;
  }
  // rank sort (the keys are distinct: every pair is found once)
  for (int t = threadIdx.x; t < cnt; t += 256) {
    const unsigned long long v = S.key[t];
    int r = 0;
    for (int u = 0; u < cnt; ++u) r += S.key[u] < v ? 1 : 0;
    inc_pairs[(size_t)e * kIncCap + r] = make_int2((int)(v >> 32), (int)(v & 0xffffffffull));
  }
  if (threadIdx.x == 0) pair_count[e] = cnt;
}

// The incremental re-checks' pairs at their systems' places in the packed
// buffer, with the droplets' values (as k_b_emit writes them).
__global__ __launch_bounds__(256) void k_b_inc_emit(
    const BEns* __restrict__ ens, const BHead* __restrict__ head, const int* __restrict__ inc,
    const int* __restrict__ pair_count, const int* __restrict__ pair_base,
    const int2* __restrict__ inc_pairs, const double* __restrict__ pos,
    const double* __restrict__ radii, int2* __restrict__ pairs, double4* __restrict__ pvals) {
  const int e = blockIdx.x;
  if (e >= head->n_ens || !inc[e]) return;
  const BEns& E = ens[e];
  const int cnt = pair_count[e];
  for (int t = threadIdx.x; t < cnt; t += 256) {
    const long long at = (long long)pair_base[e] + t;
    if (at >= head->pair_cap) break;
    const int2 p = inc_pairs[(size_t)e * kIncCap + t];
    pairs[at] = p;
    const size_t gi = (size_t)E.off + p.x, gj = (size_t)E.off + p.y;
    pvals[2 * at] = make_double4(pos[3 * gi], pos[3 * gi + 1], pos[3 * gi + 2], radii[gi]);
    pvals[2 * at + 1] = make_double4(pos[3 * gj], pos[3 * gj + 1], pos[3 * gj + 2], radii[gj]);
  }
}

// Where each system's pairs start in the packed pair buffer.
__global__ __launch_bounds__(kScanWG) void k_b_pair_base(const BHead* __restrict__ head,
                                                         const int* __restrict__ pair_count,
                                                         int* __restrict__ pair_base) {
  block_scan(pair_count, pair_base, head->n_ens, true);
}

__global__ __launch_bounds__(kRow) void k_b_emit(const BEns* __restrict__ ens,
                                                 const BDev* __restrict__ dev,
                                                 const int2* __restrict__ rows,
                                                 const BHead* __restrict__ head,
                                                 const int* __restrict__ pair_count,
                                                 const int* __restrict__ inc,
                                                 const double* __restrict__ pos,
                                                 const double* __restrict__ radii,
                                                 const int* __restrict__ start,
                                                 const int* __restrict__ sorted,
                                                 const double4* __restrict__ spos,
                                                 const int* __restrict__ pc,
                                                 const int* __restrict__ poff,
                                                 const int* __restrict__ pair_base,
                                                 int2* __restrict__ pairs,
                                                 double4* __restrict__ pvals) {
  if ((int)blockIdx.x >= head->n_rows) return;
  const int2 row = rows[blockIdx.x];
  if (pair_count[row.x] == 0 || inc[row.x]) return;
  const BEns& E = ens[row.x];
  const int i = row.y + threadIdx.x;
  if (i >= E.n) return;
  const size_t g = (size_t)E.off + i;
  if (pc[g] == 0) return;
  const long long at = (long long)pair_base[row.x] + poff[g];
  const long long room = min((long long)pc[g], (long long)head->pair_cap - at);
  if (room <= 0) return;
  search_target<true>(E, dev[row.x], pos + 3 * (size_t)E.off, radii + E.off, start + E.start_off,
                      sorted + E.off, spos + E.off, i, pairs + at, pvals + 2 * at, (int)room);
}

// ---------------------------------------------------------------- advance graph

// go = steps left, not advanced yet in this global step, and no contact in
// its last check; the rest of the advance launch reads only go, so marking
// "stepped" here is safe.
__global__ void k_b_go(const BEns* __restrict__ ens, const BHead* __restrict__ head,
                       const int* __restrict__ steps_done, const int* __restrict__ pair_count,
                       int* __restrict__ stepped, int* __restrict__ go) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= head->n_ens) return;
  const int g = (eligible(ens[e], steps_done, stepped, e) && pair_count[e] == 0) ? 1 : 0;
  go[e] = g;
  if (g) stepped[e] = 1;
}

template <int BW>
__global__ __launch_bounds__(BW) void k_b_force(const BEns* __restrict__ ens,
                                                const int4* __restrict__ items,
                                                const BHead* __restrict__ head,
                                                const int* __restrict__ go,
                                                const Rec* __restrict__ rec, int tile,
                                                double* __restrict__ part) {
  if ((int)blockIdx.x >= head->n_items) return;
  const int4 it = items[blockIdx.x];
  if (!go[it.x]) return;
  const BEns& E = ens[it.x];
  const int n = E.n;
  const Rec* r = rec + E.roff;
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

// SYM: pair-once partials (chunk sums + reversed-pair sums), else source splits.
template <bool SYM>
__global__ __launch_bounds__(kRow) void k_b_force_reduce(const BEns* __restrict__ ens,
                                                         const BDev* __restrict__ dev,
                                                         const int2* __restrict__ rows,
                                                         const BHead* __restrict__ head,
                                                         const int* __restrict__ go,
                                                         const double* __restrict__ part,
                                                         const dem::PartJ* __restrict__ part_j,
                                                         const double* __restrict__ radii,
                                                         double* __restrict__ F,
                                                         RecC* __restrict__ recc) {
  if ((int)blockIdx.x >= head->n_rows) return;
  const int2 row = rows[blockIdx.x];
  if (!go[row.x]) return;
  const BEns& E = ens[row.x];
  const int n = E.n;
  const int i = row.y + threadIdx.x;
  if (i >= n) return;
  const double* pp = part + E.part_off;
  double sum[3] = {0.0, 0.0, 0.0};
  if (SYM) {
    dem::sym_sum(pp, E.S, n, part_j + E.partj_off, (n + dem::kSymB - 1) / dem::kSymB, i, sum);
  } else {
    for (int s = 0; s < E.S; ++s)
      for (int a = 0; a < 3; ++a) sum[a] += pp[(size_t)(6 * s + a) * n + i];
  }
  const size_t g = (size_t)E.off + i;
  double Fi[3];
  float fr[3];
  dem::force_finish(sum, radii[g], dev[row.x].sc, E.m_const, Fi, fr);
  for (int a = 0; a < 3; ++a) F[3 * g + a] = Fi[a];
  const size_t q = (size_t)E.roff + i;
  recc[q].fx = fr[0];
  recc[q].fy = fr[1];
  recc[q].fz = fr[2];
}

template <int BW>
__global__ __launch_bounds__(BW) void k_b_conv(const BEns* __restrict__ ens,
                                               const int4* __restrict__ items,
                                               const BHead* __restrict__ head,
                                               const int* __restrict__ go,
                                               const RecC* __restrict__ recc, int tile,
                                               double* __restrict__ part) {
  if ((int)blockIdx.x >= head->n_items) return;
  const int4 it = items[blockIdx.x];
  if (!go[it.x]) return;
  const BEns& E = ens[it.x];
  const int n = E.n;
  const RecC* r = recc + E.roff;
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

template <bool SYM>
__global__ __launch_bounds__(kRow) void k_b_conv_euler(
    const BEns* __restrict__ ens, const BDev* __restrict__ dev, const int2* __restrict__ rows,
    const BHead* __restrict__ head, const int* __restrict__ go, const double* __restrict__ part,
    const dem::PartJ* __restrict__ part_j, const double* __restrict__ radii,
    const double* __restrict__ F, double* __restrict__ V,
    double* __restrict__ pos, Rec* __restrict__ rec, RecC* __restrict__ recc,
    int* __restrict__ steps_done, double* __restrict__ time) {
  if ((int)blockIdx.x >= head->n_rows) return;
  const int2 row = rows[blockIdx.x];
  if (!go[row.x]) return;
  const BEns& E = ens[row.x];
  const DevScales sc = dev[row.x].sc;
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
  if (SYM) {
    dem::sym_sum(pp, E.S, n, part_j + E.partj_off, (n + dem::kSymB - 1) / dem::kSymB, i, m);
  } else {
    for (int s = 0; s < E.S; ++s)
      for (int a = 0; a < 3; ++a) m[a] += pp[(size_t)(6 * s + a) * n + i];
  }
  const size_t g = (size_t)E.off + i;
  const dem::CorrScalars cs{};
  double Vi[3];
  dem::conv_finish(m, c, sc, E.eta_const, cs, F + 3 * g, Vi);
  const double ri = radii[g];
  double x[3];
  for (int a = 0; a < 3; ++a) {
    V[3 * g + a] = Vi[a];
    x[a] = dem::euler_coord(pos[3 * g + a], F[3 * g + a], Vi[a], true, ri, E.sf, E.dt, E.L, 1);
    pos[3 * g + a] = x[a];
  }
  Rec o;
  RecC cc;
  dem::pack_one(x, ri, nullptr, sc, o, cc);
  const size_t q = (size_t)E.roff + i;
  rec[q] = o;
  recc[q] = cc;
}

// Pair-once sums: item (system, target block, chunk) runs one workgroup of
// the single-system pair-once bodies on the system's padded records, so each
// system's sums equal dem_direct_* in pair-once mode 2 with the same chunks.
template <int TPL>
__global__ __launch_bounds__(dem::kSymB / TPL) void k_b_force_sym(
    const BEns* __restrict__ ens, const int4* __restrict__ items, const BHead* __restrict__ head,
    const int* __restrict__ go, const Rec* __restrict__ rec, int tile, double* __restrict__ part,
    dem::PartJ* __restrict__ part_j) {
  if ((int)blockIdx.x >= head->n_items) return;
  const int4 it = items[blockIdx.x];
  if (!go[it.x]) return;
  const BEns& E = ens[it.x];
  const int n = E.n, B = (n + dem::kSymB - 1) / dem::kSymB;
  const int e0 = it.z * E.jchunk, e1 = min(1 + dem::sym_partners(it.y, B), e0 + E.jchunk);
  dem::force_sym_block<TPL>(rec + E.roff, n, B, it.y, it.z, e0, e1, tile, part + E.part_off,
                            part_j + E.partj_off);
}

template <int TPL>
__global__ __launch_bounds__(dem::kSymB / TPL) void k_b_conv_sym(
    const BEns* __restrict__ ens, const int4* __restrict__ items, const BHead* __restrict__ head,
    const int* __restrict__ go, const RecC* __restrict__ recc, int tile, double* __restrict__ part,
    dem::PartJ* __restrict__ part_j) {
  if ((int)blockIdx.x >= head->n_items) return;
  const int4 it = items[blockIdx.x];
  if (!go[it.x]) return;
  const BEns& E = ens[it.x];
  const int n = E.n, B = (n + dem::kSymB - 1) / dem::kSymB;
  const int e0 = it.z * E.jchunk, e1 = min(1 + dem::sym_partners(it.y, B), e0 + E.jchunk);
  dem::conv_sym_block<TPL>(recc + E.roff, n, B, it.y, it.z, e0, e1, tile, part + E.part_off,
                           part_j + E.partj_off);
}

}  // namespace

// ---------------------------------------------------------------- host handle

struct dem_batch {
  int n_ens = 0;
  dem_batch_tuning tune{};
  std::vector<dem_batch_ensemble> desc;
  std::vector<int> order;  // systems by decreasing capacity (item order)
  Plan plan;
  BLayout l{};
  double *pos = nullptr, *radii = nullptr, *F = nullptr, *V = nullptr;
  std::vector<BEns> ens_h;
  std::vector<int4> items_h;
  std::vector<int2> rows_h;
  std::vector<int> stage_h;
  std::vector<BEns> ens_stage_h;
  int cap_max = 0;            // largest capacity (rows of k_b_pack_list)
  std::vector<int> fresh;     // droplets appended by each system's last merge (BEns::fresh)
  bool incremental = false;   // merges are complete greedy plans: re-checks may be incremental
  bool tables_dirty = false;  // merged systems' items / rows not rebuilt yet
  hipEvent_t go_done = nullptr;  // recorded after the "go" kernel of an advance
  hipStream_t cap_stream = nullptr;
  // 0: first check of a global step, 1: check again, 2: advance
  hipGraph_t graph[3] = {nullptr, nullptr, nullptr};
  hipGraphExec_t exec[3] = {nullptr, nullptr, nullptr};
};

namespace {

int check_tuning(const dem_batch_tuning* t) {
  if (!t) return dem::set_error(DEM_ERR_INVALID, "tuning is NULL");
  if (t->block != 64 && t->block != 128 && t->block != 256)
    return dem::set_error(DEM_ERR_INVALID, "tuning.block must be 64, 128 or 256 (got %d)", t->block);
  if (t->force_tile < 0 || t->conv_tile < 0)
    return dem::set_error(DEM_ERR_INVALID, "tiles must be >= 0");
  if (t->pair_once != 0 && t->pair_once != 1)
    return dem::set_error(DEM_ERR_INVALID, "tuning.pair_once must be 0 or 1");
  return DEM_OK;
}

int check_ensembles(const dem_batch_ensemble* ens, int n_ens) {
  if (n_ens <= 0 || !ens) return dem::set_error(DEM_ERR_INVALID, "need at least one ensemble");
  std::vector<int> idx(n_ens);
  for (int e = 0; e < n_ens; ++e) {
    const dem_batch_ensemble& d = ens[e];
    if (d.offset < 0 || d.capacity < 0 || d.n < 0 || d.n > d.capacity)
      return dem::set_error(DEM_ERR_INVALID, "ensemble %d: bad offset/n/capacity", e);
    if (d.capacity > (1 << 24))
      return dem::set_error(DEM_ERR_CAPACITY, "ensemble %d: capacity %d too large", e, d.capacity);
    if (!(d.L > 0.0)) return dem::set_error(DEM_ERR_INVALID, "ensemble %d: L must be > 0", e);
    if (d.steps < 0) return dem::set_error(DEM_ERR_INVALID, "ensemble %d: steps < 0", e);
    idx[e] = e;
  }
  // reserved ranges must not overlap
  std::sort(idx.begin(), idx.end(), [&](int a, int b) { return ens[a].offset < ens[b].offset; });
  for (int k = 1; k < n_ens; ++k) {
    const dem_batch_ensemble& a = ens[idx[k - 1]];
    const dem_batch_ensemble& b = ens[idx[k]];
    if (a.capacity > 0 && b.capacity > 0 && (long long)a.offset + a.capacity > b.offset)
      return dem::set_error(DEM_ERR_INVALID, "ensembles %d and %d overlap", idx[k - 1], idx[k]);
  }
  return DEM_OK;
}

// Rebuilds the system table, items and rows from the host descriptions and
// uploads them (synchronising `st` so the host tables can be reused).
// Device descriptor of system e from its host description and the plan.
BEns ens_entry(const dem_batch* b, int e) {
  const dem_batch_ensemble& d = b->desc[e];
  BEns E{};
  E.part_off = b->plan.part_off[e];
  E.cnt_off = b->plan.cnt_off[e];
  E.start_off = b->plan.start_off[e];
  E.partj_off = b->plan.partj_off[e];
  E.off = d.offset;
  E.roff = b->plan.rec_off[e];
  E.n = d.n;
  E.cap = d.capacity;
  if (b->tune.pair_once) {
    const int B = (d.n + dem::kSymB - 1) / dem::kSymB;
    E.S = pair_once_split(d.n, b->tune.jlen);
    E.jchunk = (dem::sym_entries_max(B > 0 ? B : 1) + E.S - 1) / E.S;
  } else {
    E.S = batch_split(d.n, b->tune.jlen);
    E.jchunk = d.n > 0 ? (d.n + E.S - 1) / E.S : 0;
  }
  E.steps = d.steps;
  E.T = next_pow2(2 * (d.capacity > 0 ? d.capacity : 1));
  E.fresh = b->fresh[e];
  E.L = d.L;
  E.dt = d.dt;
  E.sf = d.stokes_factor;
  E.m_const = d.m_const;
  E.eta_const = d.eta_const;
  return E;
}

int upload_tables(dem_batch* b, hipStream_t st) {
  const int bw = b->tune.block;
  b->ens_h.resize(b->n_ens);
  b->items_h.clear();
  b->rows_h.clear();
  for (int e = 0; e < b->n_ens; ++e) b->ens_h[e] = ens_entry(b, e);
  // items of the largest systems first (their items take longest)
  for (int e : b->order) {
    const BEns& E = b->ens_h[e];
    if (b->tune.pair_once) {
      // every chunk of every block: the reductions add all E.S chunk sums
      for (int blk = 0; blk * dem::kSymB < E.n; ++blk)
        for (int s = 0; s < E.S; ++s) b->items_h.push_back(make_int4(e, blk, s, 0));
    } else {
      for (int i0 = 0; i0 < E.n; i0 += bw)
        for (int s = 0; s < E.S; ++s)
          if (s * E.jchunk < E.n) b->items_h.push_back(make_int4(e, i0, s, 0));
    }
  }
  for (int e = 0; e < b->n_ens; ++e)
    for (int i0 = 0; i0 < b->ens_h[e].n; i0 += kRow) b->rows_h.push_back(make_int2(e, i0));
  if ((long long)b->items_h.size() > b->plan.items_cap ||
      (long long)b->rows_h.size() > b->plan.rows_cap)
    return dem::set_error(DEM_ERR_CAPACITY, "work tables exceed their capacity");
  BHead h{};
  h.n_items = (int)b->items_h.size();
  h.n_rows = (int)b->rows_h.size();
  h.n_ens = b->n_ens;
  h.contacts = b->tune.contacts ? 1 : 0;
  h.pair_cap = (int)b->plan.pair_total;
  DEM_CHECK_HIP(hipMemcpyAsync(b->l.ens, b->ens_h.data(), sizeof(BEns) * b->ens_h.size(),
                               hipMemcpyHostToDevice, st));
  if (!b->items_h.empty())
    DEM_CHECK_HIP(hipMemcpyAsync(b->l.items, b->items_h.data(), sizeof(int4) * b->items_h.size(),
                                 hipMemcpyHostToDevice, st));
  if (!b->rows_h.empty())
    DEM_CHECK_HIP(hipMemcpyAsync(b->l.rows, b->rows_h.data(), sizeof(int2) * b->rows_h.size(),
                                 hipMemcpyHostToDevice, st));
  DEM_CHECK_HIP(hipMemcpyAsync(b->l.head, &h, sizeof(h), hipMemcpyHostToDevice, st));
  DEM_CHECK_HIP(hipStreamSynchronize(st));
  return DEM_OK;
}

// Hash cells and records of every system (after any change of radii/positions).
int refresh(dem_batch* b, hipStream_t st) {
  const BLayout& l = b->l;
  hipLaunchKernelGGL(k_b_cells, dim3(b->n_ens), dim3(256), 0, st, (const BEns*)l.ens,
                     (const BHead*)l.head, (const double*)b->radii, l.dev);
  DEM_CHECK_LAUNCH("k_b_cells");
  hipLaunchKernelGGL(k_b_pack, dim3((unsigned)std::max<long long>(b->plan.rows_cap, 1)),
                     dim3(kRow), 0, st, (const BEns*)l.ens, (const BDev*)l.dev,
                     (const int2*)l.rows, (const BHead*)l.head, (const double*)b->pos,
                     (const double*)b->radii, l.rec, l.recc);
  DEM_CHECK_LAUNCH("k_b_pack");
  return DEM_OK;
}

int tile_force(const dem_batch* b) { return b->tune.force_tile > 0 ? b->tune.force_tile : 32; }
int tile_conv(const dem_batch* b) { return b->tune.conv_tile > 0 ? b->tune.conv_tile : 16; }

int enqueue_check(dem_batch* b, hipStream_t st, int new_step) {
  const BLayout& l = b->l;
  const unsigned gr = (unsigned)std::max<long long>(b->plan.rows_cap, 1);
  hipLaunchKernelGGL(k_b_begin, dim3((b->n_ens + 255) / 256), dim3(256), 0, st,
                     (const BEns*)l.ens, l.dev, (const BHead*)l.head, new_step,
                     (const int*)l.steps_done, l.stepped, l.active, l.pair_count, l.inc, l.oob,
                     l.merged, l.t0n, l.t0alive);
  DEM_CHECK_LAUNCH("k_b_begin");
  if (!b->tune.contacts) return DEM_OK;
  if (!new_step) {  // a first check is never incremental
    hipLaunchKernelGGL(k_b_inc_search, dim3(b->n_ens), dim3(256), 0, st, (const BEns*)l.ens,
                       (const BDev*)l.dev, (const BHead*)l.head, (const int*)l.active, l.inc,
                       l.oob, l.t0n, l.t0alive, (const double*)b->pos, (const double*)b->radii,
                       (const int*)l.start, (const int*)l.sorted, (const double4*)l.spos,
                       l.pair_count, l.inc_pairs);
    DEM_CHECK_LAUNCH("k_b_inc_search");
  }
  // the bucket counts are 0 here: cleared at creation and by every bucket scan
  hipLaunchKernelGGL(k_b_assign, dim3(gr), dim3(kRow), 0, st, (const BEns*)l.ens,
                     (const BDev*)l.dev, (const int2*)l.rows, (const BHead*)l.head,
                     (const int*)l.active, (const int*)l.inc, (const double*)b->pos, l.cnt,
                     l.bucket, l.rank, l.oob);
  DEM_CHECK_LAUNCH("k_b_assign");
  hipLaunchKernelGGL(k_b_bucket_scan, dim3(b->n_ens), dim3(kScanWG), 0, st, (const BEns*)l.ens,
                     (const BDev*)l.dev, (const BHead*)l.head, (const int*)l.active,
                     (const int*)l.inc, l.cnt, l.start);
  DEM_CHECK_LAUNCH("k_b_bucket_scan");
  hipLaunchKernelGGL(k_b_scatter, dim3(gr), dim3(kRow), 0, st, (const BEns*)l.ens,
                     (const BDev*)l.dev, (const int2*)l.rows, (const BHead*)l.head,
                     (const int*)l.active, (const int*)l.inc, (const int*)l.start,
                     (const int*)l.bucket, (const int*)l.rank, (const double*)b->pos,
                     (const double*)b->radii, l.sorted, l.spos);
  DEM_CHECK_LAUNCH("k_b_scatter");
  hipLaunchKernelGGL(k_b_search, dim3(gr * kSearchSplit), dim3(kRow), 0, st,
                     (const BEns*)l.ens, (const BDev*)l.dev, (const int2*)l.rows,
                     (const BHead*)l.head, (const int*)l.active, (const int*)l.inc,
                     (const int*)l.start, (const int*)l.sorted, (const double4*)l.spos, l.pc,
                     l.pair_count);
  DEM_CHECK_LAUNCH("k_b_search");
  hipLaunchKernelGGL(k_b_pair_scan, dim3(b->n_ens), dim3(kScanWG), 0, st, (const BEns*)l.ens,
                     (const BHead*)l.head, (const int*)l.pair_count, (const int*)l.inc,
                     (const int*)l.pc, l.poff);
  DEM_CHECK_LAUNCH("k_b_pair_scan");
  hipLaunchKernelGGL(k_b_pair_base, dim3(1), dim3(kScanWG), 0, st, (const BHead*)l.head,
                     (const int*)l.pair_count, l.pair_base);
  DEM_CHECK_LAUNCH("k_b_pair_base");
  hipLaunchKernelGGL(k_b_emit, dim3(gr), dim3(kRow), 0, st, (const BEns*)l.ens,
                     (const BDev*)l.dev, (const int2*)l.rows, (const BHead*)l.head,
                     (const int*)l.pair_count, (const int*)l.inc, (const double*)b->pos,
                     (const double*)b->radii, (const int*)l.start, (const int*)l.sorted,
                     (const double4*)l.spos, (const int*)l.pc, (const int*)l.poff,
                     (const int*)l.pair_base, l.pairs, l.pvals);
  DEM_CHECK_LAUNCH("k_b_emit");
  if (!new_step) {
    hipLaunchKernelGGL(k_b_inc_emit, dim3(b->n_ens), dim3(256), 0, st, (const BEns*)l.ens,
                       (const BHead*)l.head, (const int*)l.inc, (const int*)l.pair_count,
                       (const int*)l.pair_base, (const int2*)l.inc_pairs, (const double*)b->pos,
                       (const double*)b->radii, l.pairs, l.pvals);
    DEM_CHECK_LAUNCH("k_b_inc_emit");
  }
  return DEM_OK;
}

template <int BW>
int enqueue_advance_bw(dem_batch* b, hipStream_t st, bool prof) {
  const BLayout& l = b->l;
  const unsigned gi = (unsigned)std::max<long long>(b->plan.items_cap, 1);
  const unsigned gr = (unsigned)std::max<long long>(b->plan.rows_cap, 1);
  auto force = [&] {
    hipLaunchKernelGGL(k_b_force<BW>, dim3(gi), dim3(BW), 0, st, (const BEns*)l.ens,
                       (const int4*)l.items, (const BHead*)l.head, (const int*)l.go,
                       (const Rec*)l.rec, tile_force(b), l.part);
  };
  auto conv = [&] {
    hipLaunchKernelGGL(k_b_conv<BW>, dim3(gi), dim3(BW), 0, st, (const BEns*)l.ens,
                       (const int4*)l.items, (const BHead*)l.head, (const int*)l.go,
                       (const RecC*)l.recc, tile_conv(b), l.part);
  };
  if (prof) {
    DEM_PROF_LAUNCH("k_b_force", st, force());
  } else {
    force();
  }
  DEM_CHECK_LAUNCH("k_b_force");
  hipLaunchKernelGGL(k_b_force_reduce<false>, dim3(gr), dim3(kRow), 0, st, (const BEns*)l.ens,
                     (const BDev*)l.dev, (const int2*)l.rows, (const BHead*)l.head,
                     (const int*)l.go, (const double*)l.part, (const dem::PartJ*)l.part_j,
                     (const double*)b->radii, b->F, l.recc);
  DEM_CHECK_LAUNCH("k_b_force_reduce");
  if (prof) {
    DEM_PROF_LAUNCH("k_b_conv", st, conv());
  } else {
    conv();
  }
  DEM_CHECK_LAUNCH("k_b_conv");
  hipLaunchKernelGGL(k_b_conv_euler<false>, dim3(gr), dim3(kRow), 0, st, (const BEns*)l.ens,
                     (const BDev*)l.dev, (const int2*)l.rows, (const BHead*)l.head,
                     (const int*)l.go, (const double*)l.part, (const dem::PartJ*)l.part_j,
                     (const double*)b->radii, (const double*)b->F, b->V, b->pos, l.rec, l.recc,
                     l.steps_done, l.time);
  DEM_CHECK_LAUNCH("k_b_conv_euler");
  return DEM_OK;
}

int enqueue_advance_sym(dem_batch* b, hipStream_t st, bool prof) {
  const BLayout& l = b->l;
  const unsigned gi = (unsigned)std::max<long long>(b->plan.items_cap, 1);
  const unsigned gr = (unsigned)std::max<long long>(b->plan.rows_cap, 1);
  constexpr int TF = dem::kSymTplForce, TC = dem::kSymTplConv;
  auto force = [&] {
    hipLaunchKernelGGL(k_b_force_sym<TF>, dim3(gi), dim3(dem::kSymB / TF), 0, st,
                       (const BEns*)l.ens, (const int4*)l.items, (const BHead*)l.head,
                       (const int*)l.go, (const Rec*)l.rec, tile_force(b), l.part, l.part_j);
  };
  auto conv = [&] {
    hipLaunchKernelGGL(k_b_conv_sym<TC>, dim3(gi), dim3(dem::kSymB / TC), 0, st,
                       (const BEns*)l.ens, (const int4*)l.items, (const BHead*)l.head,
                       (const int*)l.go, (const RecC*)l.recc, tile_conv(b), l.part, l.part_j);
  };
  if (prof) {
    DEM_PROF_LAUNCH("k_b_force", st, force());
  } else {
    force();
  }
  DEM_CHECK_LAUNCH("k_b_force_sym");
  hipLaunchKernelGGL(k_b_force_reduce<true>, dim3(gr), dim3(kRow), 0, st, (const BEns*)l.ens,
                     (const BDev*)l.dev, (const int2*)l.rows, (const BHead*)l.head,
                     (const int*)l.go, (const double*)l.part, (const dem::PartJ*)l.part_j,
                     (const double*)b->radii, b->F, l.recc);
  DEM_CHECK_LAUNCH("k_b_force_reduce");
  if (prof) {
    DEM_PROF_LAUNCH("k_b_conv", st, conv());
  } else {
    conv();
  }
  DEM_CHECK_LAUNCH("k_b_conv_sym");
  hipLaunchKernelGGL(k_b_conv_euler<true>, dim3(gr), dim3(kRow), 0, st, (const BEns*)l.ens,
                     (const BDev*)l.dev, (const int2*)l.rows, (const BHead*)l.head,
                     (const int*)l.go, (const double*)l.part, (const dem::PartJ*)l.part_j,
                     (const double*)b->radii, (const double*)b->F, b->V, b->pos, l.rec, l.recc,
                     l.steps_done, l.time);
  DEM_CHECK_LAUNCH("k_b_conv_euler");
  return DEM_OK;
}

int enqueue_advance(dem_batch* b, hipStream_t st, bool prof) {
  if (b->tune.pair_once) return enqueue_advance_sym(b, st, prof);
  switch (b->tune.block) {
    case 64: return enqueue_advance_bw<64>(b, st, prof);
    case 128: return enqueue_advance_bw<128>(b, st, prof);
    default: return enqueue_advance_bw<256>(b, st, prof);
  }
}

int enqueue(dem_batch* b, int k, hipStream_t st, bool prof) {
  return k == 2 ? enqueue_advance(b, st, prof) : enqueue_check(b, st, k == 0 ? 1 : 0);
}

void drop_graphs(dem_batch* b) {
  for (int k = 0; k < 3; ++k) {
    if (b->exec[k]) (void)hipGraphExecDestroy(b->exec[k]);
    if (b->graph[k]) (void)hipGraphDestroy(b->graph[k]);
    b->exec[k] = nullptr;
    b->graph[k] = nullptr;
  }
}

int capture_one(dem_batch* b, int k) {
  DEM_CHECK_HIP(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeRelaxed));
  int rc = enqueue(b, k, b->cap_stream, false);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return dem::set_error(DEM_ERR_HIP, "capture: %s", hipGetErrorString(e));
  b->graph[k] = g;
  e = hipGraphInstantiate(&b->exec[k], g, nullptr, nullptr, 0);
  if (e != hipSuccess) return dem::set_error(DEM_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
  return DEM_OK;
}

// Which systems an advance moves ("go"), then an event that a second stream
// can wait on before it re-checks and merges the others.
int launch_go(dem_batch* b, hipStream_t st) {
  const BLayout& l = b->l;
  hipLaunchKernelGGL(k_b_go, dim3((b->n_ens + 255) / 256), dim3(256), 0, st,
                     (const BEns*)l.ens, (const BHead*)l.head, (const int*)l.steps_done,
                     (const int*)l.pair_count, l.stepped, l.go);
  DEM_CHECK_LAUNCH("k_b_go");
  if (!b->go_done) DEM_CHECK_HIP(hipEventCreateWithFlags(&b->go_done, hipEventDisableTiming));
  DEM_CHECK_HIP(hipEventRecord(b->go_done, st));
  return DEM_OK;
}

// k = 0: first check of a global step, 1: check again, 2: advance.  Work
// tables left stale by dem_batch_merge_async are rebuilt before a first check
// or an advance (a re-check needs only the merged systems' descriptors).
int launch(dem_batch* b, int k, hipStream_t st) {
  if (k != 1 && b->tables_dirty) {
    int rc = upload_tables(b, st);
    if (rc) return rc;
    b->tables_dirty = false;
  }
  if (k == 2) {
    int rc = launch_go(b, st);
    if (rc) return rc;
  }
  if (b->exec[k]) {
    DEM_CHECK_HIP(hipGraphLaunch(b->exec[k], st));
    return DEM_OK;
  }
  return enqueue(b, k, st, true);
}

}  // namespace

extern "C" {

// Greedy merge plans of many systems on the host (merge_plan_rows /
// droplet_solution.py:208-262 per system).  The IEEE operations are those of
// the reference's float64 scalars: r^3 = pow(r, 3), r = pow(c1 + c2, 1/3),
// centre (x1 c1 + x2 c2) / (c1 + c2) per axis, no contraction.
int dem_merge_plan(int count, const int32_t* pair_off, const int32_t* pairs,
                   const double* values, const int32_t* order, int32_t* removed_off,
                   int32_t* removed, int32_t* merged_off, double* merged) {
#pragma clang fp contract(off)
  if (count < 0) return dem::set_error(DEM_ERR_INVALID, "count = %d < 0", count);
  if (count == 0) {
    if (removed_off) removed_off[0] = 0;
    if (merged_off) merged_off[0] = 0;
    return DEM_OK;
  }
  if (!pair_off || !pairs || !values || !order || !removed_off || !removed || !merged_off ||
      !merged)
    return dem::set_error(DEM_ERR_INVALID, "null argument");
  if (pair_off[0] != 0) return dem::set_error(DEM_ERR_INVALID, "pair_off[0] must be 0");
  std::vector<int> done;
  int nr = 0, nm = 0;
  removed_off[0] = 0;
  merged_off[0] = 0;
  for (int k = 0; k < count; ++k) {
    const int a = pair_off[k], c = pair_off[k + 1] - a;
    if (c < 0) return dem::set_error(DEM_ERR_INVALID, "pair offsets must not decrease");
    done.clear();
    for (int q = 0; q < c; ++q) {
      const int o = order[a + q];
      if (o < 0 || o >= c) return dem::set_error(DEM_ERR_INVALID, "order out of range");
      const int p = a + o;
      const int i = pairs[2 * (size_t)p], j = pairs[2 * (size_t)p + 1];
      if (std::find(done.begin(), done.end(), i) != done.end() ||
          std::find(done.begin(), done.end(), j) != done.end())
        continue;
      const double* v = values + 8 * (size_t)p;
      const double c1 = pow(v[3], 3.0);
      const double c2 = pow(v[7], 3.0);
      const double sm = c1 + c2;
      double* m = merged + 4 * (size_t)nm;
      m[0] = (v[0] * c1 + v[4] * c2) / sm;
      m[1] = (v[1] * c1 + v[5] * c2) / sm;
      m[2] = (v[2] * c1 + v[6] * c2) / sm;
      m[3] = pow(sm, 1.0 / 3.0);
      ++nm;
      done.push_back(i);
      done.push_back(j);
    }
    std::sort(done.begin(), done.end());
    for (int d : done) removed[nr++] = d;
    removed_off[k + 1] = nr;
    merged_off[k + 1] = nm;
  }
  return DEM_OK;
}

int dem_batch_split(int n, int jlen) { return batch_split(n, jlen); }

int dem_batch_pair_once_split(int n, int jlen) { return pair_once_split(n, jlen); }

size_t dem_batch_workspace_bytes(const dem_batch_ensemble* ens, int n_ens,
                                 const dem_batch_tuning* tuning) {
  if (check_ensembles(ens, n_ens) || check_tuning(tuning)) return 0;
  Plan p = make_plan(ens, n_ens, tuning->block, tuning->jlen, tuning->pair_once);
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
  Plan p = make_plan(ens, n_ens, tuning->block, tuning->jlen, tuning->pair_once);
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
  b->fresh.assign(n_ens, 0);
  b->order.resize(n_ens);
  std::iota(b->order.begin(), b->order.end(), 0);
  std::stable_sort(b->order.begin(), b->order.end(),
                   [&](int x, int y) { return ens[x].capacity > ens[y].capacity; });
  b->plan = p;
  b->l = l;
  for (int e = 0; e < n_ens; ++e) b->cap_max = std::max(b->cap_max, ens[e].capacity);
  b->pos = pos;
  b->radii = radii;
  b->F = forces;
  b->V = velocities;
  hipStream_t st = dem::as_stream(stream);
  rc = upload_tables(b, st);
  if (!rc) {
    hipLaunchKernelGGL(k_b_init, dim3((n_ens + 255) / 256), dim3(256), 0, st,
                       (const BEns*)l.ens, l.dev, n_ens, l.active, l.stepped, l.go,
                       l.pair_count, l.steps_done, l.time, l.inc, l.oob, l.merged, l.t0n,
                       l.t0alive);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = dem::set_error(DEM_ERR_HIP, "k_b_init: %s", hipGetErrorString(e));
  }
  if (!rc) {  // bucket counts start at 0; each bucket scan leaves them at 0
    hipLaunchKernelGGL(k_b_zero, dim3(kZeroBlocks), dim3(256), 0, st, l.cnt, p.cnt_total);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = dem::set_error(DEM_ERR_HIP, "k_b_zero: %s", hipGetErrorString(e));
  }
  if (!rc) rc = refresh(b, st);
  if (rc) {
    delete b;
    return rc;
  }
  *out = b;
  return DEM_OK;
}

int dem_batch_resize(dem_batch_t b, int count, const int32_t* systems, const int32_t* n,
                     dem_stream_t stream) {
  if (!b) return dem::set_error(DEM_ERR_INVALID, "null handle");
  if (count < 0 || (count > 0 && (!systems || !n)))
    return dem::set_error(DEM_ERR_INVALID, "bad system list");
  for (int k = 0; k < count; ++k) {
    const int e = systems[k];
    if (e < 0 || e >= b->n_ens) return dem::set_error(DEM_ERR_INVALID, "no ensemble %d", e);
    if (n[k] < 0 || n[k] > b->desc[e].capacity)
      return dem::set_error(DEM_ERR_CAPACITY, "ensemble %d: n=%d outside [0, %d]", e, n[k],
                            b->desc[e].capacity);
  }
  hipStream_t st = dem::as_stream(stream);
  DEM_CHECK_HIP(hipStreamSynchronize(st));
  for (int k = 0; k < count; ++k) {
    b->desc[systems[k]].n = n[k];
    b->fresh[systems[k]] = 0;
  }
  int rc = upload_tables(b, st);
  if (!rc) b->tables_dirty = false;
  if (rc) return rc;
  return refresh(b, st);
}

static int merge_systems(dem_batch_t b, int count, const int32_t* systems,
                         const int32_t* removed_off, const int32_t* removed,
                         const int32_t* merged_off, const double* merged, dem_stream_t stream,
                         bool defer) {
  if (!b) return dem::set_error(DEM_ERR_INVALID, "null handle");
  if (count < 0 || count > b->n_ens) return dem::set_error(DEM_ERR_INVALID, "bad count %d", count);
  if (count == 0) return DEM_OK;
  if (!systems || !removed_off || !merged_off || (removed_off[count] > 0 && !removed) ||
      (merged_off[count] > 0 && !merged))
    return dem::set_error(DEM_ERR_INVALID, "null merge list");
  if (removed_off[0] != 0 || merged_off[0] != 0)
    return dem::set_error(DEM_ERR_INVALID, "merge lists must start at offset 0");
  std::vector<int> seen(b->n_ens, 0);
  std::vector<int> new_n(count);
  for (int k = 0; k < count; ++k) {
    const int e = systems[k];
    if (e < 0 || e >= b->n_ens || seen[e]++)
      return dem::set_error(DEM_ERR_INVALID, "bad or repeated system %d", e);
    const int n = b->desc[e].n;
    const int r0 = removed_off[k], r1 = removed_off[k + 1];
    const int m0 = merged_off[k], m1 = merged_off[k + 1];
    if (r1 < r0 || m1 < m0) return dem::set_error(DEM_ERR_INVALID, "offsets must not decrease");
    for (int q = r0; q < r1; ++q)
      if (removed[q] < 0 || removed[q] >= n || (q > r0 && removed[q] <= removed[q - 1]))
        return dem::set_error(DEM_ERR_INVALID,
                              "system %d: removed indices must be ascending in [0, %d)", e, n);
    new_n[k] = n - (r1 - r0) + (m1 - m0);
    if (new_n[k] < 0 || new_n[k] > b->desc[e].capacity)
      return dem::set_error(DEM_ERR_CAPACITY, "system %d: %d droplets after merging", e, new_n[k]);
  }
  const long long nr = removed_off[count], nm = merged_off[count];
  if (nr > b->plan.cap_total || nm > b->plan.cap_total)
    return dem::set_error(DEM_ERR_CAPACITY, "merge lists exceed the staging area");
  hipStream_t st = dem::as_stream(stream);
  DEM_CHECK_HIP(hipStreamSynchronize(st));  // stage_h may still feed a copy
  b->stage_h.assign(systems, systems + count);
  b->stage_h.insert(b->stage_h.end(), removed_off, removed_off + count + 1);
  b->stage_h.insert(b->stage_h.end(), merged_off, merged_off + count + 1);
  if (nr) b->stage_h.insert(b->stage_h.end(), removed, removed + nr);
  DEM_CHECK_HIP(hipMemcpyAsync(b->l.stage_i, b->stage_h.data(), sizeof(int) * b->stage_h.size(),
                               hipMemcpyHostToDevice, st));
  if (nm)
    DEM_CHECK_HIP(hipMemcpyAsync(b->l.stage_d, merged, sizeof(double) * 4 * nm,
                                 hipMemcpyHostToDevice, st));
  const int rows_max = std::max(1, (b->cap_max + kRow - 1) / kRow);
  const long long blocks = (long long)count * rows_max;
  if (blocks > 0x7fffffffLL) return dem::set_error(DEM_ERR_CAPACITY, "too many rows to merge");
  hipLaunchKernelGGL(k_b_compact_gather, dim3((unsigned)blocks), dim3(256), 0, st,
                     (const BEns*)b->l.ens, count, rows_max, (const int*)b->l.stage_i,
                     (const double*)b->pos, (const double*)b->radii, b->l.cscr, b->l.sorted,
                     (const int*)b->l.t0n, b->l.t0alive);
  DEM_CHECK_LAUNCH("k_b_compact_gather");
  hipLaunchKernelGGL(k_b_compact_scatter, dim3((unsigned)blocks), dim3(256), 0, st,
                     (const BEns*)b->l.ens, count, rows_max, (const int*)b->l.stage_i,
                     (const double*)b->l.stage_d, (const double4*)b->l.cscr, b->pos, b->radii,
                     b->l.merged);
  DEM_CHECK_LAUNCH("k_b_compact_scatter");
  for (int k = 0; k < count; ++k) {
    b->desc[systems[k]].n = new_n[k];
    b->fresh[systems[k]] = b->incremental ? merged_off[k + 1] - merged_off[k] : 0;
  }
  if (!defer) {
    int rc = upload_tables(b, st);  // synchronises: the host lists may be freed after return
    if (rc) return rc;
    b->tables_dirty = false;
    return refresh(b, st);
  }
  // only the merged systems' descriptors, hash cells and records change; the
  // work tables are rebuilt before the next first check or advance
  const BLayout& l = b->l;
  b->ens_stage_h.resize(count);
  for (int k = 0; k < count; ++k) {
    const int e = systems[k];
    b->ens_h[e] = ens_entry(b, e);
    b->ens_stage_h[k] = b->ens_h[e];
  }
  DEM_CHECK_HIP(hipMemcpyAsync(l.ens_stage, b->ens_stage_h.data(), sizeof(BEns) * count,
                               hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_b_ens_scatter, dim3((count + 255) / 256), dim3(256), 0, st, count,
                     (const int*)l.stage_i, (const BEns*)l.ens_stage, l.ens);
  DEM_CHECK_LAUNCH("k_b_ens_scatter");
  hipLaunchKernelGGL(k_b_cells_list, dim3(count), dim3(256), 0, st, (const BEns*)l.ens, count,
                     (const int*)l.stage_i, (const double*)b->radii, l.dev);
  DEM_CHECK_LAUNCH("k_b_cells_list");
  hipLaunchKernelGGL(k_b_pack_list, dim3((unsigned)blocks), dim3(kRow), 0, st,
                     (const BEns*)l.ens, (const BDev*)l.dev, count, rows_max,
                     (const int*)l.stage_i, (const double*)b->pos, (const double*)b->radii, l.rec,
                     l.recc);
  DEM_CHECK_LAUNCH("k_b_pack_list");
  b->tables_dirty = true;
  DEM_CHECK_HIP(hipStreamSynchronize(st));  // the host lists may be freed after return
  return DEM_OK;
}

int dem_batch_merge(dem_batch_t b, int count, const int32_t* systems, const int32_t* removed_off,
                    const int32_t* removed, const int32_t* merged_off, const double* merged,
                    dem_stream_t stream) {
  return merge_systems(b, count, systems, removed_off, removed, merged_off, merged, stream, false);
}

int dem_batch_merge_async(dem_batch_t b, int count, const int32_t* systems,
                          const int32_t* removed_off, const int32_t* removed,
                          const int32_t* merged_off, const double* merged, dem_stream_t stream) {
  return merge_systems(b, count, systems, removed_off, removed, merged_off, merged, stream, true);
}

int dem_batch_set_incremental(dem_batch_t b, int on) {
  if (!b) return dem::set_error(DEM_ERR_INVALID, "null handle");
  if (on != 0 && on != 1) return dem::set_error(DEM_ERR_INVALID, "on must be 0 or 1");
  b->incremental = on != 0;
  if (!on) std::fill(b->fresh.begin(), b->fresh.end(), 0);
  return DEM_OK;
}

int dem_batch_wait_go(dem_batch_t b, dem_stream_t stream) {
  if (!b) return dem::set_error(DEM_ERR_INVALID, "null handle");
  if (!b->go_done) return dem::set_error(DEM_ERR_INVALID, "no advance has been launched");
  DEM_CHECK_HIP(hipStreamWaitEvent(dem::as_stream(stream), b->go_done, 0));
  return DEM_OK;
}

int dem_batch_check(dem_batch_t b, int new_step, dem_stream_t stream) {
  if (!b) return dem::set_error(DEM_ERR_INVALID, "null handle");
  return launch(b, new_step ? 0 : 1, dem::as_stream(stream));
}

int dem_batch_advance(dem_batch_t b, dem_stream_t stream) {
  if (!b) return dem::set_error(DEM_ERR_INVALID, "null handle");
  return launch(b, 2, dem::as_stream(stream));
}

int dem_batch_capture(dem_batch_t b) {
  if (!b) return dem::set_error(DEM_ERR_INVALID, "null handle");
  drop_graphs(b);
  if (!b->cap_stream) DEM_CHECK_HIP(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
  int rc = DEM_OK;
  for (int k = 0; k < 3 && !rc; ++k) rc = capture_one(b, k);
  if (rc) drop_graphs(b);
  return rc;
}

int dem_batch_run(dem_batch_t b, int steps, dem_stream_t stream) {
  if (!b) return dem::set_error(DEM_ERR_INVALID, "null handle");
  if (steps < 0) return dem::set_error(DEM_ERR_INVALID, "steps < 0");
  hipStream_t st = dem::as_stream(stream);
  for (int k = 0; k < steps; ++k) {
    int rc = launch(b, 0, st);
    if (!rc) rc = launch(b, 2, st);
    if (rc) return rc;
  }
  return DEM_OK;
}

int dem_batch_field(dem_batch_t b, const char* name, void** ptr, size_t* bytes) {
  if (!b || !name || !ptr || !bytes) return dem::set_error(DEM_ERR_INVALID, "null argument");
  const size_t ne = (size_t)b->n_ens;
  const size_t pc = (size_t)std::max<long long>(b->plan.pair_total, 1);
  struct F {
    const char* n;
    void* p;
    size_t sz;
  } fields[] = {{"active", b->l.active, ne * sizeof(int)},
                {"stepped", b->l.stepped, ne * sizeof(int)},
                {"pair_count", b->l.pair_count, ne * sizeof(int)},
                {"pair_base", b->l.pair_base, (ne + 1) * sizeof(int)},
                {"incremental", b->l.inc, ne * sizeof(int)},
                {"steps_done", b->l.steps_done, ne * sizeof(int)},
                {"time", b->l.time, ne * sizeof(double)},
                {"pairs", b->l.pairs, pc * sizeof(int2)},
                {"pair_values", b->l.pvals, 2 * pc * sizeof(double4)}};
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
  drop_graphs(b);
  if (b->go_done) (void)hipEventDestroy(b->go_done);
  if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
  delete b;
  return DEM_OK;
}

}  // extern "C"
