// Flat Barnes-Hut octree for gfx950: build, aggregates and stackless walks.
//
// Reference: dem/octree/flat_tree.py (FlatOctree) and octree_node.py.
//
// Structure parity.  The reference builds level by level: every node with
// count > mpl is split into 8 children (empty ones included), nodes of a level
// are processed in index order, children get consecutive indices, and a split
// is skipped once node_count + 8 would exceed max_nodes (flat_tree.py:308-354).
// Node indices are therefore breadth-first with children in octant order,
// which within a level is Morton order.  This build reproduces that exact
// numbering:
//   1. every particle computes its octant path (21 levels, 63-bit key) by
//      replaying the reference's float64 bound recursion (cx = (min+max)*0.5,
//      bit = x >= cx; :283-306, octree_node.py:83-95), so particles exactly on a
//      split plane land where the reference puts them;
//   2. a stable radix sort orders particles by key: every node owns a
//      contiguous range of the sorted array;
//   3. level by level, a scan over the level's split candidates gives each
//      split node its rank, and the capacity clamp becomes "rank < budget";
//      children ranges come from binary searches of the next octant digit.
// Leaf slots (leaf_start / leaf_indices, :370-397), next pointers (two
// top-down passes, :399-463) and the bottom-up R^3 / force-moment aggregates
// (:469-777) follow the reference's definitions; leaf contents are listed in
// sorted order (the reference's order inside a leaf is an atomic race).
//
// Walks (:783-1612).  One target per lane, targets in Morton order so the
// lanes of a wave walk similar paths.  The opening test
// `s*s < theta^2 * D^2 and D^2 > 0` and the skip tests are evaluated in
// float64 exactly as the reference writes them, so the set of opened nodes
// matches; the interaction arithmetic is FP32 (fixed-point positions as in
// the direct kernel), accumulated in float64.
#include <math.h>
#include <string.h>

#include "dem_common.h"
#include "pair_math.h"
#include "profile.h"
#include "scan.h"
#include "sort.h"
#include "tile_scan.h"

namespace {

constexpr int kKeyLevels = 21;  // 3 * 21 = 63 key bits
constexpr int kMaxLevels = 40;  // reference MAX_LEVELS
constexpr double kTwo32 = 4294967296.0;
constexpr double kSigma = 0x1p-22;
constexpr double kSigma3 = 0x1p-66;
constexpr double kR3Skip = 1e-30;   // flat_tree.py:29
constexpr double kFsqSkip = 1e-60;  // flat_tree.py:32

// kExactMac: a node smaller than 2^22 fixed units, whose opening test always
// takes the float64 path (its fixed-point centre is not precise enough for
// the far-field separation).
int g_exact_mac = 0;  // test switch: every opening test on the float64 path

enum NodeFlags { kSkipR3 = 1, kSkipF = 2, kExactMac = 4 };

// Per-node walk record (64 B): what the walks read at every visit.
struct alignas(16) NodeW {
  uint32_t qx, qy, qz;  // centre in fixed point, quantized as the particles
  float s2f;            // s^2 in fixed units^2, s = largest bbox extent
  float sf;             // s in fixed units
  int first_child, next, count, mstart;
  float r3w;            // R3_sum in work units^3
  int flags;
  int pad[5];
};
static_assert(sizeof(NodeW) == 64, "node record");

// Per-node float64 centre and s^2: the reference's opening test, read when
// the FP32 test cannot decide.
struct alignas(16) NodeC {
  double cx, cy, cz;  // R^3-weighted centre (geometric centre if R3_sum == 0)
  double s2;
};

// Per-node Stokeslet data: force sum / Fs and the centred dipole tensor
// D_ab = P_ab - cm_a F_b divided by Fs*U (fixed units), kept as the parts the
// far-field velocity uses (formed in float64): trace, antisymmetric part
// a_xy = D_yx - D_xy, a_xz = D_zx - D_xz, a_yz = D_zy - D_yz, and the
// symmetric quadratic form q(n) = sum_ab D_ab n_a n_b as diagonal and doubled
// off-diagonal coefficients.
struct alignas(16) NodeF {
  float fx, fy, fz, tr;
  float axy, axz, ayz;
  float sxx, syy, szz, sxy2, sxz2, syz2;
  float pad[3];
};
static_assert(sizeof(NodeF) == 64, "node force record");

// FP32 opening test of a node at a target from the fixed-point separation R
// (centre - x_i, fixed units): 1 take, 0 open, -1 undecided.  With both
// points quantized to within half a unit, |R|^2 is within 2^-21 |D|^2 +
// 3.5 |D| + 3.1 of the exact |D|^2, so m = theta^2 |R|^2 - s^2 is within
// band = 2^-17 s^2 + 8 theta s + 8 of the exact value wherever
// |D| <= 2 s / theta; farther out m > 0 with a margin far above the error.
// A decision outside the band is therefore the float64 test's (s^2 <
// theta^2 d^2, d^2 > 0).  A component within 2^11 units of half the box
// could take the other periodic image than the float64 minimum image:
// undecided too.
__device__ __forceinline__ int fast_open_test(float Rx, float Ry, float Rz, float th2f,
                                              float band_s, float s2f, float sf) {
  constexpr float kHalfBand = 2147481600.f;  // 2^31 - 2^11
  const float d2 = Rx * Rx + Ry * Ry + Rz * Rz;
  const float m = fmaf(th2f, d2, -s2f);
  const float band = fmaf(0x1p-17f, s2f, fmaf(band_s, sf, 8.f));
  const float amax = fmaxf(fabsf(Rx), fmaxf(fabsf(Ry), fabsf(Rz)));
  if (!(fabsf(m) > band) || amax >= kHalfBand) return -1;
  return m > 0.f ? 1 : 0;
}

// Far-field velocity of a node at a target (FP32, fixed units): the Oseen
// monopole of the force sum and the dipole correction (flat_tree.py:
// 1242-1265) for r = x_i - centre = -R, written with the unit vector
// n = r / |r| to stay inside the FP32 range:
//   v += |r|^-2 ((D^T - D) n - tr(D) n + 3 n (n.D n)).
// Outside the walk kernels so that FP contraction applies.
__device__ __forceinline__ void node_velocity(float Rx, float Ry, float Rz, const NodeF& nf,
                                              float& vx, float& vy, float& vz) {
  dem::oseen_pair(Rx, Ry, Rz, nf.fx, nf.fy, nf.fz, true, vx, vy, vz);
  const float ir = __builtin_amdgcn_rsqf(Rx * Rx + Ry * Ry + Rz * Rz);
  const float ir2 = ir * ir;
  const float nx = -Rx * ir, ny = -Ry * ir, nz = -Rz * ir;
  const float ex = nf.axy * ny + nf.axz * nz;
  const float ey = nf.ayz * nz - nf.axy * nx;
  const float ez = -(nf.axz * nx + nf.ayz * ny);
  const float q = nf.sxx * nx * nx + nf.syy * ny * ny + nf.szz * nz * nz +
                  nf.sxy2 * nx * ny + nf.sxz2 * nx * nz + nf.syz2 * ny * nz;
  const float q3 = 3.f * q;
  vx += ir2 * ((ex - nf.tr * nx) + q3 * nx);
  vy += ir2 * ((ey - nf.tr * ny) + q3 * ny);
  vz += ir2 * ((ez - nf.tr * nz) + q3 * nz);
}

// Sorted source record: fixed-point position, radius and r^3 in work units,
// and the source force / Fs (set when forces are propagated to the tree).
struct RecF {
  uint32_t x, y, z;
  float r, r3;
  float f[3];
};

// Two consecutive records, read with one 64-byte load.
typedef uint32_t RecWords __attribute__((ext_vector_type(16), aligned(4)));
struct RecF2 {
  RecF a, b;
};

__device__ __forceinline__ RecF2 load_two(const RecF* __restrict__ srec, int j) {
  const RecWords w = *reinterpret_cast<const RecWords*>(srec + j);
  RecF2 o;
  o.a.x = w[0];
  o.a.y = w[1];
  o.a.z = w[2];
  o.a.r = __uint_as_float(w[3]);
  o.a.r3 = __uint_as_float(w[4]);
  o.a.f[0] = __uint_as_float(w[5]);
  o.a.f[1] = __uint_as_float(w[6]);
  o.a.f[2] = __uint_as_float(w[7]);
  o.b.x = w[8];
  o.b.y = w[9];
  o.b.z = w[10];
  o.b.r = __uint_as_float(w[11]);
  o.b.r3 = __uint_as_float(w[12]);
  o.b.f[0] = __uint_as_float(w[13]);
  o.b.f[1] = __uint_as_float(w[14]);
  o.b.f[2] = __uint_as_float(w[15]);
  return o;
}

struct Scales {
  double origin[3];
  double inv_U, U, lambda, Fs;
  int periodic, pad;
};

// State of the level-by-level split kept on the device, so that levels run
// back to back without a host round trip per level.
struct LevelRec {
  int node_count, parent_count;  // after the levels split so far
  int ls, le;                    // node range of the next level to examine
  int stop;                      // 1: a level split nothing; later levels do nothing
  int num_levels, max_depth;
  int pad;
  int level_offsets[DEM_TREE_MAX_LEVELS];
};

// Workspace layout.
struct TreeLayout {
  Scales* sc;
  double* bbox;                  // [256][6]
  // nodes (M = max_nodes)
  double* bmin;                  // [3M]
  double* bmax;                  // [3M]
  double* r3sum;                 // [M]
  double* r3c;                   // [3M]
  double* fsum;                  // [3M]
  double* fmom;                  // [9M] (xx,xy,xz,yx,yy,yz,zx,zy,zz)
  double* maxr;                  // [M]
  int* first_child;              // [M]
  int* next;                     // [M]
  int* leaf_start;               // [M]
  int* count;                    // [M]
  int* mstart;                   // [M]
  int* parents;                  // [M]
  int* pfc;                      // [M]
  int* flag;                     // [M]
  int* rank;                     // [M]
  int* leafcnt;                  // [M]
  NodeW* nw;                     // [M]
  NodeC* nc;                     // [M]
  NodeF* nf;                     // [M]
  int* counter;                  // [16]
  struct LevelRec* lrec;         // the level loop's device-side record, two slots
  // particles (N = capacity)
  unsigned long long* keys;      // [N]
  unsigned long long* keys_alt;  // [N]
  int* perm;                     // [N] sorted -> original
  int* perm_alt;                 // [N]
  int* leaf_indices;             // [N]
  double* spos;                  // [3N] sorted float64 positions
  double* srad;                  // [N]
  RecF* srec;                    // [N + 1]
  double* pforces;               // [3N] forces of the last force walk (original order)
  double4* packed;               // [N] (x, y, z, r) in original order (build input copy)
  int* scan_ws;
  int* sort_ws;
  size_t bytes;
};

TreeLayout carve(void* ws, int ncap, int M) {
  dem::Carver c(ws);
  TreeLayout l;
  const size_t N = (size_t)(ncap > 0 ? ncap : 1);
  const size_t m = (size_t)(M > 0 ? M : 1);
  l.sc = c.take<Scales>(1);
  l.bbox = c.take<double>(256 * 6);
  l.bmin = c.take<double>(3 * m);
  l.bmax = c.take<double>(3 * m);
  l.r3sum = c.take<double>(m);
  l.r3c = c.take<double>(3 * m);
  l.fsum = c.take<double>(3 * m);
  l.fmom = c.take<double>(9 * m);
  l.maxr = c.take<double>(m);
  l.first_child = c.take<int>(m);
  l.next = c.take<int>(m);
  l.leaf_start = c.take<int>(m);
  l.count = c.take<int>(m);
  l.mstart = c.take<int>(m);
  l.parents = c.take<int>(m);
  l.pfc = c.take<int>(m);
  l.flag = c.take<int>(m);
  l.rank = c.take<int>(m);
  l.leafcnt = c.take<int>(m);
  l.nw = c.take<NodeW>(m);
  l.nc = c.take<NodeC>(m);
  l.nf = c.take<NodeF>(m);
  l.counter = c.take<int>(16);
  l.lrec = c.take<LevelRec>(2);
  l.keys = c.take<unsigned long long>(N);
  l.keys_alt = c.take<unsigned long long>(N);
  l.perm = c.take<int>(N);
  l.perm_alt = c.take<int>(N);
  l.leaf_indices = c.take<int>(N);
  l.spos = c.take<double>(3 * N);
  l.srad = c.take<double>(N);
  l.srec = c.take<RecF>(N + 1);  // + one spare record (leaf loops read members in twos)
  l.pforces = c.take<double>(3 * N);
  l.packed = c.take<double4>(N);
  const size_t big = N > m ? N : m;
  l.scan_ws = c.take<int>(dem::scan_workspace_ints((int)big));
  l.sort_ws = c.take<int>(dem::radix_sort_workspace_ints((int)N));
  l.bytes = dem::align_up(c.used, 256);
  return l;
}

// ---------------------------------------------------------------- helpers

__device__ __forceinline__ double mic64(double d, double L, int periodic) {
  if (periodic && fabs(d) > L * 0.5) d -= copysign(L, d);
  return d;
}

__global__ void k_fill_i(int* p, int n, int v) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_fill_d(double* p, size_t n, double v) {
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------- build

// Octant path of every particle: replays the reference's bound recursion
// from the root [0, L]^3 (flat_tree.py:250-259, :338-349).
__global__ __launch_bounds__(256) void k_keys(const double* __restrict__ pos,
                                              const double* __restrict__ radii, int n, double L,
                                              unsigned long long* __restrict__ keys,
                                              int* __restrict__ perm, double4* __restrict__ packed) {
#pragma clang fp contract(off)
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double x = pos[3 * (size_t)i], y = pos[3 * (size_t)i + 1], z = pos[3 * (size_t)i + 2];
  // one aligned 32-byte record per droplet: the sorted gather reads one line
  packed[i] = make_double4(x, y, z, radii[i]);
  double mnx = 0.0, mny = 0.0, mnz = 0.0, mxx = L, mxy = L, mxz = L;
  unsigned long long key = 0ull;
  for (int l = 0; l < kKeyLevels; ++l) {
    const double cx = (mnx + mxx) * 0.5, cy = (mny + mxy) * 0.5, cz = (mnz + mxz) * 0.5;
    const int bx = x >= cx, by = y >= cy, bz = z >= cz;
    key = (key << 3) | (unsigned long long)(bx | (by << 1) | (bz << 2));
    if (bx) mnx = cx; else mxx = cx;
    if (by) mny = cy; else mxy = cy;
    if (bz) mnz = cz; else mxz = cz;
  }
  keys[i] = key;
  perm[i] = i;
}

__device__ __forceinline__ int digit_at(unsigned long long key, int level) {
  // octant digit that selects the child at `level` (1..21)
  return (int)((key >> (3 * (kKeyLevels - level))) & 7ull);
}

// The root node [0, L]^3 holding all n particles, and the level loop's start.
__device__ void root_init(double L, int n, double* bmin, double* bmax, int* first_child, int* next,
                          int* leaf_start, int* count, int* mstart, int* counter, LevelRec* rec) {
  counter[1] = 0;  // leaf slots (atomic total)
  counter[2] = 0;  // split candidates of level 0 (atomic total; levels alternate 2, 3)
  LevelRec r{};
  r.node_count = 1;
  r.ls = 0;
  r.le = 1;
  *rec = r;
  bmin[0] = bmin[1] = bmin[2] = 0.0;
  bmax[0] = bmax[1] = bmax[2] = L;
  first_child[0] = -1;
  next[0] = -1;
  leaf_start[0] = -1;
  count[0] = n;
  mstart[0] = 0;
}

// Split candidates of one level: leaves with count > mpl (depth-limited by the key).
__device__ __forceinline__ bool split_candidate(int level, int mpl, int count) {
  return level < kKeyLevels && count > mpl;
}

// Tile scan of the level's candidate flags: local ranks, tile sums and the
// candidate total (*cand, cleared beforehand).  The level's node range comes
// from its device record; a stopped loop scans nothing.
__global__ __launch_bounds__(dem::kScanThreads) void k_level_scan(
    int mpl, int level, const LevelRec* __restrict__ rec, const int* __restrict__ count,
    int* __restrict__ rank, int* __restrict__ tile_sums, int* __restrict__ cand) {
  if (rec->stop) return;
  const int ls = rec->ls, cnt = rec->le - rec->ls;
  if ((int)blockIdx.x >= dem::scan_tiles(cnt)) return;
  dem::ts_tile_scan([=](int k) { return split_candidate(level, mpl, count[ls + k]) ? 1 : 0; },
                    cnt, rank, tile_sums, cand);
}

// Split the first k candidates (in node order) of the level, k the capacity
// clamp of the candidates (node_count + 8 must stay within max_nodes,
// flat_tree.py:325-327); eight lanes per node: lane `oct` finds where its
// octant starts in the node's sorted range (a binary search on the next
// octant digit) and writes that child; its end is the next lane's start.  The
// candidate rank is the tile-local rank plus the prefix of the tile sums.
// Every block derives k and the next level's record from this level's record
// (read only); block 0 writes the next record (the other slot) and clears the
// next level's candidate total.  A level that splits nothing stops the loop,
// as the reference's does.
__global__ __launch_bounds__(256) void k_level_split(
    int level, int mpl, int max_nodes, const LevelRec* __restrict__ rec,
    LevelRec* __restrict__ rec_next, const int* __restrict__ cand, int* __restrict__ cand_next,
    const int* __restrict__ rank_local, const int* __restrict__ tile_sums,
    const unsigned long long* __restrict__ keys, double* bmin, double* bmax, int* first_child,
    int* next, int* leaf_start, int* count, int* mstart, int* parents, int* pfc) {
#pragma clang fp contract(off)
  extern __shared__ int pre[];
  const LevelRec r = *rec;
  int budget = 0;
  if (!r.stop) {
    const int c = *cand;
    const int room = (max_nodes - r.node_count) / 8;
    budget = c < room ? c : (room > 0 ? room : 0);
    if (r.le <= r.ls || r.num_levels >= DEM_TREE_MAX_LEVELS) budget = 0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    LevelRec nx = r;
    if (r.stop || budget == 0) {
      nx.stop = 1;
    } else {
      nx.level_offsets[nx.num_levels++] = r.parent_count;
      nx.parent_count = r.parent_count + budget;
      nx.ls = r.le;
      nx.le = r.node_count + 8 * budget;
      nx.node_count = nx.le;
      nx.max_depth = level + 1;
    }
    *rec_next = nx;
    *cand_next = 0;
  }
  if (budget == 0) return;  // uniform
  const int ls = r.ls, le = r.le;
  const int node_base = r.node_count, pc_base = r.parent_count;
  const int cnt = le - ls;
  if ((int)blockIdx.x * 32 >= cnt) return;  // 32 nodes per block; uniform per block
  dem::ts_tile_prefix(tile_sums, dem::scan_tiles(cnt), pre);
  const int gid = blockIdx.x * 256 + threadIdx.x;
  const int kk = gid >> 3, oct = gid & 7;
  const int node = ls + kk;
  bool work = node < le && split_candidate(level, mpl, count[node]);
  int q = 0;
  if (work) {
    q = rank_local[kk] + pre[kk / dem::kScanTile];
    work = q < budget;
  }
  int start = 0, s1 = 0;
  if (work) {
    const int s0 = mstart[node];
    s1 = s0 + count[node];
    // first position in [s0, s1) whose digit is >= oct
    int a = s0, b = s1;
    if (oct > 0)
      while (a < b) {
        const int mid = (a + b) >> 1;
        if (digit_at(keys[mid], level + 1) < oct) a = mid + 1; else b = mid;
      }
    start = a;
  }
  // the lanes of a node are eight consecutive lanes of one wave
  const int nxt = __shfl_down(start, 1, 8);
  if (!work) return;
  const int end = oct == 7 ? s1 : nxt;
  const int base = node_base + 8 * q;
  if (oct == 0) {
    first_child[node] = base;
    parents[pc_base + q] = node;
    pfc[pc_base + q] = base;
  }
  const double mnx = bmin[3 * node], mny = bmin[3 * node + 1], mnz = bmin[3 * node + 2];
  const double mxx = bmax[3 * node], mxy = bmax[3 * node + 1], mxz = bmax[3 * node + 2];
  const double cx = (mnx + mxx) * 0.5, cy = (mny + mxy) * 0.5, cz = (mnz + mxz) * 0.5;
  const int child = base + oct;
  bmin[3 * child] = (oct & 1) ? cx : mnx;
  bmin[3 * child + 1] = (oct & 2) ? cy : mny;
  bmin[3 * child + 2] = (oct & 4) ? cz : mnz;
  bmax[3 * child] = (oct & 1) ? mxx : cx;
  bmax[3 * child + 1] = (oct & 2) ? mxy : cy;
  bmax[3 * child + 2] = (oct & 4) ? mxz : cz;
  first_child[child] = -1;
  next[child] = -1;
  leaf_start[child] = -1;
  count[child] = end - start;
  mstart[child] = start;
}

__device__ __forceinline__ int leaf_slots(int first_child, int count) {
  return (first_child < 0 && count > 0) ? count : 0;
}

// Tile scan of the leaves' particle counts; the total (counter[1], cleared by
// k_root) is the number of leaf slots.
__global__ __launch_bounds__(dem::kScanThreads) void k_leaf_scan(
    int nn, const int* __restrict__ first_child, const int* __restrict__ count,
    int* __restrict__ off_local, int* __restrict__ tile_sums, int* __restrict__ counter) {
  dem::ts_tile_scan([=](int i) { return leaf_slots(first_child[i], count[i]); }, nn, off_local,
                    tile_sums, counter + 1);
}

// leaf_start from the scan of leaf counts (flat_tree.py:370-385),
// leaf_indices filled in sorted order (:387-397).
__global__ __launch_bounds__(256) void k_leaf_fill(
    int nn, const int* __restrict__ first_child, const int* __restrict__ count,
    const int* __restrict__ mstart, const int* __restrict__ off_local,
    const int* __restrict__ tile_sums, int nt, const int* __restrict__ perm,
    int* __restrict__ leaf_start, int* __restrict__ leaf_indices) {
  extern __shared__ int pre[];
  dem::ts_tile_prefix(tile_sums, nt, pre);
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nn) return;
  if (first_child[i] < 0 && count[i] > 0) {
    const int ls = off_local[i] + pre[i / dem::kScanTile];
    leaf_start[i] = ls;
    const int m0 = mstart[i];
    for (int k = 0; k < count[i]; ++k) leaf_indices[ls + k] = perm[m0 + k];
  } else {
    leaf_start[i] = -1;
  }
}

// Next pointers and first children (flat_tree.py:399-463) in one launch.  The
// reference sets them in two top-down sweeps: pass 1 chains each parent's
// eight children (child 7 takes the parent's next), pass 2 relinks the
// occupied children (each to the next occupied sibling, the last to the
// parent's next) and points the parent at its first occupied child.  Both
// have closed forms, evaluated here per node by walking up the tree: children
// of the P-th parent (global parent index) are nodes 1 + 8P .. 8 + 8P.
//   pass-1 next of x: x + 1 if x is not its parent's last child, else the
//     pass-1 next of the parent (root: -1);
//   final next of an occupied x: its next occupied sibling, else the final
//     next of its (occupied) parent (root: -1); an empty node keeps its
//     pass-1 next;
//   first child of a split node: its first occupied child (an empty split
//     node, possible only with mpl < 1: its pass-1 next).
__device__ __forceinline__ int next_pass1(int x, const int* __restrict__ parents) {
  while (x != 0) {
    const int P = (x - 1) >> 3;
    if (((x - 1) & 7) < 7) return x + 1;
    x = parents[P];
  }
  return -1;
}

__device__ __forceinline__ int next_occupied(int x, const int* __restrict__ parents,
                                             const int* __restrict__ count) {
  while (x != 0) {
    const int P = (x - 1) >> 3;
    const int base = 1 + 8 * P;
    for (int s = ((x - 1) & 7) + 1; s < 8; ++s)
      if (count[base + s] > 0) return base + s;
    x = parents[P];
  }
  return -1;
}

__global__ __launch_bounds__(256) void k_next_links(int nn, const int* __restrict__ parents,
                                                    const int* __restrict__ count,
                                                    int* __restrict__ first_child,
                                                    int* __restrict__ next) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= nn) return;
  if (c != 0) next[c] = count[c] > 0 ? next_occupied(c, parents, count) : next_pass1(c, parents);
  const int base = first_child[c];
  if (base < 0) return;  // a leaf
  if (count[c] == 0) {
    first_child[c] = c == 0 ? -1 : next_pass1(c, parents);
    return;
  }
  for (int j = 0; j < 8; ++j)
    if (count[base + j] > 0) {
      first_child[c] = base + j;
      return;
    }
}

// Sum of a value over four consecutive lanes in a fixed order,
// ((v0 + v1) + (v2 + v3)), valid on the first of the four.
__device__ __forceinline__ double quad_sum(double v) {
  v += __shfl_xor(v, 1);
  v += __shfl_xor(v, 2);
  return v;
}

// Leaf R^3 aggregates, zeroing every node first (:469-511).  Four lanes per
// node: lane k takes the leaf's particles k, k + 4, ...; the lanes' sums are
// combined in a fixed order.
__global__ __launch_bounds__(256) void k_leaf_r3(int nn, const int* __restrict__ first_child,
                                                 const int* __restrict__ count,
                                                 const int* __restrict__ mstart,
                                                 const double* __restrict__ spos,
                                                 const double* __restrict__ srad, double* r3sum,
                                                 double* r3c, double* maxr) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  const int i = g >> 2, k0 = g & 3;
  if (i >= nn) return;  // the four lanes of a node leave together
  double s = 0.0, wx = 0.0, wy = 0.0, wz = 0.0, mr = 0.0;
  if (first_child[i] < 0 && count[i] > 0) {
    const int m0 = mstart[i];
    for (int k = k0; k < count[i]; k += 4) {
      const int p = m0 + k;
      const double r = srad[p];
      const double r3 = r * r * r;
      s += r3;
      wx += spos[3 * (size_t)p] * r3;
      wy += spos[3 * (size_t)p + 1] * r3;
      wz += spos[3 * (size_t)p + 2] * r3;
      if (r > mr) mr = r;
    }
  }
  s = quad_sum(s);
  wx = quad_sum(wx);
  wy = quad_sum(wy);
  wz = quad_sum(wz);
  mr = fmax(mr, __shfl_xor(mr, 1));
  mr = fmax(mr, __shfl_xor(mr, 2));
  if (k0) return;
  r3sum[i] = s;
  r3c[3 * i] = wx;
  r3c[3 * i + 1] = wy;
  r3c[3 * i + 2] = wz;
  maxr[i] = mr;
}

// One parent level of the bottom-up R^3 sweep (:513-542).
__global__ void k_sweep_r3(int p0, int p1, const int* __restrict__ parents,
                           const int* __restrict__ pfc, const int* __restrict__ count,
                           double* r3sum, double* r3c, double* maxr) {
  int p = p0 + blockIdx.x * 256 + threadIdx.x;
  if (p >= p1) return;
  const int parent = parents[p], base = pfc[p];
  double s = 0.0, wx = 0.0, wy = 0.0, wz = 0.0, mr = 0.0;
  for (int j = 0; j < 8; ++j) {
    const int c = base + j;
    if (count[c] > 0) {
      s += r3sum[c];
      wx += r3c[3 * c];
      wy += r3c[3 * c + 1];
      wz += r3c[3 * c + 2];
      if (maxr[c] > mr) mr = maxr[c];
    }
  }
  r3sum[parent] = s;
  r3c[3 * parent] = wx;
  r3c[3 * parent + 1] = wy;
  r3c[3 * parent + 2] = wz;
  maxr[parent] = mr;
}

// Leaf force sums and raw moments P_ab = sum x_a F_b (:556-626, :696-766);
// four lanes per node as in k_leaf_r3.
__global__ __launch_bounds__(256) void k_leaf_forces(int nn, const int* __restrict__ first_child,
                                                     const int* __restrict__ count,
                                                     const int* __restrict__ mstart,
                                                     const int* __restrict__ perm,
                                                     const double* __restrict__ spos,
                                                     const double* __restrict__ F, double* fsum,
                                                     double* fmom) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  const int i = g >> 2, k0 = g & 3;
  if (i >= nn) return;  // the four lanes of a node leave together
  double f[3] = {0.0, 0.0, 0.0}, P[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (first_child[i] < 0 && count[i] > 0) {
    const int m0 = mstart[i];
    for (int k = k0; k < count[i]; k += 4) {
      const int p = m0 + k;
      const int o = perm[p];
      const double Fx = F[3 * (size_t)o], Fy = F[3 * (size_t)o + 1], Fz = F[3 * (size_t)o + 2];
      const double rx = spos[3 * (size_t)p], ry = spos[3 * (size_t)p + 1], rz = spos[3 * (size_t)p + 2];
      f[0] += Fx;
      f[1] += Fy;
      f[2] += Fz;
      P[0] += rx * Fx; P[1] += rx * Fy; P[2] += rx * Fz;
      P[3] += ry * Fx; P[4] += ry * Fy; P[5] += ry * Fz;
      P[6] += rz * Fx; P[7] += rz * Fy; P[8] += rz * Fz;
    }
  }
  for (int a = 0; a < 3; ++a) f[a] = quad_sum(f[a]);
  for (int a = 0; a < 9; ++a) P[a] = quad_sum(P[a]);
  if (k0) return;
  for (int a = 0; a < 3; ++a) fsum[3 * (size_t)i + a] = f[a];
  for (int a = 0; a < 9; ++a) fmom[9 * (size_t)i + a] = P[a];
}

__global__ void k_sweep_forces(int p0, int p1, const int* __restrict__ parents,
                               const int* __restrict__ pfc, const int* __restrict__ count,
                               double* fsum, double* fmom) {
  int p = p0 + blockIdx.x * 256 + threadIdx.x;
  if (p >= p1) return;
  const int parent = parents[p], base = pfc[p];
  double f[3] = {0.0, 0.0, 0.0}, P[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int j = 0; j < 8; ++j) {
    const int c = base + j;
    if (count[c] > 0) {
      for (int a = 0; a < 3; ++a) f[a] += fsum[3 * (size_t)c + a];
      for (int a = 0; a < 9; ++a) P[a] += fmom[9 * (size_t)c + a];
    }
  }
  for (int a = 0; a < 3; ++a) fsum[3 * (size_t)parent + a] = f[a];
  for (int a = 0; a < 9; ++a) fmom[9 * (size_t)parent + a] = P[a];
}

// ---------------------------------------------------------------- scales / packing

__global__ __launch_bounds__(256) void k_bbox(const double* __restrict__ pos, int n,
                                              double* __restrict__ part) {
  __shared__ double lo_s[3][256], hi_s[3][256];
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
    for (int a = 0; a < 3; ++a) {
      const double v = pos[3 * (size_t)i + a];
      lo[a] = fmin(lo[a], v);
      hi[a] = fmax(hi[a], v);
    }
  for (int a = 0; a < 3; ++a) {
    lo_s[a][threadIdx.x] = lo[a];
    hi_s[a][threadIdx.x] = hi[a];
  }
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int a = 0; a < 3; ++a) {
        lo_s[a][threadIdx.x] = fmin(lo_s[a][threadIdx.x], lo_s[a][threadIdx.x + w]);
        hi_s[a][threadIdx.x] = fmax(hi_s[a][threadIdx.x], hi_s[a][threadIdx.x + w]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int a = 0; a < 3; ++a) {
      part[6 * blockIdx.x + a] = lo_s[a][0];
      part[6 * blockIdx.x + 3 + a] = hi_s[a][0];
    }
}

// Scales of the fixed-point / work units (same scheme as the direct kernel).
// Scales of the fixed-point records, and the root node (root_init).
__global__ void k_scales(const double* __restrict__ bbox, int nblk, double L, int periodic,
                         double m_const, Scales* __restrict__ sc, int n, double* bmin,
                         double* bmax, int* first_child, int* next, int* leaf_start, int* count,
                         int* mstart, int* counter, LevelRec* rec) {
  if (threadIdx.x != 0) return;
  root_init(L, n, bmin, bmax, first_child, next, leaf_start, count, mstart, counter, rec);
  Scales s;
  s.periodic = periodic;
  s.pad = 0;
  if (periodic) {
    s.origin[0] = s.origin[1] = s.origin[2] = 0.0;
    s.U = L / kTwo32;
    s.inv_U = kTwo32 / L;
  } else {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int b = 0; b < nblk; ++b)
      for (int a = 0; a < 3; ++a) {
        lo[a] = fmin(lo[a], bbox[6 * b + a]);
        hi[a] = fmax(hi[a], bbox[6 * b + 3 + a]);
      }
    double ext = 0.0;
    for (int a = 0; a < 3; ++a) ext = fmax(ext, hi[a] - lo[a]);
    if (!(ext > 0.0) || !isfinite(ext)) ext = 1.0;
    for (int a = 0; a < 3; ++a) s.origin[a] = isfinite(lo[a]) ? lo[a] : 0.0;
    s.U = ext / (2147483648.0 - 4096.0);
    s.inv_U = 1.0 / s.U;
  }
  s.lambda = s.U / kSigma;
  const double fs = m_const * s.lambda * s.lambda;
  s.Fs = (fs > 0.0 && isfinite(fs)) ? fs : 1.0;
  *sc = s;
}

// Fixed-point coordinate of x along axis a (the sorted records' and the node
// centres' quantization): periodic, the fraction of the box times 2^32;
// otherwise units of U from the bounding-box origin.
__device__ __forceinline__ uint32_t quantize(double x, int a, const Scales& sc) {
#pragma clang fp contract(off)
  if (sc.periodic) {
    double u = x * sc.inv_U * 0x1p-32;
    u -= floor(u);
    return (uint32_t)(unsigned long long)rint(u * kTwo32);
  }
  return (uint32_t)(long long)rint((x - sc.origin[a]) * sc.inv_U);
}

// Sorted particle data: float64 positions / radii and fixed-point records.
__global__ __launch_bounds__(256) void k_sorted_particles(
    const double4* __restrict__ packed, int n, const int* __restrict__ perm,
    const Scales* __restrict__ scp, double* __restrict__ spos, double* __restrict__ srad,
    RecF* __restrict__ srec) {
  int p = blockIdx.x * 256 + threadIdx.x;
  if (p == 0) srec[n] = RecF{};  // the spare record after the last
  if (p >= n) return;
  const Scales sc = *scp;
  const double4 d = packed[perm[p]];
  const double xs[3] = {d.x, d.y, d.z};
  uint32_t q[3];
  for (int a = 0; a < 3; ++a) {
    spos[3 * (size_t)p + a] = xs[a];
    q[a] = quantize(xs[a], a, sc);
  }
  const double r = d.w;
  srad[p] = r;
  const double rl = r / sc.lambda;
  RecF rr;
  rr.x = q[0];
  rr.y = q[1];
  rr.z = q[2];
  rr.r = (float)rl;
  rr.r3 = (float)(rl * rl * rl);
  rr.f[0] = rr.f[1] = rr.f[2] = 0.f;
  srec[p] = rr;
}

// Walk record of every node (centre as get_center, octree_node.py:97-114).
__global__ void k_node_walk(int nn, const double* __restrict__ bmin, const double* __restrict__ bmax,
                            const double* __restrict__ r3sum, const double* __restrict__ r3c,
                            const int* __restrict__ first_child, const int* __restrict__ next,
                            const int* __restrict__ count, const int* __restrict__ mstart,
                            const Scales* __restrict__ scp, NodeW* __restrict__ nw,
                            NodeC* __restrict__ nc) {
#pragma clang fp contract(off)
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nn) return;
  const Scales sc = *scp;
  NodeW w;
  NodeC c;
  const double s3 = r3sum[i];
  if (s3 > 0.0) {
    const double inv = 1.0 / s3;
    c.cx = r3c[3 * i] * inv;
    c.cy = r3c[3 * i + 1] * inv;
    c.cz = r3c[3 * i + 2] * inv;
  } else {
    c.cx = (bmin[3 * i] + bmax[3 * i]) * 0.5;
    c.cy = (bmin[3 * i + 1] + bmax[3 * i + 1]) * 0.5;
    c.cz = (bmin[3 * i + 2] + bmax[3 * i + 2]) * 0.5;
  }
  const double ex = bmax[3 * i] - bmin[3 * i], ey = bmax[3 * i + 1] - bmin[3 * i + 1],
               ez = bmax[3 * i + 2] - bmin[3 * i + 2];
  const double s = fmax(fmax(ex, ey), ez);
  c.s2 = s * s;
  nc[i] = c;
  w.qx = quantize(c.cx, 0, sc);
  w.qy = quantize(c.cy, 1, sc);
  w.qz = quantize(c.cz, 2, sc);
  const double su = s * sc.inv_U;
  w.sf = (float)su;
  w.s2f = (float)(su * su);
  w.first_child = first_child[i];
  w.next = next[i];
  w.count = count[i];
  w.mstart = mstart[i];
  const double lam = scp->lambda;
  w.r3w = (float)(s3 / (lam * lam * lam));
  w.flags = (count[i] == 0 || s3 < kR3Skip) ? kSkipR3 : 0;
  if (!(su >= 0x1p22)) w.flags |= kExactMac;
  for (int k = 0; k < 5; ++k) w.pad[k] = 0;
  nw[i] = w;
}

// Stokeslet node data + the skip flag on the force sum (:1021-1029, :1215-1240).
__global__ void k_node_force(int nn, const double* __restrict__ fsum, const double* __restrict__ fmom,
                             const NodeW* __restrict__ nwr, const NodeC* __restrict__ ncr,
                             const Scales* __restrict__ scp, NodeF* __restrict__ nf,
                             int* __restrict__ flags_out) {
#pragma clang fp contract(off)
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nn) return;
  const Scales sc = *scp;
  const double Fx = fsum[3 * (size_t)i], Fy = fsum[3 * (size_t)i + 1], Fz = fsum[3 * (size_t)i + 2];
  const double fs2 = Fx * Fx + Fy * Fy + Fz * Fz;
  const NodeW w = nwr[i];
  const NodeC c = ncr[i];
  NodeF o;
  o.fx = (float)(Fx / sc.Fs);
  o.fy = (float)(Fy / sc.Fs);
  o.fz = (float)(Fz / sc.Fs);
  const double cm[3] = {c.cx, c.cy, c.cz};
  const double F[3] = {Fx, Fy, Fz};
  const double k = 1.0 / (sc.Fs * sc.U);
  double D[9];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) D[3 * a + b] = (fmom[9 * (size_t)i + 3 * a + b] - cm[a] * F[b]) * k;
  o.tr = (float)(D[0] + D[4] + D[8]);
  o.axy = (float)(D[3] - D[1]);
  o.axz = (float)(D[6] - D[2]);
  o.ayz = (float)(D[7] - D[5]);
  o.sxx = (float)D[0];
  o.syy = (float)D[4];
  o.szz = (float)D[8];
  o.sxy2 = (float)(D[1] + D[3]);
  o.sxz2 = (float)(D[2] + D[6]);
  o.syz2 = (float)(D[5] + D[7]);
  o.pad[0] = o.pad[1] = o.pad[2] = 0.f;
  nf[i] = o;
  int fl = w.flags & ~kSkipF;
  if (w.count == 0 || fs2 < kFsqSkip) fl |= kSkipF;
  flags_out[i] = fl;
}

__global__ void k_apply_flags(int nn, const int* __restrict__ fl, NodeW* nw) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < nn) nw[i].flags = fl[i];
}

// Source forces / Fs into the sorted records.
__global__ void k_sorted_forces(int n, const int* __restrict__ perm, const double* __restrict__ F,
                                const Scales* __restrict__ scp, RecF* __restrict__ srec) {
  int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  const double fs = scp->Fs;
  const int o = perm[p];
  for (int a = 0; a < 3; ++a) srec[p].f[a] = (float)(F[3 * (size_t)o + a] / fs);
}

// ---------------------------------------------------------------- walks

struct CorrArgs {
  dem::CorrGrid g;
  int enabled;
};

struct WalkArgs {
  const NodeW* nw;
  const NodeC* nc;
  const NodeF* nf;
  const RecF* srec;
  const double* spos;
  const double* srad;
  const int* perm;
  int n;
  int guard;  // iteration bound of a walk (2 * node_count + 8)
  double L;
  int periodic;
  double theta_sq;
  float th2f;     // fl32(theta^2)
  float band_s;   // 8 theta (rounded up): the band's term in s
  int exact_mac;  // every opening test on the float64 path
};

// Wave-coherent walks.  The 64 lanes of a wave (targets consecutive in
// Morton order, hence close in space) traverse one node sequence together:
// at an internal node every lane still in the walk applies its own opening
// test; a lane that takes the node's approximation (or prunes it) records
// next[node] as the node where it resumes, and the wave descends when any
// lane opens the node.  A preorder walk reaches next[X] exactly after X's
// subtree, so each lane meets the same nodes, in the same order, as its own
// walk would: the interaction sets and summation order are the reference's
// per-particle ones, while node records and leaf members are loaded at
// wave-uniform addresses (scalar loads) and control flow does not diverge.
constexpr int kActive = -2;  // lane state: in the walk; otherwise the resume node

__device__ __forceinline__ int wave_node(int node) { return __builtin_amdgcn_readfirstlane(node); }

// Walk workgroup: one wave, so the dispatcher places waves one at a time and
// the launch's tail is balanced per wave.
constexpr int kWalkWG = 64;

// Force walk (:783-906).  Writes forces (original order) and the stored copy.
// TPL targets per lane: a wave walks for 64 TPL consecutive (Morton-ordered)
// targets, lane l holding targets base + 64 h + l; node loads, leaf-member
// loads and the walk's control are shared by all of them.
template <int TPL>
__global__ __launch_bounds__(kWalkWG) void k_walk_forces(WalkArgs wa, const Scales* __restrict__ scp,
                                                     double m_const, double* __restrict__ F,
                                                     double* __restrict__ pforces) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63;
  const int base = (blockIdx.x * (kWalkWG / 64) + (threadIdx.x >> 6)) * 64 * TPL;
  const Scales sc = *scp;
  int t[TPL], resume[TPL];
  bool live[TPL];
  RecF me[TPL];
  double p[TPL][3], acc[TPL][3];
  for (int h = 0; h < TPL; ++h) {
    const int t0 = base + 64 * h + lane;
    live[h] = t0 < wa.n;
    t[h] = live[h] ? t0 : wa.n - 1;
    me[h] = wa.srec[t[h]];
    for (int a = 0; a < 3; ++a) {
      p[h][a] = wa.spos[3 * (size_t)t[h] + a];
      acc[h][a] = 0.0;
    }
    resume[h] = live[h] ? kActive : -1;
  }
  int node = 0;
  int steps = wa.guard;  // a preorder walk visits each node at most once
  while (node >= 0 && --steps > 0) {
    bool act[TPL];
    bool any = false;
    for (int h = 0; h < TPL; ++h) {
      if (resume[h] == node) resume[h] = kActive;
      act[h] = resume[h] == kActive;
      any = any || act[h];
    }
    const NodeW nd = wa.nw[node];
    if (nd.flags & kSkipR3) {
      node = wave_node(nd.next);
      continue;
    }
    if (nd.first_child < 0) {
      if (any) {
        float f[TPL][3] = {};
        const int e = nd.mstart + nd.count;
        // members in twos, both records in one scalar load (the record
        // array ends with a spare record); lanes outside the walk compute
        // too and drop their sums below
        for (int j = nd.mstart; j < e; j += 2) {
          const RecF2 two = load_two(wa.srec, j);
          for (int u = 0; u < 2; ++u) {
            const RecF& sj = u == 0 ? two.a : two.b;
            const int jj = j + u;
            // tree leaves keep r^2 >= (r_i + r_j)^2 (:835), which also drops
            // j == i (r^2 = 0); the spare past the leaf adds nothing
            for (int h = 0; h < TPL; ++h)
              dem::dipole_pair<true>(dem::diff32(sj.x, me[h].x), dem::diff32(sj.y, me[h].y),
                                     dem::diff32(sj.z, me[h].z), me[h].r, sj.r, sj.r3, jj < e,
                                     f[h][0], f[h][1], f[h][2]);
          }
        }
        for (int h = 0; h < TPL; ++h)
          if (act[h])
            for (int a = 0; a < 3; ++a) acc[h][a] += (double)f[h][a];
      }
      node = wave_node(nd.next);
      continue;
    }
    bool open = false;
    const bool exact = wa.exact_mac || (nd.flags & kExactMac);
    for (int h = 0; h < TPL; ++h) {
      if (!act[h]) continue;
      // R = centre - x_i (fixed units)
      float Rx = dem::diff32(nd.qx, me[h].x), Ry = dem::diff32(nd.qy, me[h].y),
            Rz = dem::diff32(nd.qz, me[h].z);
      int take = exact ? -1 : fast_open_test(Rx, Ry, Rz, wa.th2f, wa.band_s, nd.s2f, nd.sf);
      if (take < 0) {
        // the reference's float64 test (:871)
        const NodeC c = wa.nc[node];
        double dx = mic64(c.cx - p[h][0], wa.L, wa.periodic);
        double dy = mic64(c.cy - p[h][1], wa.L, wa.periodic);
        double dz = mic64(c.cz - p[h][2], wa.L, wa.periodic);
        double d2 = dx * dx;
        d2 = d2 + dy * dy;
        d2 = d2 + dz * dz;
        take = c.s2 < wa.theta_sq * d2 && d2 > 0.0;
        Rx = (float)(dx * sc.inv_U);
        Ry = (float)(dy * sc.inv_U);
        Rz = (float)(dz * sc.inv_U);
      }
      if (take) {
        float fx = 0.f, fy = 0.f, fz = 0.f;
        dem::dipole_monopole(Rx, Ry, Rz, nd.r3w, fx, fy, fz);
        acc[h][0] += (double)fx;
        acc[h][1] += (double)fy;
        acc[h][2] += (double)fz;
        resume[h] = nd.next;
      } else {
        open = true;
      }
    }
    node = wave_node(__any(open) ? nd.first_child : nd.next);
  }
  for (int h = 0; h < TPL; ++h) {
    if (!live[h]) continue;
    double ax = acc[h][0], ay = acc[h][1], az = acc[h][2];
    if (steps <= 0) ax = ay = az = NAN;  // malformed next chain: fail loudly
    const double rl = wa.srad[t[h]] / sc.lambda;
    const double c = m_const * sc.lambda * sc.lambda * kSigma3 * (rl * rl * rl);
    const int o = wa.perm[t[h]];
    const double Fx = c * ax, Fy = c * ay, Fz = c * az;
    F[3 * (size_t)o] = Fx;
    F[3 * (size_t)o + 1] = Fy;
    F[3 * (size_t)o + 2] = Fz;
    pforces[3 * (size_t)o] = Fx;
    pforces[3 * (size_t)o + 1] = Fy;
    pforces[3 * (size_t)o + 2] = Fz;
  }
}

// Stokeslet / total-velocity walk (:1000-1280, :1326-1588).
// MODE 0: Stokeslet (+ COMSOL self term); MODE 1: total velocity (migration
// F_i / (stokes r_i) added, no self term — as the reference).
template <int MODE, bool CORR>
__global__ __launch_bounds__(kWalkWG) void k_walk_stokes(WalkArgs wa, const Scales* __restrict__ scp,
                                                     CorrArgs ca, double eta_const, double common,
                                                     double self_coeff, double stokes_factor,
                                                     const double* __restrict__ Fself,
                                                     double* __restrict__ V) {
#pragma clang fp contract(off)
  const int t0 = blockIdx.x * kWalkWG + threadIdx.x;
  const bool live = t0 < wa.n;
  const int t = live ? t0 : wa.n - 1;
  const Scales sc = *scp;
  const RecF me = wa.srec[t];
  const float rif = me.r * 0x1p22f;  // radius, fixed units
  const double px = wa.spos[3 * (size_t)t], py = wa.spos[3 * (size_t)t + 1],
               pz = wa.spos[3 * (size_t)t + 2];
  double ax = 0.0, ay = 0.0, az = 0.0;  // Oseen sum, fixed units
  double cx = 0.0, cy = 0.0, cz = 0.0;  // COMSOL sum
  int resume = live ? kActive : -1;
  int node = 0;
  int steps = wa.guard;  // a preorder walk visits each node at most once
  while (node >= 0 && --steps > 0) {
    if (resume == node) resume = kActive;
    const bool act = resume == kActive;
    const NodeW nd = wa.nw[node];
    if (nd.flags & kSkipF) {
      node = wave_node(nd.next);
      continue;
    }
    if (nd.first_child < 0) {
      if (act) {
        float vx = 0.f, vy = 0.f, vz = 0.f, ux = 0.f, uy = 0.f, uz = 0.f;
        const int e = nd.mstart + nd.count;
        for (int j = nd.mstart; j < e; j += 2) {
          // members in twos: one wave-uniform (scalar) load of both records
          // (the array ends with a spare record), then the lane's tests
          const RecF2 two = load_two(wa.srec, j);
          for (int u = 0; u < 2; ++u) {
            const RecF& sj = u == 0 ? two.a : two.b;
            const int jj = j + u;
            const float Fx = sj.f[0], Fy = sj.f[1], Fz = sj.f[2];
            // D = x_j - x_i (fixed units); the Oseen tensor is even in D
            const float dx = dem::diff32(sj.x, me.x);
            const float dy = dem::diff32(sj.y, me.y);
            const float dz = dem::diff32(sj.z, me.z);
            // >= (:1055); j == i, contacts and the spare past the leaf add
            // nothing.  2^22 (r_j + r_i) in one rounding, as 2^22 fl(r_i + r_j).
            const bool keep = dem::oseen_member(dx, dy, dz, Fx, Fy, Fz,
                                                fmaf(sj.r, 0x1p22f, rif), jj < e, vx, vy, vz);
            if (CORR && keep) dem::corr_lookup(ca.g, dx, dy, dz, Fx, Fy, Fz, ux, uy, uz);
          }
        }
        ax += (double)vx;
        ay += (double)vy;
        az += (double)vz;
        if (CORR) {
          cx += (double)ux;
          cy += (double)uy;
          cz += (double)uz;
        }
      }
      node = wave_node(nd.next);
      continue;
    }
    // R = centre - x_i (fixed units)
    float Rx = dem::diff32(nd.qx, me.x), Ry = dem::diff32(nd.qy, me.y),
          Rz = dem::diff32(nd.qz, me.z);
    int take = (wa.exact_mac || (nd.flags & kExactMac))
                   ? -1
                   : fast_open_test(Rx, Ry, Rz, wa.th2f, wa.band_s, nd.s2f, nd.sf);
    if (act && take < 0) {
      // the reference's float64 test, x_i - centre (:1139-1154, :1479)
      const NodeC c = wa.nc[node];
      double dx = mic64(px - c.cx, wa.L, wa.periodic);
      double dy = mic64(py - c.cy, wa.L, wa.periodic);
      double dz = mic64(pz - c.cz, wa.L, wa.periodic);
      double d2 = dx * dx;
      d2 = d2 + dy * dy;
      d2 = d2 + dz * dz;
      take = c.s2 < wa.theta_sq * d2 && d2 > 0.0;
      Rx = (float)(-dx * sc.inv_U);
      Ry = (float)(-dy * sc.inv_U);
      Rz = (float)(-dz * sc.inv_U);
    }
    const bool approx = act && take > 0;
    const NodeF nf = wa.nf[node];  // wave-uniform: loaded for every lane
    if (approx) {
      float vx = 0.f, vy = 0.f, vz = 0.f;
      node_velocity(Rx, Ry, Rz, nf, vx, vy, vz);
      ax += (double)vx;
      ay += (double)vy;
      az += (double)vz;
      if (CORR) {
        float ux2 = 0.f, uy2 = 0.f, uz2 = 0.f;
        dem::corr_lookup(ca.g, Rx, Ry, Rz, nf.fx, nf.fy, nf.fz, ux2, uy2, uz2);
        cx += (double)ux2;
        cy += (double)uy2;
        cz += (double)uz2;
      }
      resume = nd.next;
    }
    node = wave_node(__any(act && !approx) ? nd.first_child : nd.next);
  }
  if (!live) return;
  if (steps <= 0) ax = ay = az = NAN;  // malformed next chain: fail loudly
  const int o = wa.perm[t];
  const double km = eta_const * sc.Fs / sc.U;
  const double kc = sc.Fs * common;
  double v[3] = {km * ax, km * ay, km * az};
  if (CORR) {
    v[0] += kc * cx;
    v[1] += kc * cy;
    v[2] += kc * cz;
  }
  if (MODE == 0 && CORR) {
    for (int a = 0; a < 3; ++a) v[a] += self_coeff * Fself[3 * (size_t)o + a];
  }
  if (MODE == 1) {
    const double inv = 1.0 / (stokes_factor * wa.srad[t]);
    for (int a = 0; a < 3; ++a) v[a] = Fself[3 * (size_t)o + a] * inv + v[a];
  }
  for (int a = 0; a < 3; ++a) V[3 * (size_t)o + a] = v[a];
}

// Collision counts between volume classes (TreeCollisionFrequency.
// _count_collisions, pbm/collision_frequency/tree_collision_freq.py:62-140).
// One target per lane in sorted (Morton) order, so a wave's walks are close;
// float64 throughout, contraction off, as the reference writes the tests.
// Counts are integers: the atomic adds commute, the result is exact.
__device__ __forceinline__ double box_gap_sq(const double* lo, const double* hi, const double* p,
                                             double L, int periodic) {
#pragma clang fp contract(off)
  double g[3];
  for (int a = 0; a < 3; ++a) {
    double d = 0.0;
    if (p[a] < lo[a])
      d = lo[a] - p[a];
    else if (p[a] > hi[a])
      d = p[a] - hi[a];
    if (periodic && L > 0.0 && d > L * 0.5) d = L - d;  // :167-173
    g[a] = d;
  }
  return g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
}

__global__ __launch_bounds__(256) void k_walk_collisions(
    int n, double L, int periodic, const int* __restrict__ count,
    const int* __restrict__ first_child, const int* __restrict__ next,
    const int* __restrict__ leaf_start, const int* __restrict__ mstart,
    const int* __restrict__ leaf_indices, const int* __restrict__ perm,
    const double* __restrict__ bmin, const double* __restrict__ bmax,
    const double* __restrict__ maxr, const double* __restrict__ spos,
    const double* __restrict__ srad, const int* __restrict__ bins, int n_bins,
    int* __restrict__ counts) {
#pragma clang fp contract(off)
  const int p0 = blockIdx.x * 256 + threadIdx.x;
  const bool live = p0 < n;
  const int p = live ? p0 : n - 1;
  const int i = perm[p];
  const double pi[3] = {spos[3 * (size_t)p], spos[3 * (size_t)p + 1], spos[3 * (size_t)p + 2]};
  const double ri = srad[p];
  const int bi = bins[i];
  const bool bi_ok = bi >= 0 && bi < n_bins;
  // wave-coherent walk (see k_walk_forces): lanes that prune a node resume at
  // its next; the wave descends when any lane does
  int resume = live ? kActive : -1;
  int node = 0;
  while (node >= 0) {
    if (resume == node) resume = kActive;
    const bool act = resume == kActive;
    const int cnt = count[node];
    if (cnt == 0) {
      node = wave_node(next[node]);
      continue;
    }
    const int fc = first_child[node];
    if (fc < 0) {
      const int ls = leaf_start[node], m0 = mstart[node];
      for (int k = 0; act && k < cnt; ++k) {
        const int j = leaf_indices[ls + k];
        if (j >= 0 && j > i) {
          const double* q = spos + 3 * (size_t)(m0 + k);
          const double dx = mic64(q[0] - pi[0], L, periodic);
          const double dy = mic64(q[1] - pi[1], L, periodic);
          const double dz = mic64(q[2] - pi[2], L, periodic);
          const double d2 = dx * dx + dy * dy + dz * dz;
          const double c = ri + srad[m0 + k];
          if (d2 <= c * c) {
            const int bj = bins[j];
            if (bi_ok && bj >= 0 && bj < n_bins) {
              atomicAdd(&counts[bi * n_bins + bj], 1);
              atomicAdd(&counts[bj * n_bins + bi], 1);
            }
          }
        }
      }
      node = wave_node(next[node]);
      continue;
    }
    const double mc = ri + maxr[node];
    bool descend = false;
    if (act) {
      if (box_gap_sq(bmin + 3 * (size_t)node, bmax + 3 * (size_t)node, pi, L, periodic) > mc * mc)
        resume = next[node];
      else
        descend = true;
    }
    node = wave_node(__any(descend) ? fc : next[node]);
  }
}

// ---------------------------------------------------------------- host helpers

inline int nb256(long long n) { return (int)((n + 255) / 256); }
// workgroups of 256 lanes walking `tpl` targets each
constexpr int kWalkTpl = 1;  // 2 measured slower: the wider union of the walks costs more than it hides

int check_info(const dem_tree_info* info, const void* ws, size_t ws_bytes, TreeLayout& l) {
  if (!info) return dem::set_error(DEM_ERR_INVALID, "tree info is NULL");
  l = carve(const_cast<void*>(ws), info->n_capacity, info->max_nodes);
  if (!ws || ws_bytes < l.bytes)
    return dem::set_error(DEM_ERR_WORKSPACE, "octree workspace %zu < %zu bytes", ws_bytes, l.bytes);
  if (info->node_count < 1) return dem::set_error(DEM_ERR_INVALID, "octree has not been built");
  return DEM_OK;
}

int propagate_forces(const dem_tree_info* info, const TreeLayout& l, const double* F,
                     hipStream_t st) {
  const int nn = info->node_count;
  hipLaunchKernelGGL(k_leaf_forces, dim3(nb256(4 * nn)), dim3(256), 0, st, nn,
                     (const int*)l.first_child, (const int*)l.count, (const int*)l.mstart,
                     (const int*)l.perm, (const double*)l.spos, F, l.fsum, l.fmom);
  DEM_CHECK_LAUNCH("k_leaf_forces");
  for (int lev = info->num_levels - 1; lev >= 0; --lev) {
    const int p0 = info->level_offsets[lev];
    const int p1 = lev + 1 < info->num_levels ? info->level_offsets[lev + 1] : info->parent_count;
    if (p1 > p0) {
      hipLaunchKernelGGL(k_sweep_forces, dim3(nb256(p1 - p0)), dim3(256), 0, st, p0, p1,
                         (const int*)l.parents, (const int*)l.pfc, (const int*)l.count, l.fsum,
                         l.fmom);
      DEM_CHECK_LAUNCH("k_sweep_forces");
    }
  }
  hipLaunchKernelGGL(k_node_force, dim3(nb256(nn)), dim3(256), 0, st, nn, (const double*)l.fsum,
                     (const double*)l.fmom, (const NodeW*)l.nw, (const NodeC*)l.nc,
                     (const Scales*)l.sc, l.nf, l.flag);
  DEM_CHECK_LAUNCH("k_node_force");
  hipLaunchKernelGGL(k_apply_flags, dim3(nb256(nn)), dim3(256), 0, st, nn, (const int*)l.flag,
                     l.nw);
  DEM_CHECK_LAUNCH("k_apply_flags");
  hipLaunchKernelGGL(k_sorted_forces, dim3(nb256(info->n)), dim3(256), 0, st, info->n,
                     (const int*)l.perm, F, (const Scales*)l.sc, l.srec);
  DEM_CHECK_LAUNCH("k_sorted_forces");
  return DEM_OK;
}

WalkArgs walk_args(const dem_tree_info* info, const TreeLayout& l, double theta_sq) {
  WalkArgs w{};
  w.nw = l.nw;
  w.nc = l.nc;
  w.nf = l.nf;
  w.srec = l.srec;
  w.spos = l.spos;
  w.srad = l.srad;
  w.perm = l.perm;
  w.n = info->n;
  w.guard = 2 * info->node_count + 8;
  w.L = info->L;
  w.periodic = info->periodic;
  w.theta_sq = theta_sq;
  w.th2f = (float)theta_sq;
  w.band_s = (float)(8.0 * sqrt(theta_sq) * (1.0 + 0x1p-20));
  w.exact_mac = g_exact_mac;
  return w;
}

int corr_args(const dem_correction* corr, const dem_tree_info* info, CorrArgs& ca,
              double& common, double& self_coeff) {
  ca = CorrArgs{};
  common = 0.0;
  self_coeff = 0.0;
  if (!corr || !corr->enabled) return DEM_OK;
  if (!corr->grid || corr->n < 2)
    return dem::set_error(DEM_ERR_INVALID, "correction enabled without a float grid");
  ca.g.grid = reinterpret_cast<const float4*>(corr->grid);
  ca.g.n = corr->n;
  ca.enabled = 1;
  ca.g.nmax = (float)(corr->n - 1);
  ca.g.ga = 0.f;  // set from the device-side fixed unit below
  ca.g.gb = (float)(-corr->grid_min * corr->inv_dx);
  common = corr->Fz_inv * corr->eta_ratio * corr->L_ratio;
  self_coeff = corr->self_coeff;
  (void)info;
  return DEM_OK;
}

__global__ void k_corr_scale(const Scales* __restrict__ scp, double L_ratio, double inv_dx,
                             float* out) {
  if (threadIdx.x == 0) out[0] = (float)(-scp->U * L_ratio * inv_dx);
}

}  // namespace

extern "C" {

int dem_octree_set_exact_mac(int on) {
  const int prev = g_exact_mac;
  g_exact_mac = on ? 1 : 0;
  return prev;
}

size_t dem_octree_workspace_bytes(int n_capacity, int max_nodes) {
  return carve(nullptr, n_capacity < 0 ? 0 : n_capacity, max_nodes < 1 ? 1 : max_nodes).bytes;
}

int dem_octree_build(const double* pos, const double* radii, int n, double L, int periodic,
                     int mpl, double m_const, dem_tree_info* info, void* ws, size_t ws_bytes,
                     dem_stream_t stream) {
  if (!info) return dem::set_error(DEM_ERR_INVALID, "tree info is NULL");
  if (n < 0) return dem::set_error(DEM_ERR_INVALID, "n must be >= 0");
  if (n > info->n_capacity)
    return dem::set_error(DEM_ERR_CAPACITY,
                          "Number of particles (%d) exceeds allocated capacity (%d)", n,
                          info->n_capacity);
  if (info->max_nodes < 1) return dem::set_error(DEM_ERR_INVALID, "max_nodes must be >= 1");
  if (n > 0 && (!pos || !radii)) return dem::set_error(DEM_ERR_INVALID, "null positions/radii");
  if (!(L > 0.0)) return dem::set_error(DEM_ERR_INVALID, "L must be > 0 (got %g)", L);
  TreeLayout l = carve(ws, info->n_capacity, info->max_nodes);
  if (!ws || ws_bytes < l.bytes)
    return dem::set_error(DEM_ERR_WORKSPACE, "octree workspace %zu < %zu bytes", ws_bytes, l.bytes);
  hipStream_t st = dem::as_stream(stream);
  const int M = info->max_nodes;
  info->n = n;
  info->L = L;
  info->periodic = periodic ? 1 : 0;
  info->mpl = mpl;
  info->node_count = 1;
  info->parent_count = 0;
  info->num_levels = 0;
  info->leaf_slot_count = 0;
  info->max_depth = 0;

  // scales + keys + sort
  int nblk = n > 0 ? nb256(n) : 1;
  if (nblk > 256) nblk = 256;
  if (!periodic && n > 0) {
    hipLaunchKernelGGL(k_bbox, dim3(nblk), dim3(256), 0, st, pos, n, l.bbox);
    DEM_CHECK_LAUNCH("k_bbox");
  }
  hipLaunchKernelGGL(k_scales, dim3(1), dim3(64), 0, st, (const double*)l.bbox, nblk, L,
                     periodic ? 1 : 0, m_const, l.sc, n, l.bmin, l.bmax, l.first_child, l.next,
                     l.leaf_start, l.count, l.mstart, l.counter, l.lrec);
  DEM_CHECK_LAUNCH("k_scales");
  // Keys sorted first by their top kPartialSortBits bits: that orders every
  // node of levels <= kPartialLevels as the full key does (a node's range,
  // and its children's by the next octant digit).  Only a tree that splits
  // deeper takes the remaining passes (a full stable sort of the partially
  // sorted keys, which keeps those ranges).
  constexpr int kPartialSortBits = 32;  // four 8-bit digits: bits 32..63
  // the key has 3 kKeyLevels = 63 bits, so bits 32..62 hold levels 1..10
  constexpr int kPartialLevels = (3 * kKeyLevels - (64 - kPartialSortBits)) / 3;
  static_assert(kPartialLevels == 10, "levels ordered by the first sort");
  bool full_sort = false;
  unsigned long long* skeys = l.keys;  // the buffer holding the sorted keys
  if (n > 0) {
    hipLaunchKernelGGL(k_keys, dim3(nb256(n)), dim3(256), 0, st, pos, radii, n, L, l.keys, l.perm,
                       l.packed);
    DEM_CHECK_LAUNCH("k_keys");
    int rc = dem::radix_sort_u64(l.keys, l.perm, l.keys_alt, l.perm_alt, n,
                                 64 - kPartialSortBits, 64, l.sort_ws, st, &skeys);
    if (rc) return rc;
  }

  // level-by-level split (flat_tree.py:209-236).  Each level is two launches
  // — candidate scan, split — working from a device record of the loop state
  // (two slots: level `level` reads slot level & 1, its split writes the
  // other); the host launches them for an upper bound of the level's size
  // (8 x the previous bound) and reads the record only every kLevelsPerSync
  // levels to stop (a tree of N = 2^20 with leaves of 16 is 7 levels deep).
  constexpr int kLevelsPerSync = 8;
  int* tile_sums = l.flag;  // the flags' storage holds the scan-tile sums
  LevelRec rec{};
  int launched = 0;         // levels launched
  bool rec_read = false;    // rec holds the record the launched levels left
  long long bound = 1;      // nodes the level can hold at most
  for (int level = 0; level < kMaxLevels; ++level) {
    if (level > 0 && (level % kLevelsPerSync == 0 || level == kPartialLevels)) {
      DEM_CHECK_HIP(hipMemcpyAsync(&rec, l.lrec + (level & 1), sizeof(LevelRec),
                                   hipMemcpyDeviceToHost, st));
      DEM_CHECK_HIP(hipStreamSynchronize(st));
      rec_read = true;
      if (rec.stop) break;
    }
    rec_read = false;
    if (level == kPartialLevels && !full_sort && n > 1) {
      // the next split reads octant digits below the sorted bits
      int rc = dem::radix_sort_u64(skeys, l.perm, skeys == l.keys ? l.keys_alt : l.keys,
                                   l.perm_alt, n, 0, 64, l.sort_ws, st, &skeys);
      if (rc) return rc;
      full_sort = true;
    }
    const int cb = (int)(bound < M ? bound : M);
    const int nt = dem::scan_tiles(cb);
    const int cur = level & 1, nxt = cur ^ 1;
    hipLaunchKernelGGL(k_level_scan, dim3(nt), dim3(dem::kScanThreads), 0, st, mpl, level,
                       (const LevelRec*)(l.lrec + cur), (const int*)l.count, l.rank, tile_sums,
                       l.counter + 2 + cur);
    DEM_CHECK_LAUNCH("k_level_scan");
    hipLaunchKernelGGL(k_level_split, dim3((cb + 31) / 32), dim3(256), nt * sizeof(int), st, level,
                       mpl, M, (const LevelRec*)(l.lrec + cur), l.lrec + nxt,
                       (const int*)(l.counter + 2 + cur), l.counter + 2 + nxt,
                       (const int*)l.rank, (const int*)tile_sums,
                       (const unsigned long long*)skeys, l.bmin, l.bmax, l.first_child, l.next,
                       l.leaf_start, l.count, l.mstart, l.parents, l.pfc);
    DEM_CHECK_LAUNCH("k_level_split");
    launched = level + 1;
    bound = bound * 8 < M ? bound * 8 : M;
  }
  if (!rec_read) {
    DEM_CHECK_HIP(hipMemcpyAsync(&rec, l.lrec + (launched & 1), sizeof(LevelRec),
                                 hipMemcpyDeviceToHost, st));
    DEM_CHECK_HIP(hipStreamSynchronize(st));
  }
  info->node_count = rec.node_count;
  info->parent_count = rec.parent_count;
  info->num_levels = rec.num_levels;
  info->max_depth = rec.max_depth;
  for (int k = 0; k < rec.num_levels; ++k) info->level_offsets[k] = rec.level_offsets[k];
  const int nn = info->node_count;
  if (n > 0) {
    hipLaunchKernelGGL(k_sorted_particles, dim3(nb256(n)), dim3(256), 0, st,
                       (const double4*)l.packed, n, (const int*)l.perm, (const Scales*)l.sc,
                       l.spos, l.srad, l.srec);
    DEM_CHECK_LAUNCH("k_sorted_particles");
  }

  // leaf slots (flat_tree.py:370-397)
  const int ntl = dem::scan_tiles(nn);
  hipLaunchKernelGGL(k_leaf_scan, dim3(ntl), dim3(dem::kScanThreads), 0, st, nn,
                     (const int*)l.first_child, (const int*)l.count, l.rank, tile_sums, l.counter);
  DEM_CHECK_LAUNCH("k_leaf_scan");
  hipLaunchKernelGGL(k_leaf_fill, dim3(nb256(nn)), dim3(256), ntl * sizeof(int), st, nn,
                     (const int*)l.first_child, (const int*)l.count, (const int*)l.mstart,
                     (const int*)l.rank, (const int*)tile_sums, ntl, (const int*)l.perm,
                     l.leaf_start, l.leaf_indices);
  DEM_CHECK_LAUNCH("k_leaf_fill");

  // next pointers and first children (flat_tree.py:399-463)
  hipLaunchKernelGGL(k_next_links, dim3(nb256(nn)), dim3(256), 0, st, nn, (const int*)l.parents,
                     (const int*)l.count, l.first_child, l.next);
  DEM_CHECK_LAUNCH("k_next_links");

  // R^3 aggregates (flat_tree.py:544-554)
  hipLaunchKernelGGL(k_leaf_r3, dim3(nb256(4 * nn)), dim3(256), 0, st, nn, (const int*)l.first_child,
                     (const int*)l.count, (const int*)l.mstart, (const double*)l.spos,
                     (const double*)l.srad, l.r3sum, l.r3c, l.maxr);
  DEM_CHECK_LAUNCH("k_leaf_r3");
  for (int lev = info->num_levels - 1; lev >= 0; --lev) {
    const int p0 = info->level_offsets[lev];
    const int p1 = lev + 1 < info->num_levels ? info->level_offsets[lev + 1] : info->parent_count;
    if (p1 > p0) {
      hipLaunchKernelGGL(k_sweep_r3, dim3(nb256(p1 - p0)), dim3(256), 0, st, p0, p1,
                         (const int*)l.parents, (const int*)l.pfc, (const int*)l.count, l.r3sum,
                         l.r3c, l.maxr);
      DEM_CHECK_LAUNCH("k_sweep_r3");
    }
  }
  hipLaunchKernelGGL(k_node_walk, dim3(nb256(nn)), dim3(256), 0, st, nn, (const double*)l.bmin,
                     (const double*)l.bmax, (const double*)l.r3sum, (const double*)l.r3c,
                     (const int*)l.first_child, (const int*)l.next, (const int*)l.count,
                     (const int*)l.mstart, (const Scales*)l.sc, l.nw, l.nc);
  DEM_CHECK_LAUNCH("k_node_walk");
  int lsc = 0;
  DEM_CHECK_HIP(hipMemcpyAsync(&lsc, l.counter + 1, sizeof(int), hipMemcpyDeviceToHost, st));
  DEM_CHECK_HIP(hipStreamSynchronize(st));
  info->leaf_slot_count = lsc;
  info->forces_valid = 0;
  return DEM_OK;
}

int dem_octree_forces(dem_tree_info* info, double theta_sq, double m_const, double* forces,
                      void* ws, size_t ws_bytes, dem_stream_t stream) {
  TreeLayout l;
  int rc = check_info(info, ws, ws_bytes, l);
  if (rc) return rc;
  if (info->n == 0) return DEM_OK;
  if (!forces) return dem::set_error(DEM_ERR_INVALID, "null forces");
  hipStream_t st = dem::as_stream(stream);
  WalkArgs wa = walk_args(info, l, theta_sq);
  DEM_PROF_LAUNCH("k_walk_forces", st,
                  hipLaunchKernelGGL(k_walk_forces<kWalkTpl>,
                                     dim3((unsigned)((info->n + (long long)kWalkWG * kWalkTpl - 1) /
                                                     ((long long)kWalkWG * kWalkTpl))),
                                     dim3(kWalkWG), 0, st, wa,
                                     (const Scales*)l.sc, m_const, forces, l.pforces));
  DEM_CHECK_LAUNCH("k_walk_forces");
  info->forces_valid = 1;
  return DEM_OK;
}

static int stokes_common(int mode, const dem_tree_info* info, double theta_sq, const double* forces,
                         double eta_const, double stokes_factor, const dem_correction* corr,
                         double* vel, void* ws, size_t ws_bytes, dem_stream_t stream) {
  TreeLayout l;
  int rc = check_info(info, ws, ws_bytes, l);
  if (rc) return rc;
  if (info->n == 0) return DEM_OK;
  if (!vel) return dem::set_error(DEM_ERR_INVALID, "null velocities");
  hipStream_t st = dem::as_stream(stream);
  const double* F = forces ? forces : (const double*)l.pforces;
  if (!forces && !info->forces_valid)
    return dem::set_error(DEM_ERR_INVALID, "no forces: pass them or run dem_octree_forces first");
  if ((rc = propagate_forces(info, l, F, st))) return rc;
  CorrArgs ca;
  double common, self_coeff;
  if ((rc = corr_args(corr, info, ca, common, self_coeff))) return rc;
  if (ca.enabled) {
    // ga = -U * L_ratio * inv_dx needs U, which lives on the device for open boxes
    float* tmp = reinterpret_cast<float*>(l.counter + 8);
    hipLaunchKernelGGL(k_corr_scale, dim3(1), dim3(64), 0, st, (const Scales*)l.sc, corr->L_ratio,
                       corr->inv_dx, tmp);
    DEM_CHECK_LAUNCH("k_corr_scale");
    float ga = 0.f;
    DEM_CHECK_HIP(hipMemcpyAsync(&ga, tmp, sizeof(float), hipMemcpyDeviceToHost, st));
    DEM_CHECK_HIP(hipStreamSynchronize(st));
    ca.g.ga = ga;
  }
  WalkArgs wa = walk_args(info, l, theta_sq);
  const dim3 g((unsigned)((info->n + kWalkWG - 1) / kWalkWG)), b(kWalkWG);
  const char* name = mode == 0 ? "k_walk_stokeslet" : "k_walk_total";
  int slot = dem::prof_begin(name, st);
  if (mode == 0 && ca.enabled)
    hipLaunchKernelGGL((k_walk_stokes<0, true>), g, b, 0, st, wa, (const Scales*)l.sc, ca,
                       eta_const, common, self_coeff, stokes_factor, F, vel);
  else if (mode == 0)
    hipLaunchKernelGGL((k_walk_stokes<0, false>), g, b, 0, st, wa, (const Scales*)l.sc, ca,
                       eta_const, common, self_coeff, stokes_factor, F, vel);
  else if (ca.enabled)
    hipLaunchKernelGGL((k_walk_stokes<1, true>), g, b, 0, st, wa, (const Scales*)l.sc, ca,
                       eta_const, common, self_coeff, stokes_factor, F, vel);
  else
    hipLaunchKernelGGL((k_walk_stokes<1, false>), g, b, 0, st, wa, (const Scales*)l.sc, ca,
                       eta_const, common, self_coeff, stokes_factor, F, vel);
  dem::prof_end(slot, st);
  DEM_CHECK_LAUNCH(name);
  return DEM_OK;
}

int dem_octree_stokeslet(const dem_tree_info* info, double theta_sq, const double* forces,
                         double eta_const, const dem_correction* corr, double* vel, void* ws,
                         size_t ws_bytes, dem_stream_t stream) {
  return stokes_common(0, info, theta_sq, forces, eta_const, 0.0, corr, vel, ws, ws_bytes, stream);
}

int dem_octree_total_velocity(const dem_tree_info* info, double theta_sq, const double* forces,
                              double eta_const, double stokes_factor, const dem_correction* corr,
                              double* vel, void* ws, size_t ws_bytes, dem_stream_t stream) {
  return stokes_common(1, info, theta_sq, forces, eta_const, stokes_factor, corr, vel, ws, ws_bytes,
                       stream);
}

int dem_octree_collision_counts(const dem_tree_info* info, const int32_t* bins, int n_bins,
                                int32_t* counts, void* ws, size_t ws_bytes, dem_stream_t stream) {
  TreeLayout l;
  int rc = check_info(info, ws, ws_bytes, l);
  if (rc) return rc;
  if (n_bins < 1 || (long long)n_bins * n_bins > (1ll << 30))
    return dem::set_error(DEM_ERR_INVALID, "n_bins %d out of range", n_bins);
  if (!counts) return dem::set_error(DEM_ERR_INVALID, "null counts");
  if (info->n > 0 && !bins) return dem::set_error(DEM_ERR_INVALID, "null bins");
  hipStream_t st = dem::as_stream(stream);
  const int nb2 = n_bins * n_bins;
  hipLaunchKernelGGL(k_fill_i, dim3(nb256(nb2)), dim3(256), 0, st, counts, nb2, 0);
  DEM_CHECK_LAUNCH("k_fill_i");
  if (info->n == 0) return DEM_OK;
  DEM_PROF_LAUNCH("k_walk_collisions", st,
                  hipLaunchKernelGGL(k_walk_collisions, dim3(nb256(info->n)), dim3(256), 0, st,
                                     info->n, info->L, info->periodic, (const int*)l.count,
                                     (const int*)l.first_child, (const int*)l.next,
                                     (const int*)l.leaf_start, (const int*)l.mstart,
                                     (const int*)l.leaf_indices, (const int*)l.perm,
                                     (const double*)l.bmin, (const double*)l.bmax,
                                     (const double*)l.maxr, (const double*)l.spos,
                                     (const double*)l.srad, bins, n_bins, counts));
  DEM_CHECK_LAUNCH("k_walk_collisions");
  return DEM_OK;
}

int dem_octree_field(const dem_tree_info* info, const char* name, size_t* offset, size_t* bytes) {
  if (!info || !name || !offset || !bytes)
    return dem::set_error(DEM_ERR_INVALID, "dem_octree_field: null argument");
  TreeLayout l = carve(nullptr, info->n_capacity, info->max_nodes);
  const size_t M = (size_t)info->max_nodes, N = (size_t)(info->n_capacity > 0 ? info->n_capacity : 1);
  struct F {
    const char* n;
    const void* p;
    size_t b;
  } table[] = {
      {"min", l.bmin, 3 * M * 8},          {"max", l.bmax, 3 * M * 8},
      {"R3_sum", l.r3sum, M * 8},          {"R3_c", l.r3c, 3 * M * 8},
      {"force_sum", l.fsum, 3 * M * 8},    {"fmom", l.fmom, 9 * M * 8},
      {"max_radius", l.maxr, M * 8},       {"first_child", l.first_child, M * 4},
      {"next", l.next, M * 4},             {"leaf_start", l.leaf_start, M * 4},
      {"count", l.count, M * 4},           {"parents", l.parents, M * 4},
      {"parent_first_child", l.pfc, M * 4}, {"leaf_indices", l.leaf_indices, N * 4},
      {"perm", l.perm, N * 4},             {"particle_forces", l.pforces, 3 * N * 8},
  };
  for (const F& f : table)
    if (strcmp(f.n, name) == 0) {
      *offset = (size_t)reinterpret_cast<uintptr_t>(f.p);
      *bytes = f.b;
      return DEM_OK;
    }
  return dem::set_error(DEM_ERR_INVALID, "unknown octree field '%s'", name);
}

}  // extern "C"
