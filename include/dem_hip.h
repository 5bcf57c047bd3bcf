/*
 * dem_hip.h — C ABI of libdemhip.so, the MI355X (gfx950) implementation of the
 * DEM droplet hot path: dipole force + Oseen convection (direct O(N^2)),
 * spatial-hash contact detection, the explicit Euler step and the flat
 * Barnes-Hut octree.
 *
 * Conventions (all entry points):
 *   - Buffers are device pointers owned by the caller (PyTorch ROCm tensors in
 *     the Python host layer).  The library never frees caller memory and never
 *     allocates device memory: every scratch buffer is a caller-provided
 *     workspace whose size is returned by the matching *_workspace_bytes().
 *   - Positions are (N,3) row-major float64 in metres, radii (N,) float64,
 *     forces (N,3) float64 in newtons, velocities (N,3) float64 in m/s — the
 *     same arrays the reference passes between its Python operators.
 *   - Every call is asynchronous on `stream` (a hipStream_t; NULL = default
 *     stream) and returns DEM_OK or a negative status; dem_last_error() gives a
 *     thread-local message for the last failure on the calling thread.
 *   - There is no global mutable state: calls with distinct workspaces may run
 *     concurrently on distinct streams.
 */
#ifndef DEM_HIP_H
#define DEM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* dem_stream_t; /* hipStream_t */

enum dem_status {
  DEM_OK = 0,
  DEM_ERR_INVALID = -1,   /* bad argument (n < 0, null pointer, L <= 0 ...)        */
  DEM_ERR_CAPACITY = -2,  /* output / workspace capacity exceeded                    */
  DEM_ERR_HIP = -3,       /* a HIP runtime call failed                               */
  DEM_ERR_WORKSPACE = -4  /* workspace smaller than *_workspace_bytes() demands       */
};

/* Physical constants of one force evaluation.
 * Reference: dem/force_calculator/direct_droplet_force_calculator.py:23-54
 * (m_const = 12*pi*eps0*eps_oil*E^2, eta_const = 1/(8*pi*eta_oil), L,
 *  boundary_mode_int) and dem/octree/force_tree.py:54-57. */
typedef struct dem_phys {
  double m_const;   /* dipole force constant (N/m^2)                 */
  double eta_const; /* Oseen prefactor 1/(8 pi eta_oil) (1/(Pa s))    */
  double L;         /* periodic box edge (m)                           */
  int32_t periodic; /* 1: minimum-image convention, 0: open boundary   */
  int32_t reserved;
} dem_phys;

/* COMSOL periodic-image correction of the Oseen sum (trilinear lookups on a
 * regular n^3 grid over [grid_min, -grid_min]^3).
 * Reference: direct_droplet_force_calculator.py:74-181, :413-464 and
 * dem/octree/flat_tree.py:141-167.  `grid` is float32 (u,v,w,0) per node,
 * node (ix,iy,iz) at ((ix*n)+iy)*n+iz; `grid64` the same in float64 (used only
 * by the FP64 oracle kernel; may be NULL otherwise). */
typedef struct dem_correction {
  const float* grid;
  const double* grid64;
  int32_t n;
  int32_t enabled;
  double grid_min;
  double inv_dx;
  double Fz_inv;
  double L_ratio;   /* L_comsol / L_sim                                   */
  double eta_ratio; /* eta_comsol / eta_oil                               */
  double self_coeff;/* w(0,0,0)/Fz * eta_ratio * L_ratio                  */
} dem_correction;

const char* dem_last_error(void);
int dem_version(void);
/* Number of HIP devices visible to the library's runtime (or a negative
 * status).  Used by the host layer to fail loudly without a GPU. */
int dem_device_count(void);

/* ------------------------------------------------------------------------ */
/* Direct O(N^2) dipole force + Oseen convection (FP32 compute, FP64 I/O)    */
/* ------------------------------------------------------------------------ */

size_t dem_direct_workspace_bytes(int n);

/* Forces only.  Replaces DirectDropletForceCalculator._calculate_forces
 * (direct_droplet_force_calculator.py:202-253) behind calculate() (:260-305). */
int dem_direct_forces(const double* pos, const double* radii, int n,
                      const dem_phys* phys, double* forces,
                      void* ws, size_t ws_bytes, dem_stream_t stream);

/* Convection from caller-supplied forces.  Replaces
 * _calculate_convection_velocities (:354-468) behind calculate_convection
 * (:307-352).  `corr` may be NULL (no periodic correction). */
int dem_direct_convection(const double* pos, const double* radii,
                          const double* forces, int n, const dem_phys* phys,
                          const dem_correction* corr, double* velocities,
                          void* ws, size_t ws_bytes, dem_stream_t stream);

/* Fused force + convection.  Replaces calculate_forces_and_convection
 * (:470-527): phase 1 (forces), then phase 2 (Oseen sum of those forces). */
int dem_direct_forces_and_convection(const double* pos, const double* radii,
                                     int n, const dem_phys* phys,
                                     const dem_correction* corr,
                                     double* forces, double* velocities,
                                     void* ws, size_t ws_bytes,
                                     dem_stream_t stream);

/* FP64 oracle twin of dem_direct_forces_and_convection: the reference's
 * arithmetic in float64 (same branches, masks and minimum-image rule), one
 * target per thread.  Used as the checker at sizes the CPU oracle cannot
 * reach in seconds.  No workspace. */
int dem_direct_forces_and_convection_f64(const double* pos, const double* radii,
                                         int n, const dem_phys* phys,
                                         const dem_correction* corr,
                                         double* forces, double* velocities,
                                         dem_stream_t stream);

/* FP64 oracle restricted to the targets listed in targets[0..nt): forces of
 * those targets (rows of F, nt x 3) over all n sources and, when V is not
 * NULL, their Oseen velocities (nt x 3) from the given forces of all n
 * sources.  Checker for sizes where the full FP64 sum is too slow. */
int dem_direct_f64_targets(const double* pos, const double* radii, int n,
                           const dem_phys* phys, const dem_correction* corr,
                           const int32_t* targets, int nt,
                           const double* sources_forces, double* F, double* V,
                           dem_stream_t stream);

/* Collision counts between volume classes over all pairs
 * (DirectCollisionFrequency.compute,
 * pbm/collision_frequency/direct_collision_freq.py:15-41): every pair i < j with
 * sqrt(dx^2 + dy^2 + dz^2) <= r_i + r_j, the minimum image taken as
 * d_a -= L rint(d_a / L) when periodic and L > 0 (float64), adds 1 to
 * counts[bins[i]][bins[j]] and counts[bins[j]][bins[i]].  counts: n_bins x
 * n_bins int32, overwritten.  A bin outside [0, n_bins) contributes nothing. */
int dem_direct_collision_counts(const double* pos, const double* radii, int n, double L,
                                int periodic, const int32_t* bins, int n_bins, int32_t* counts,
                                dem_stream_t stream);

/* Tuning knob of the direct kernels (j-split factor and FP32 tile length);
 * 0 keeps the built-in heuristic.  Returns the previous values packed as
 * (split << 16) | tile. */
int dem_direct_set_tuning(int split, int tile);

/* Pair-once (Newton's third law) evaluation of the periodic force sum and of
 * the Oseen sum without the COMSOL correction: each unordered pair is
 * evaluated once and credited to both droplets (same accuracy contract,
 * different rounding than the one-sided kernels).  mode 0: one-sided kernels
 * only; 1 (default): pair-once for 10240 <= n <= 131072 when no explicit split
 * is set; 2: pair-once for every n, the split of dem_direct_set_tuning (when
 * > 0) giving the number of chunks of each block's partner list (the batched
 * pair-once sums equal these bit for bit with the split of
 * dem_batch_pair_once_split).  Workspace sizes depend on the mode.  Returns
 * the previous mode.  No counterpart in the reference (a tuning switch). */
int dem_direct_set_symmetric(int on);

/* Separations of the pair-once Oseen sum in mode 1 of
 * dem_direct_set_symmetric: on (default) takes them from 32-bit high words of
 * the positions -- exact in an open box; in a periodic box the pairs that lie
 * within one unit of half the box on some axis (where the 32-bit difference
 * can name the other periodic image) are found by an exact-match lookup of
 * the high words during dem_direct_forces_and_convection and their terms
 * replaced by the 64-bit ones, and a call where such a word repeats falls back
 * to 64-bit separations; dem_direct_convection alone keeps 64-bit separations
 * in a periodic box.  off: 64-bit separations, the kernels of mode 2.
 * Returns the previous setting.  No counterpart in the reference (a tuning
 * switch of this build's kernels). */
int dem_direct_set_image32(int on);

/* ------------------------------------------------------------------------ */
/* Spatial-hash contact detection                                            */
/* ------------------------------------------------------------------------ */

/* Workspace for n particles and a hash table of `table_size` buckets
 * (table_size = next_pow2(2 * capacity) in the reference,
 * spatial_hash_collision_detector.py:42-43). */
size_t dem_hash_workspace_bytes(int n, int table_size);

/* Contact pairs (i < j, dist <= r_i + r_j) found through the reference's
 * cell hash and 27-cell stencil (spatial_hash_collision_detector.py:73-229):
 * cell = 2*max(r), g = max(1, ceil(L/cell)), periodic wrap or open clamp,
 * hash (ix*73856093 ^ iy*19349663 ^ iz*83492791) & (table_size-1).
 * Pairs are written unique and lexicographically sorted to
 * pairs[0 .. 2*min(count, pair_capacity)); *count_out receives the number
 * of distinct pairs (may exceed pair_capacity: the caller then retries with a
 * larger buffer).  count_out is a device address: device memory, or pinned
 * host memory through its device address (hipHostGetDevicePointer).  The
 * count is stored (at system scope) as soon as it is known, before the last
 * kernel has written all pairs: a host may poll a pinned count word instead
 * of synchronising, and reads the pairs in stream order.  periodic: 1
 * periodic, 0 open. */
int dem_hash_pairs(const double* pos, const double* radii, int n, double L,
                   int periodic, int table_size, int32_t* pairs,
                   int pair_capacity, int32_t* count_out, void* ws,
                   size_t ws_bytes, dem_stream_t stream);

/* Bytes of the built-in scratch of dem_hash_pairs_noalloc (per GPU). */
#define DEM_HASH_NOALLOC_SCRATCH_BYTES (32u << 20)

/* dem_hash_pairs with no workspace argument and no allocation: the scratch is
 * a fixed array of DEM_HASH_NOALLOC_SCRATCH_BYTES in device memory, one per
 * GPU, held until the process ends.  Same results as dem_hash_pairs.  Sizes
 * whose workspace (dem_hash_workspace_bytes(n, table_size)) exceeds the
 * scratch are rejected with DEM_ERR_CAPACITY (n = 65536 with table_size
 * 131072 needs about 7 MB; n up to ~300000 fits with the reference's table
 * size).  Calls that overlap on one GPU share the scratch and race: issue
 * them on one stream or serialise them.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). */
int dem_hash_pairs_noalloc(const double* pos, const double* radii, int n, double L,
                           int periodic, int table_size, int32_t* pairs,
                           int pair_capacity, int32_t* count_out,
                           dem_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Explicit Euler step (float64, bit-exact with the reference's numpy)        */
/* ------------------------------------------------------------------------ */

/* pos += (forces / (stokes_factor * r) + conv) * dt, then pos %= L if
 * periodic (Python float-mod semantics).  euler_droplet_solver.py:124-131.
 * conv may be NULL (treated as zero). */
int dem_euler_step(double* pos, const double* forces, const double* conv,
                   const double* radii, int n, double stokes_factor, double dt,
                   double L, int periodic, dem_stream_t stream);

/* pos += vel * dt, then pos %= L if periodic (tree path,
 * euler_droplet_solver.py:119, :128-131). */
int dem_euler_step_total(double* pos, const double* vel, int n, double dt,
                         double L, int periodic, dem_stream_t stream);

/* Volume-conserving merge, compaction step (droplet_solution.py:263-268,
 * merge_collided's np.delete / np.append after the greedy plan): survivors of
 * (pos [n][3], radii [n]) keep their order, then the n_merged merged droplets
 * (merged [n_merged][4] = x, y, z, r) are appended, into out_pos
 * [n - n_removed + n_merged][3] and out_radii (distinct from the inputs).
 * removed: the n_removed distinct removed indices, ascending (device).
 * All pointers are device pointers. */
int dem_merge_compact(const double* pos, const double* radii, int n,
                      const int32_t* removed, int n_removed, const double* merged,
                      int n_merged, double* out_pos, double* out_radii,
                      dem_stream_t stream);

/* out[4 q ..] = (x, y, z, r) of droplet idx[q], q < k (NaN for an index
 * outside [0, n)): the droplet values a host-side merge plan reads
 * (droplet_solution.py:230-250).  Device pointers. */
int dem_gather_droplets(const double* pos, const double* radii, int n, const int32_t* idx,
                        int k, double* out, dem_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Flat Barnes-Hut octree                                                     */
/* ------------------------------------------------------------------------ */

#define DEM_TREE_MAX_LEVELS 64

/* Host-side description of a built tree.  The caller sets n_capacity and
 * max_nodes (reference FlatOctree.__init__: max_nodes =
 * max(4*max(N,64)//max(mpl,1), 1024), flat_tree.py:63-67) before the first
 * build; dem_octree_build fills the rest.  level_offsets[k] is the index in
 * the parents array where parent level k starts (the reference's
 * _parent_level_offsets). */
typedef struct dem_tree_info {
  int32_t n_capacity;
  int32_t max_nodes;
  int32_t n;
  int32_t node_count;
  int32_t parent_count;
  int32_t leaf_slot_count;
  int32_t num_levels;
  int32_t max_depth;
  int32_t level_offsets[DEM_TREE_MAX_LEVELS];
  int32_t periodic;
  int32_t mpl;
  int32_t forces_valid; /* particle_forces holds the last dem_octree_forces output */
  int32_t reserved;
  double L;
} dem_tree_info;

size_t dem_octree_workspace_bytes(int n_capacity, int max_nodes);

/* Opening tests of the walks: by default an FP32 test on fixed-point
 * separations decides wherever its error band allows and the reference's
 * float64 test (flat_tree.py:871, :1154, :1479) decides the rest, so the
 * interaction sets are the float64 test's.  on = 1 sends every test to the
 * float64 path (a checking switch for tests).  Returns the previous setting.
 * No counterpart in the reference. */
int dem_octree_set_exact_mac(int on);

/* Builds the tree of the root cube [0, L]^3 (FlatOctree.build,
 * flat_tree.py:180-246): same node numbering (breadth-first, children in
 * octant order, split when count > mpl, capacity clamp), leaf slots, next
 * pointers and R^3 aggregates.  m_const only sets the internal force unit.
 * Synchronises `stream` once per tree level. */
int dem_octree_build(const double* pos, const double* radii, int n, double L, int periodic,
                     int mpl, double m_const, dem_tree_info* info, void* ws, size_t ws_bytes,
                     dem_stream_t stream);

/* Dipole forces by the stackless walk with opening test s^2 < theta_sq * D^2
 * (_compute_forces_kernel,
 * flat_tree.py:783-906); also kept as the tree's particle_forces. */
int dem_octree_forces(dem_tree_info* info, double theta_sq, double m_const, double* forces,
                      void* ws, size_t ws_bytes, dem_stream_t stream);

/* Oseen velocities (monopole + dipole correction, optional COMSOL term and
 * self term; _compute_stokeslet_kernel, flat_tree.py:1000-1312).  forces ==
 * NULL uses particle_forces. */
int dem_octree_stokeslet(const dem_tree_info* info, double theta_sq, const double* forces,
                         double eta_const, const dem_correction* corr, double* vel, void* ws,
                         size_t ws_bytes, dem_stream_t stream);

/* Migration + Oseen velocity in one walk, without the COMSOL self term
 * (_compute_total_velocity_kernel, flat_tree.py:1326-1612). */
int dem_octree_total_velocity(const dem_tree_info* info, double theta_sq, const double* forces,
                              double eta_const, double stokes_factor,
                              const dem_correction* corr, double* vel, void* ws,
                              size_t ws_bytes, dem_stream_t stream);

/* Collision counts between volume classes by a pruned walk of the built tree
 * (TreeCollisionFrequency._count_collisions,
 * pbm/collision_frequency/tree_collision_freq.py:62-175).  For every particle
 * i the stackless walk skips empty nodes and internal nodes whose bounding box
 * is farther from x_i than r_i + max_radius (squared per-axis gaps; when
 * periodic a gap above L/2 becomes L - gap, :143-175), and at leaves every
 * j > i with |d|^2 <= (r_i + r_j)^2 (minimum image |d_a| > L/2 ->
 * d_a -= sign(d_a) L, float64) adds 1 to counts[bins[i]][bins[j]] and
 * counts[bins[j]][bins[i]].  bins: int32 class of each particle of the last
 * build (original order, info->n entries); counts: n_bins x n_bins int32,
 * overwritten.  A bin outside [0, n_bins) contributes nothing. */
int dem_octree_collision_counts(const dem_tree_info* info, const int32_t* bins, int n_bins,
                                int32_t* counts, void* ws, size_t ws_bytes, dem_stream_t stream);

/* Byte offset and size inside the octree workspace of a named node / particle
 * array ("count", "first_child", "next", "leaf_start", "min", "max",
 * "R3_sum", "R3_c", "force_sum", "fmom", "max_radius", "parents",
 * "parent_first_child", "leaf_indices", "perm", "particle_forces"). */
int dem_octree_field(const dem_tree_info* info, const char* name, size_t* offset,
                     size_t* bytes);

/* ------------------------------------------------------------------------ */
/* Batched ensembles: many independent periodic systems stepped together     */
/* ------------------------------------------------------------------------ */

/* The reference runs independent systems one after another (main_multi.py:
 * 43-71, a host loop over EulerDropletSolver.solve, euler_droplet_solver.py:
 * 43-159, with SpatialHashCollisionDetector.detect and
 * DirectDropletForceCalculator.calculate_forces_and_convection per step).
 * The batch keeps that per-system semantics — contact check first; on contact
 * the caller merges and checks again without advancing; then force, Oseen
 * velocity and the Euler step with periodic wrap — for all systems at once.
 * One global step:
 *   dem_batch_check(new_step=1)  hash pass of every active system: pair
 *                                counts and the packed pair list;
 *   dem_batch_advance            systems without contacts advance (this can
 *                                run while the host plans the merges);
 *   while some system has pairs: dem_batch_merge (the caller's merge plan),
 *                                dem_batch_check(new_step=0) (only the
 *                                systems the previous check found in contact);
 *                                or, on a second stream beside the advance:
 *                                dem_batch_wait_go, then dem_batch_merge_async
 *                                and the re-checks;
 *   dem_batch_advance            the remaining systems advance.
 * Each system advances at most once per global step ("stepped" flag).  Launch
 * sizes are fixed at the capacities, so the three captured HIP graphs stay
 * valid while merges shrink the systems.  Each system's sums use the source
 * split dem_batch_split(n, jlen); with that split passed to
 * dem_direct_set_tuning, dem_direct_forces_and_convection + dem_euler_step give
 * bit-identical results. */
typedef struct dem_batch_ensemble {
  int32_t offset;   /* first droplet of the system in the concatenated arrays  */
  int32_t n;        /* droplets now (<= capacity)                              */
  int32_t capacity; /* droplets reserved; also fixes the hash table size
                       next_pow2(2*capacity) (spatial_hash...py:42-43)         */
  int32_t steps;    /* Euler steps to take; the system stops after that many   */
  double L;         /* periodic box edge (m)                                   */
  double dt;        /* time step (s)                                           */
  double stokes_factor;
  double m_const;
  double eta_const;
} dem_batch_ensemble;

typedef struct dem_batch_tuning {
  int32_t block;      /* targets per work item: 64, 128 or 256                  */
  int32_t jlen;       /* sources per work item (source split); <= 0: no split   */
  int32_t force_tile; /* FP32 sum length before the float64 hand-off; 0: 32     */
  int32_t conv_tile;  /* same for the Oseen sum; 0: 16                          */
  int32_t contacts;   /* 1: contact pass; 0: every active system advances (the
                         solver without a collision detector)                   */
  int32_t pair_once;  /* 1: pair-once sums (Newton's third law, blocks of 256;
                         jlen / 256 partner blocks per work item; block unused);
                         0: one-sided sums                                      */
} dem_batch_tuning;

typedef struct dem_batch* dem_batch_t; /* opaque host handle */

/* Pair-once mode: chunks of each 256-droplet block's partner list of a system
 * of n droplets for a given jlen (the `split` to pass to dem_direct_set_tuning
 * with dem_direct_set_symmetric(2) for bit-identical single-system sums). */
int dem_batch_pair_once_split(int n, int jlen);

/* Source split of a system of n droplets for a given jlen (also the `split`
 * to pass to dem_direct_set_tuning for bit-identical single-system sums). */
int dem_batch_split(int n, int jlen);

size_t dem_batch_workspace_bytes(const dem_batch_ensemble* ens, int n_ens,
                                 const dem_batch_tuning* tuning);

/* Creates the handle over caller arrays pos (sum capacity x 3), radii, F, V
 * (float64, device) and the workspace; uploads the work tables and quantises
 * every system.  The handle owns host tables, a capture stream and the graphs;
 * the caller owns all device memory. */
int dem_batch_create(const dem_batch_ensemble* ens, int n_ens, const dem_batch_tuning* tuning,
                     double* pos, double* radii, double* forces, double* velocities,
                     void* ws, size_t ws_bytes, dem_stream_t stream, dem_batch_t* out);

/* Contact pass; new_step = 1 starts a global step (clears "stepped") and
 * checks every system with steps left, 0 re-checks only the systems the
 * previous check found in contact (merged by the caller since).  Results: fields
 * "pair_count", "pair_base" (exclusive scan of pair_count) and the packed
 * "pairs" / "pair_values" (system after system, each lexicographically
 * sorted; values are (x, y, z, r) of i then of j).  A total above the
 * buffer's capacity is truncated. */
int dem_batch_check(dem_batch_t b, int new_step, dem_stream_t stream);

/* Force + Oseen + Euler of every active system whose last check found no
 * contact and that has not advanced in this global step. */
int dem_batch_advance(dem_batch_t b, dem_stream_t stream);

/* Applies the caller's merges (droplet_solution.py:208-287 computed on the
 * host): for systems[k], the droplets removed[removed_off[k] ..
 * removed_off[k+1]) (local indices, ascending) leave, survivors keep their
 * order, and the merged droplets merged[4*merged_off[k] .. 4*merged_off[k+1])
 * ((x, y, z, r) each) are appended.  Host arrays; rebuilds the work tables,
 * hash cells and records.  Synchronises `stream`. */
int dem_batch_merge(dem_batch_t b, int count, const int32_t* systems, const int32_t* removed_off,
                    const int32_t* removed, const int32_t* merged_off, const double* merged,
                    dem_stream_t stream);

/* dem_batch_merge without the rebuild of the work tables: only the merged
 * systems' descriptors, hash cells and records are rewritten (on `stream`),
 * so it may run on a second stream while an advance of the other systems
 * runs (after dem_batch_wait_go on that stream).  The tables are rebuilt by
 * the next first check or advance; a re-check (new_step=0) does not need
 * them.  Synchronises `stream`.  No reference counterpart (scheduling). */
int dem_batch_merge_async(dem_batch_t b, int count, const int32_t* systems,
                          const int32_t* removed_off, const int32_t* removed,
                          const int32_t* merged_off, const double* merged,
                          dem_stream_t stream);

/* on = 1 declares that every later dem_batch_merge / dem_batch_merge_async
 * applies, for each listed system, the greedy merge of the complete pair list
 * of that system's last check (as dem_merge_plan computes it; the reference's
 * merge, droplet_solution.py:208-262).  A re-check of such a system may then
 * search only around the merged droplets: no pair of two survivors is left,
 * so when the hash cell is unchanged the re-check's pairs are exactly those
 * with a merged droplet, found without searching the survivors (same pair
 * lists as a full re-check).  0 (the default) makes every re-check a full one.
 * No reference counterpart (scheduling). */
int dem_batch_set_incremental(dem_batch_t b, int on);

/* Makes `stream` wait until the last dem_batch_advance has decided which
 * systems it moves; from then on the systems it leaves alone (those found in
 * contact) may be merged and re-checked on `stream` while it runs.  No
 * reference counterpart (scheduling). */
int dem_batch_wait_go(dem_batch_t b, dem_stream_t stream);

/* Greedy merge plans of `count` systems, host only (no GPU): system k's
 * pairs are pairs[pair_off[k] .. pair_off[k+1]) (local (i, j)) with
 * values[8 p ..] = (x, y, z, r) of i then of j, visited in the order
 * order[pair_off[k] + q] (positions inside the system's segment: the
 * reference's np.argsort of the first column).  A pair whose droplet already
 * merged is skipped; otherwise the merged droplet r = (r_i^3 + r_j^3)^(1/3) at
 * the r^3-weighted centre is appended to merged[4 m ..] and i, j are removed.
 * Outputs: removed[removed_off[k] .. removed_off[k+1]) ascending and
 * merged_off (count + 1 each; at most 2 removed and 1 merged per pair): the
 * lists dem_batch_merge takes.  merge_plan_rows / droplet_solution.py:208-262
 * with the same IEEE operations. */
int dem_merge_plan(int count, const int32_t* pair_off, const int32_t* pairs,
                   const double* values, const int32_t* order, int32_t* removed_off,
                   int32_t* removed, int32_t* merged_off, double* merged);

/* Systems systems[k] now hold n[k] droplets (the caller rewrote their slices
 * of pos/radii itself): rebuilds the work tables, hash cells and records.
 * Synchronises `stream`. */
int dem_batch_resize(dem_batch_t b, int count, const int32_t* systems, const int32_t* n,
                     dem_stream_t stream);

/* Captures the first check, the re-check and the advance as HIP graphs (on
 * the handle's own stream); later check/advance/run calls launch them. */
int dem_batch_capture(dem_batch_t b);

/* `steps` x (check, advance) without merging: systems with a contact stay. */
int dem_batch_run(dem_batch_t b, int steps, dem_stream_t stream);

/* Device address and size of "active", "stepped", "pair_count", "steps_done",
 * "incremental" (int32 per system: 1 when the last re-check searched only
 * around the merged droplets), "pair_base" (int32, n_ens + 1), "time" (float64 per
 * system: sum of dt over advanced steps), "pairs" (int32 (i, j), local
 * indices), "pair_values" (float64 (x, y, z, r) x 2 per pair). */
int dem_batch_field(dem_batch_t b, const char* name, void** ptr, size_t* bytes);

int dem_batch_destroy(dem_batch_t b);

/* ------------------------------------------------------------------------ */
/* Differentiation: adjoint of one direct Euler step, near-contact loss      */
/* ------------------------------------------------------------------------ */

/* The reference has no gradients; these support differentiating trajectories
 * of the direct path (forces direct_droplet_force_calculator.py:183-253,
 * Oseen :354-411, Euler euler_droplet_solver.py:113-131) through a
 * hand-written adjoint.  Periodic boxes only. */
size_t dem_adjoint_workspace_bytes(int n);

/* For positions x (the state of one step), the forward forces F at x and an
 * adjoint vector a (all (n,3) float64), with v = F/(sf r) + V(x, F):
 *   grad_x = (dv/dx)^T a                                   ((n,3) float64)
 *   sums[0] = sum_j c_j . F_j   (dJ/dm_const  = sums[0] / m_const)
 *   sums[1] = sum_j F_j . u_j   (dJ/deta_const = sums[1] / eta_const)
 *   sums[2] = sum_i b_i . F_i   (dJ/dstokes   = -sums[2] / stokes_factor)
 * for J = sum_i a_i . v_i, where b_i = a_i/(sf r_i), u = eta sum G a and
 * c = b + u.  sums is a device array of 3 doubles.  The contact mask is taken
 * as locally constant.  FP32 pair arithmetic, float64 accumulation. */
int dem_direct_step_vjp(const double* pos, const double* radii, const double* forces,
                        const double* adj, int n, const dem_phys* phys, double stokes_factor,
                        double* grad_x, double* sums, void* ws, size_t ws_bytes,
                        dem_stream_t stream);

/* loss = sum_{i<j} softplus(unit * (factor (r_i + r_j) - |d_ij|)) (minimum
 * image) into *loss (device double) and its gradient d loss / d x into grad
 * ((n,3) float64, per metre).  unit = 1e6 measures lengths in micrometres. */
int dem_contact_loss(const double* pos, const double* radii, int n, const dem_phys* phys,
                     double factor, double unit, double* loss, double* grad, void* ws,
                     size_t ws_bytes, dem_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Kernel timing (measurement support; no counterpart in the reference)       */
/* ------------------------------------------------------------------------ */

/* While started, every launch of a hot kernel is bracketed by a HIP event
 * pair recorded on the launch stream.  dem_profile_read() synchronises those
 * events and returns the summed duration (ms) and launch count of `kernel`
 * (e.g. "k_force_tile", "k_conv_tile").  Process-wide, single-threaded use. */
int dem_profile_start(void);
int dem_profile_stop(void);
int dem_profile_read(const char* kernel, double* total_ms, int* launches);

#ifdef __cplusplus
}
#endif

#endif /* DEM_HIP_H */
