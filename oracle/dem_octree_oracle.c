/*
 * dem_octree_oracle.c — FP64 CPU restatement of the reference's flat
 * Barnes-Hut octree (dem/octree/flat_tree.py, octree_node.py).
 *
 * TEST INFRASTRUCTURE ONLY (see dem_oracle.c).  The build follows the
 * reference's own algorithm — per level: octant assignment, serial split of
 * overfull leaves in node order with the capacity check, scatter — rather
 * than the GPU's sort-based construction, so structural parity between the
 * two is a real check.  Particles inside a leaf are listed in ascending index
 * order (the reference's order there is an atomic race).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct oracle_tree {
  int n, max_nodes;
  const double* pos; /* (n,3) */
  const double* r;   /* (n,)  */
  double L;
  int periodic, mpl;
  /* nodes */
  double* bmin;  /* 3*max_nodes */
  double* bmax;  /* 3*max_nodes */
  double* r3sum; /* max_nodes */
  double* r3c;   /* 3*max_nodes */
  double* fsum;  /* 3*max_nodes */
  double* fmom;  /* 9*max_nodes */
  double* maxr;  /* max_nodes */
  int* first_child;
  int* next;
  int* leaf_start;
  int* count;
  int* parents;
  int* pfc;
  int* leaf_indices; /* n */
  int level_offsets[64];
  int num_levels;
  int node_count, parent_count, leaf_slot_count;
} oracle_tree;

static inline double mic(double d, double L, int periodic) {
  if (periodic && fabs(d) > L * 0.5) d -= (d > 0.0 ? L : -L);
  return d;
}

static void init_node(oracle_tree* t, int i, double mnx, double mny, double mnz, double mxx,
                      double mxy, double mxz) {
  t->bmin[3 * i] = mnx;
  t->bmin[3 * i + 1] = mny;
  t->bmin[3 * i + 2] = mnz;
  t->bmax[3 * i] = mxx;
  t->bmax[3 * i + 1] = mxy;
  t->bmax[3 * i + 2] = mxz;
  t->first_child[i] = -1;
  t->next[i] = -1;
  t->leaf_start[i] = -1;
  t->count[i] = 0;
}

/* FlatOctree.build (flat_tree.py:180-246).  Returns 0 or -1 on a bad size. */
int oracle_tree_build(oracle_tree* t) {
  const int n = t->n, M = t->max_nodes;
  int* part_node = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int* part_oct = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int* child_cnt = (int*)calloc((size_t)M * 8, sizeof(int));
  t->node_count = 1;
  t->parent_count = 0;
  t->leaf_slot_count = 0;
  t->num_levels = 0;
  init_node(t, 0, 0.0, 0.0, 0.0, t->L, t->L, t->L);
  t->count[0] = n;
  for (int i = 0; i < n; ++i) part_node[i] = 0;
  int lstart = 0, lend = 1;
  for (int level = 0; level < 40; ++level) {
    int overflow = 0;
    const int pc_before = t->parent_count;
    /* assign octants */
    for (int node = lstart; node < lend; ++node)
      for (int o = 0; o < 8; ++o) child_cnt[8 * node + o] = 0;
    for (int i = 0; i < n; ++i) {
      const int node = part_node[i];
      if (node >= lstart && node < lend) {
        const double cx = (t->bmin[3 * node] + t->bmax[3 * node]) * 0.5;
        const double cy = (t->bmin[3 * node + 1] + t->bmax[3 * node + 1]) * 0.5;
        const double cz = (t->bmin[3 * node + 2] + t->bmax[3 * node + 2]) * 0.5;
        int o = 0;
        if (t->pos[3 * i] >= cx) o |= 1;
        if (t->pos[3 * i + 1] >= cy) o |= 2;
        if (t->pos[3 * i + 2] >= cz) o |= 4;
        part_oct[i] = o;
        child_cnt[8 * node + o]++;
      }
    }
    /* serial split */
    for (int node = lstart; node < lend; ++node) {
      if (t->first_child[node] < 0 && t->count[node] > t->mpl) {
        const int base = t->node_count;
        if (base + 8 > M) continue;
        t->node_count = base + 8;
        t->first_child[node] = base;
        t->parents[t->parent_count] = node;
        t->pfc[t->parent_count] = base;
        t->parent_count++;
        const double mnx = t->bmin[3 * node], mny = t->bmin[3 * node + 1], mnz = t->bmin[3 * node + 2];
        const double mxx = t->bmax[3 * node], mxy = t->bmax[3 * node + 1], mxz = t->bmax[3 * node + 2];
        const double cx = (mnx + mxx) * 0.5, cy = (mny + mxy) * 0.5, cz = (mnz + mxz) * 0.5;
        for (int o = 0; o < 8; ++o) {
          init_node(t, base + o, (o & 1) ? cx : mnx, (o & 2) ? cy : mny, (o & 4) ? cz : mnz,
                    (o & 1) ? mxx : cx, (o & 2) ? mxy : cy, (o & 4) ? mxz : cz);
          t->count[base + o] = child_cnt[8 * node + o];
        }
        overflow = 1;
      }
    }
    /* scatter */
    for (int i = 0; i < n; ++i) {
      const int node = part_node[i];
      if (node >= 0 && t->first_child[node] >= 0) part_node[i] = t->first_child[node] + part_oct[i];
    }
    lstart = lend;
    lend = t->node_count;
    if (t->parent_count > pc_before) t->level_offsets[t->num_levels++] = pc_before;
    if (!overflow) break;
  }
  /* exact leaf slots, then fill in ascending particle order */
  for (int node = 0; node < t->node_count; ++node) {
    if (t->first_child[node] < 0 && t->count[node] > 0) {
      t->leaf_start[node] = t->leaf_slot_count;
      t->leaf_slot_count += t->count[node];
      t->count[node] = 0;
    }
  }
  for (int i = 0; i < n; ++i) {
    const int node = part_node[i];
    if (node >= 0 && t->first_child[node] < 0 && t->leaf_start[node] >= 0)
      t->leaf_indices[t->leaf_start[node] + t->count[node]++] = i;
  }
  /* next pointers: pass 1 (sibling chains) then pass 2 (unlink empties), top-down */
  for (int pass = 0; pass < 2; ++pass)
    for (int lev = 0; lev < t->num_levels; ++lev) {
      const int p0 = t->level_offsets[lev];
      const int p1 = lev + 1 < t->num_levels ? t->level_offsets[lev + 1] : t->parent_count;
      for (int p = p0; p < p1; ++p) {
        const int parent = t->parents[p], base = t->pfc[p];
        if (pass == 0) {
          const int pn = t->next[parent];
          for (int o = 0; o < 8; ++o) t->next[base + o] = (o == 7) ? pn : base + o + 1;
        } else {
          if (t->count[parent] == 0) {
            t->first_child[parent] = t->next[parent];
            continue;
          }
          int first = -1, last = -1;
          for (int j = 0; j < 8; ++j) {
            const int c = base + j;
            if (t->count[c] > 0) {
              if (first < 0) first = c;
              else t->next[last] = c;
              last = c;
            }
          }
          if (last >= 0) {
            t->next[last] = t->next[parent];
            t->first_child[parent] = first;
          }
        }
      }
    }
  /* R^3 aggregates: leaves, then parent levels bottom-up */
  for (int i = 0; i < t->node_count; ++i) {
    double s = 0.0, wx = 0.0, wy = 0.0, wz = 0.0, mr = 0.0;
    if (t->first_child[i] < 0 && t->count[i] > 0 && t->leaf_start[i] >= 0) {
      for (int k = 0; k < t->count[i]; ++k) {
        const int p = t->leaf_indices[t->leaf_start[i] + k];
        const double ri = t->r[p], r3 = ri * ri * ri;
        s += r3;
        wx += t->pos[3 * p] * r3;
        wy += t->pos[3 * p + 1] * r3;
        wz += t->pos[3 * p + 2] * r3;
        if (ri > mr) mr = ri;
      }
    }
    t->r3sum[i] = s;
    t->r3c[3 * i] = wx;
    t->r3c[3 * i + 1] = wy;
    t->r3c[3 * i + 2] = wz;
    t->maxr[i] = mr;
  }
  for (int lev = t->num_levels - 1; lev >= 0; --lev) {
    const int p0 = t->level_offsets[lev];
    const int p1 = lev + 1 < t->num_levels ? t->level_offsets[lev + 1] : t->parent_count;
    for (int p = p0; p < p1; ++p) {
      const int parent = t->parents[p], base = t->pfc[p];
      double s = 0.0, wx = 0.0, wy = 0.0, wz = 0.0, mr = 0.0;
      for (int j = 0; j < 8; ++j) {
        const int c = base + j;
        if (t->count[c] > 0) {
          s += t->r3sum[c];
          wx += t->r3c[3 * c];
          wy += t->r3c[3 * c + 1];
          wz += t->r3c[3 * c + 2];
          if (t->maxr[c] > mr) mr = t->maxr[c];
        }
      }
      t->r3sum[parent] = s;
      t->r3c[3 * parent] = wx;
      t->r3c[3 * parent + 1] = wy;
      t->r3c[3 * parent + 2] = wz;
      t->maxr[parent] = mr;
    }
  }
  free(part_node);
  free(part_oct);
  free(child_cnt);
  return 0;
}

/* Force sums and raw moments, leaves then parent levels (flat_tree.py:556-694). */
void oracle_tree_propagate_forces(oracle_tree* t, const double* F) {
  for (int i = 0; i < t->node_count; ++i) {
    double f[3] = {0, 0, 0}, P[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (t->first_child[i] < 0 && t->count[i] > 0 && t->leaf_start[i] >= 0) {
      for (int k = 0; k < t->count[i]; ++k) {
        const int p = t->leaf_indices[t->leaf_start[i] + k];
        for (int a = 0; a < 3; ++a) f[a] += F[3 * p + a];
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) P[3 * a + b] += t->pos[3 * p + a] * F[3 * p + b];
      }
    }
    for (int a = 0; a < 3; ++a) t->fsum[3 * i + a] = f[a];
    for (int a = 0; a < 9; ++a) t->fmom[9 * i + a] = P[a];
  }
  for (int lev = t->num_levels - 1; lev >= 0; --lev) {
    const int p0 = t->level_offsets[lev];
    const int p1 = lev + 1 < t->num_levels ? t->level_offsets[lev + 1] : t->parent_count;
    for (int p = p0; p < p1; ++p) {
      const int parent = t->parents[p], base = t->pfc[p];
      double f[3] = {0, 0, 0}, P[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      for (int j = 0; j < 8; ++j) {
        const int c = base + j;
        if (t->count[c] > 0) {
          for (int a = 0; a < 3; ++a) f[a] += t->fsum[3 * c + a];
          for (int a = 0; a < 9; ++a) P[a] += t->fmom[9 * c + a];
        }
      }
      for (int a = 0; a < 3; ++a) t->fsum[3 * parent + a] = f[a];
      for (int a = 0; a < 9; ++a) t->fmom[9 * parent + a] = P[a];
    }
  }
}

/* Mistakes planted by tests (0 in every other use), to show that the
 * far-field order tests detect them: bit 1 drops the trace term of the
 * Stokeslet dipole, bit 2 drops the whole dipole, bit 4 evaluates the force
 * far field at the node's geometric centre instead of its R^3 centroid. */
static int g_plant = 0;
void oracle_set_plant(int plant) { g_plant = plant; }

static void get_center(const oracle_tree* t, int i, double c[3]) {
  if (t->r3sum[i] > 0.0) {
    const double inv = 1.0 / t->r3sum[i];
    for (int a = 0; a < 3; ++a) c[a] = t->r3c[3 * i + a] * inv;
  } else {
    for (int a = 0; a < 3; ++a) c[a] = (t->bmin[3 * i + a] + t->bmax[3 * i + a]) * 0.5;
  }
}

static double node_size(const oracle_tree* t, int i) {
  double s = t->bmax[3 * i] - t->bmin[3 * i];
  const double ey = t->bmax[3 * i + 1] - t->bmin[3 * i + 1];
  const double ez = t->bmax[3 * i + 2] - t->bmin[3 * i + 2];
  if (ey > s) s = ey;
  if (ez > s) s = ez;
  return s;
}

/* _compute_forces_kernel (flat_tree.py:783-894), targets [i0, i1). */
void oracle_tree_forces(const oracle_tree* t, double theta, double m_const, double* Fout, int i0,
                        int i1) {
  const double theta_sq = theta * theta;
  for (int i = i0; i < i1; ++i) {
    const double* pi = t->pos + 3 * i;
    const double ri3 = t->r[i] * t->r[i] * t->r[i];
    double fx = 0.0, fy = 0.0, fz = 0.0;
    int node = 0;
    while (node >= 0) {
      const int cnt = t->count[node];
      if (cnt == 0 || t->r3sum[node] < 1e-30) {
        node = t->next[node];
        continue;
      }
      if (t->first_child[node] < 0) {
        const int ls = t->leaf_start[node];
        for (int k = 0; k < cnt; ++k) {
          const int j = t->leaf_indices[ls + k];
          if (j < 0 || j == i) continue;
          const double dx = mic(t->pos[3 * j] - pi[0], t->L, t->periodic);
          const double dy = mic(t->pos[3 * j + 1] - pi[1], t->L, t->periodic);
          const double dz = mic(t->pos[3 * j + 2] - pi[2], t->L, t->periodic);
          const double r_sq = dx * dx + dy * dy + dz * dz;
          const double rs = t->r[i] + t->r[j];
          if (r_sq >= rs * rs) {
            const double rr = sqrt(r_sq);
            const double r7 = r_sq * r_sq * r_sq * rr;
            const double rj3 = t->r[j] * t->r[j] * t->r[j];
            const double factor = m_const * ri3 * rj3 / r7;
            const double dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
            const double axy = 4.0 * dz2 - dx2 - dy2;
            fx += factor * dx * axy;
            fy += factor * dy * axy;
            fz += factor * dz * (2.0 * dz2 - 3.0 * dx2 - 3.0 * dy2);
          }
        }
        node = t->next[node];
        continue;
      }
      const double s = node_size(t, node);
      double c[3];
      get_center(t, node, c);
      if (g_plant & 4)
        for (int a = 0; a < 3; ++a) c[a] = (t->bmin[3 * node + a] + t->bmax[3 * node + a]) * 0.5;
      const double dx = mic(c[0] - pi[0], t->L, t->periodic);
      const double dy = mic(c[1] - pi[1], t->L, t->periodic);
      const double dz = mic(c[2] - pi[2], t->L, t->periodic);
      const double D_sq = dx * dx + dy * dy + dz * dz;
      if (s * s < theta_sq * D_sq && D_sq > 0.0) {
        const double rr = sqrt(D_sq);
        const double r7 = D_sq * D_sq * D_sq * rr;
        const double factor = m_const * ri3 * t->r3sum[node] / r7;
        const double dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
        const double axy = 4.0 * dz2 - dx2 - dy2;
        fx += factor * dx * axy;
        fy += factor * dy * axy;
        fz += factor * dz * (2.0 * dz2 - 3.0 * dx2 - 3.0 * dy2);
        node = t->next[node];
      } else {
        node = t->first_child[node];
      }
    }
    Fout[3 * i] = fx;
    Fout[3 * i + 1] = fy;
    Fout[3 * i + 2] = fz;
  }
}

typedef struct oracle_corr oracle_corr; /* dem_oracle.c */
void oracle_corr_pair_term(const oracle_corr* c, double x, double y, double z, double fx,
                           double fy, double fz, double* vx, double* vy, double* vz);
double oracle_corr_self(const oracle_corr* c);

/* _compute_stokeslet_kernel (mode 0, :1000-1280) and
 * _compute_total_velocity_kernel (mode 1, :1326-1588); needs
 * oracle_tree_propagate_forces(F) first.  Fself: forces used for the self /
 * migration terms. */
void oracle_tree_stokes(const oracle_tree* t, int mode, double theta, const double* F,
                        double eta, double stokes, const oracle_corr* corr, double* V, int i0,
                        int i1) {
  const double theta_sq = theta * theta;
  for (int i = i0; i < i1; ++i) {
    const double* pi = t->pos + 3 * i;
    double vx = 0.0, vy = 0.0, vz = 0.0;
    if (mode == 1) {
      const double inv = 1.0 / (stokes * t->r[i]);
      vx = F[3 * i] * inv;
      vy = F[3 * i + 1] * inv;
      vz = F[3 * i + 2] * inv;
    }
    int node = 0;
    while (node >= 0) {
      const int cnt = t->count[node];
      const double fs2 = t->fsum[3 * node] * t->fsum[3 * node] +
                         t->fsum[3 * node + 1] * t->fsum[3 * node + 1] +
                         t->fsum[3 * node + 2] * t->fsum[3 * node + 2];
      if (cnt == 0 || fs2 < 1e-60) {
        node = t->next[node];
        continue;
      }
      if (t->first_child[node] < 0) {
        const int ls = t->leaf_start[node];
        for (int k = 0; k < cnt; ++k) {
          const int j = t->leaf_indices[ls + k];
          if (j < 0 || j == i) continue;
          const double dx = mic(pi[0] - t->pos[3 * j], t->L, t->periodic);
          const double dy = mic(pi[1] - t->pos[3 * j + 1], t->L, t->periodic);
          const double dz = mic(pi[2] - t->pos[3 * j + 2], t->L, t->periodic);
          const double r_sq = dx * dx + dy * dy + dz * dz;
          const double rs = t->r[i] + t->r[j];
          if (r_sq >= rs * rs) {
            const double rr = sqrt(r_sq), inv_r = 1.0 / rr, inv_r2 = inv_r * inv_r;
            const double Fx = F[3 * j], Fy = F[3 * j + 1], Fz = F[3 * j + 2];
            const double dot = dx * Fx + dy * Fy + dz * Fz;
            const double coeff = eta * inv_r;
            vx += coeff * (Fx + dx * inv_r2 * dot);
            vy += coeff * (Fy + dy * inv_r2 * dot);
            vz += coeff * (Fz + dz * inv_r2 * dot);
            if (corr) oracle_corr_pair_term(corr, dx, dy, dz, Fx, Fy, Fz, &vx, &vy, &vz);
          }
        }
        node = t->next[node];
        continue;
      }
      const double s = node_size(t, node);
      double c[3];
      get_center(t, node, c);
      const double dx = mic(pi[0] - c[0], t->L, t->periodic);
      const double dy = mic(pi[1] - c[1], t->L, t->periodic);
      const double dz = mic(pi[2] - c[2], t->L, t->periodic);
      const double D_sq = dx * dx + dy * dy + dz * dz;
      if (s * s < theta_sq * D_sq && D_sq > 0.0) {
        const double rr = sqrt(D_sq), inv_r = 1.0 / rr, inv_r2 = inv_r * inv_r;
        const double Fx = t->fsum[3 * node], Fy = t->fsum[3 * node + 1], Fz = t->fsum[3 * node + 2];
        const double dot = dx * Fx + dy * Fy + dz * Fz;
        const double coeff = eta * inv_r;
        vx += coeff * (Fx + dx * inv_r2 * dot);
        vy += coeff * (Fy + dy * inv_r2 * dot);
        vz += coeff * (Fz + dz * inv_r2 * dot);
        if (corr) oracle_corr_pair_term(corr, dx, dy, dz, Fx, Fy, Fz, &vx, &vy, &vz);
        double cm[3];
        get_center(t, node, cm);
        const double* P = t->fmom + 9 * node;
        const double Fv[3] = {Fx, Fy, Fz};
        double D[9];
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) D[3 * a + b] = P[3 * a + b] - cm[a] * Fv[b];
        if (g_plant & 2)
          for (int a = 0; a < 9; ++a) D[a] = 0.0;
        const double rx = dx, ry = dy, rz = dz;
        const double ux = D[0] * rx + D[3] * ry + D[6] * rz;
        const double uy = D[1] * rx + D[4] * ry + D[7] * rz;
        const double uz = D[2] * rx + D[5] * ry + D[8] * rz;
        const double wx = D[0] * rx + D[1] * ry + D[2] * rz;
        const double wy = D[3] * rx + D[4] * ry + D[5] * rz;
        const double wz = D[6] * rx + D[7] * ry + D[8] * rz;
        const double s_tr = (g_plant & 1) ? 0.0 : D[0] + D[4] + D[8];
        const double q = rx * wx + ry * wy + rz * wz;
        const double inv_r3 = inv_r * inv_r2, inv_r5 = inv_r3 * inv_r2;
        vx += eta * (inv_r3 * (ux - wx - s_tr * rx) + 3.0 * rx * q * inv_r5);
        vy += eta * (inv_r3 * (uy - wy - s_tr * ry) + 3.0 * ry * q * inv_r5);
        vz += eta * (inv_r3 * (uz - wz - s_tr * rz) + 3.0 * rz * q * inv_r5);
        node = t->next[node];
      } else {
        node = t->first_child[node];
      }
    }
    if (mode == 0 && corr) {
      const double g = oracle_corr_self(corr);
      vx += g * F[3 * i];
      vy += g * F[3 * i + 1];
      vz += g * F[3 * i + 2];
    }
    V[3 * i] = vx;
    V[3 * i + 1] = vy;
    V[3 * i + 2] = vz;
  }
}

/* ------------------------------------------------------------------ PBM */

/* Squared distance from p to a node box with the reference's periodic rule
 * (TreeCollisionFrequency._min_dist_sq_to_bbox,
 * pbm/collision_frequency/tree_collision_freq.py:143-175): per axis the gap
 * outside [bmin, bmax]; when periodic, a gap above L/2 becomes L - gap. */
static double min_dist_sq_box(const oracle_tree* t, int node, const double p[3]) {
  double g[3];
  for (int a = 0; a < 3; ++a) {
    const double lo = t->bmin[3 * node + a], hi = t->bmax[3 * node + a];
    double d = 0.0;
    if (p[a] < lo)
      d = lo - p[a];
    else if (p[a] > hi)
      d = p[a] - hi;
    if (t->periodic && t->L > 0.0 && d > t->L * 0.5) d = t->L - d;
    g[a] = d;
  }
  return g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
}

/* TreeCollisionFrequency._count_collisions (tree_collision_freq.py:62-140):
 * for every particle i the stackless walk skips empty nodes, prunes internal
 * nodes whose box is farther than r_i + max_radius, and at leaves counts every
 * j > i with |d|^2 <= (r_i + r_j)^2 (minimum image |d| > L/2 -> d -= sign(d) L)
 * into counts[bin_i][bin_j] and counts[bin_j][bin_i].  counts: n_bins^2,
 * zeroed here. */
void oracle_tree_collision_counts(const oracle_tree* t, const int* bins, int n_bins,
                                  int* counts) {
  memset(counts, 0, sizeof(int) * (size_t)n_bins * (size_t)n_bins);
  for (int i = 0; i < t->n; ++i) {
    const double* pi = t->pos + 3 * i;
    const double ri = t->r[i];
    const int bi = bins[i];
    int node = 0;
    while (node >= 0) {
      const int cnt = t->count[node];
      if (cnt == 0) {
        node = t->next[node];
        continue;
      }
      if (t->first_child[node] < 0) {
        const int ls = t->leaf_start[node];
        for (int k = 0; k < cnt; ++k) {
          const int j = t->leaf_indices[ls + k];
          if (j >= 0 && j > i) {
            const double* pj = t->pos + 3 * j;
            const double dx = mic(pj[0] - pi[0], t->L, t->periodic);
            const double dy = mic(pj[1] - pi[1], t->L, t->periodic);
            const double dz = mic(pj[2] - pi[2], t->L, t->periodic);
            const double d2 = dx * dx + dy * dy + dz * dz;
            const double c = ri + t->r[j];
            if (d2 <= c * c) {
              const int bj = bins[j];
              counts[bi * n_bins + bj] += 1;
              counts[bj * n_bins + bi] += 1;
            }
          }
        }
        node = t->next[node];
        continue;
      }
      const double mc = ri + t->maxr[node];
      if (min_dist_sq_box(t, node, pi) > mc * mc)
        node = t->next[node];
      else
        node = t->first_child[node];
    }
  }
}
