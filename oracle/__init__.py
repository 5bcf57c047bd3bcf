"""FP64 CPU oracle of the reference DEM hot path — TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import this package, and only as the checker (or the
timed CPU baseline).  The product package ``dem`` never imports it; the HIP
path fails loudly when its extension is missing instead of falling back here.

What each function restates (reference file:line) and what pins it
(tests/test_oracle_*.py):

* ``direct_forces`` — ``_calculate_forces`` + ``_calculate_one_force``
  (dem/force_calculator/direct_droplet_force_calculator.py:183-253).  Pinned by
  the two-body closed forms along z and x, Newton's third law, F = -grad U of
  U = (m/3) r_i^3 r_j^3 (r^2 - 3 dz^2)/r^5 (central finite differences), the
  minimum-image sign flip (reference test_direct...py:77), zero force on a
  symmetric cubic lattice, E^2 / length^2 scaling and a pure-Python loop.
* ``direct_convection`` — ``_calculate_convection_velocities`` (:354-468).
  Pinned by the two-body Oseen closed form v = eta/r (F + (F.r^)r^), symmetry
  of the Oseen tensor, v proportional to 1/eta, and a pure-Python loop.  The
  COMSOL term is pinned by trilinear interpolation reproducing grid nodes and
  the reference's own symmetry / convolution identities
  (tests/periodic_correction/test_comsol_lattice_correction.py:62-142).
* ``hash_pairs`` — ``SpatialHashCollisionDetector.detect``
  (dem/collision_detector/spatial_hash_collision_detector.py:73-316).  Pinned
  by the reference's exact small cases (test_spatial_hash...py:24-113) and by
  brute force: every hash pair is a true contact, and the hash set equals the
  brute-force set filtered by the 27-cell stencil (the reference's periodic
  miss included).
* ``merge`` — ``DropletSolution.generate_next_solution``
  (dem/solution/droplet_solution.py:208-287).  Pinned by the reference's
  test_collision_merges_two_drops (3 -> 2 drops, r^3 conserved, merged drop
  appended last) plus the greedy-skip rule on chains.
* ``euler_step`` — ``EulerDropletSolver.solve`` update
  (dem/solver/euler_droplet_solver.py:119-131).  Pinned by hand-computed
  single steps and the periodic wrap.
* Octree functions (``octree_*``) — ``FlatOctree`` (dem/octree/flat_tree.py).
  Pinned by the build invariants of the reference tests, theta = 0 equalling
  the direct sum, and the reference's tree-vs-direct error bands.
* ``direct_collision_counts`` / ``OracleTree.collision_counts`` — the
  population-balance collision counts (pbm/collision_frequency/
  direct_collision_freq.py:15-41, tree_collision_freq.py:62-175).  Pinned by
  the reference's own cases (tests/pbm/collision_frequency/*.py), a
  pure-Python restatement of both loops on small systems, and brute force:
  for open boxes the pruned walk counts exactly the brute-force pairs under
  |d|^2 <= (r_i + r_j)^2.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "build" / "liboracle.so"

_lib = None


def build(force: bool = False) -> Path:
    """Compile the C restatement (gcc, OpenMP) into oracle/build/."""
    if force or not LIB_PATH.exists() or any(
        p.stat().st_mtime > LIB_PATH.stat().st_mtime for p in HERE.glob("*.c")
    ):
        subprocess.run(["make", "-C", str(HERE), "-B" if force else "all"], check=True,
                       capture_output=True)
    return LIB_PATH


class _Corr(ctypes.Structure):
    _fields_ = [
        ("u", ctypes.c_void_p),
        ("v", ctypes.c_void_p),
        ("w", ctypes.c_void_p),
        ("n", ctypes.c_int),
        ("enabled", ctypes.c_int),
        ("gmin", ctypes.c_double),
        ("inv_dx", ctypes.c_double),
        ("Fz_inv", ctypes.c_double),
        ("L_ratio", ctypes.c_double),
        ("eta_ratio", ctypes.c_double),
        ("self_coeff", ctypes.c_double),
    ]


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        build()
        _lib = ctypes.CDLL(str(LIB_PATH))
        d, i, p = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
        _lib.oracle_direct_forces_range.argtypes = [i, i, i, p, p, d, d, i, p]
        _lib.oracle_direct_convection_range.argtypes = [i, i, i, p, p, d, d, i, p, p]
        _lib.oracle_corr_eval.argtypes = [p, i, p, p]
        _lib.oracle_hash_pairs.argtypes = [i, p, p, d, i, i, p, i]
        _lib.oracle_hash_pairs.restype = i
        _lib.oracle_brute_pairs.argtypes = [i, p, p, d, i, p, i]
        _lib.oracle_brute_pairs.restype = i
        _lib.oracle_num_threads.restype = i
        _lib.oracle_set_threads.argtypes = [i]
        _lib.oracle_set_plant.argtypes = [i]
        _lib.oracle_tree_build.argtypes = [p]
        _lib.oracle_tree_build.restype = i
        _lib.oracle_tree_propagate_forces.argtypes = [p, p]
        _lib.oracle_tree_forces.argtypes = [p, d, d, p, i, i]
        _lib.oracle_tree_stokes.argtypes = [p, i, d, p, d, d, p, p, i, i]
        _lib.oracle_tree_collision_counts.argtypes = [p, p, i, p]
        _lib.oracle_direct_collision_counts.argtypes = [p, p, i, d, i, p, i, p]
    return _lib


def _c(a: np.ndarray, dtype=np.float64) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=dtype)


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def num_threads() -> int:
    return int(lib().oracle_num_threads())


def set_plant(plant: int) -> None:
    """Plant a mistake in the tree walks (tests only; 0 restores them): bit 1
    drops the trace term of the Stokeslet dipole, bit 2 the whole dipole,
    bit 4 evaluates the force far field at the node's geometric centre."""
    lib().oracle_set_plant(int(plant))


def set_threads(t: int) -> None:
    lib().oracle_set_threads(int(t))


# ---------------------------------------------------------------- direct


def direct_forces(pos, radii, m_const, L, periodic, targets=None) -> np.ndarray:
    """Float64 dipole forces; ``targets`` = (i0, i1) limits the computed rows."""
    pos = _c(pos).reshape(-1, 3)
    radii = _c(radii)
    n = pos.shape[0]
    i0, i1 = targets if targets is not None else (0, n)
    F = np.zeros((n, 3))
    lib().oracle_direct_forces_range(n, i0, i1, _ptr(pos), _ptr(radii), float(m_const), float(L),
                                     int(bool(periodic)), _ptr(F))
    return F


class Correction:
    """Float64 grids + scalars of the COMSOL correction, in the reference's
    layout (grid_u/v/w of shape (n,n,n), direct...py:87-108)."""

    def __init__(self, grid_u, grid_v, grid_w, grid_min, grid_dx, Fz_comsol, L_ratio, eta_ratio,
                 self_coeff):
        self.u, self.v, self.w = _c(grid_u), _c(grid_v), _c(grid_w)
        self.st = _Corr(_ptr(self.u), _ptr(self.v), _ptr(self.w), self.u.shape[0], 1,
                        float(grid_min), 1.0 / float(grid_dx), 1.0 / float(Fz_comsol),
                        float(L_ratio), float(eta_ratio), float(self_coeff))

    def eval(self, pts) -> np.ndarray:
        pts = _c(pts).reshape(-1, 3)
        out = np.zeros_like(pts)
        lib().oracle_corr_eval(ctypes.byref(self.st), pts.shape[0], _ptr(pts), _ptr(out))
        return out


def direct_convection(pos, forces, eta_const, L, periodic, corr: Correction | None = None,
                      targets=None) -> np.ndarray:
    pos = _c(pos).reshape(-1, 3)
    forces = _c(forces).reshape(-1, 3)
    n = pos.shape[0]
    i0, i1 = targets if targets is not None else (0, n)
    V = np.zeros((n, 3))
    cp = ctypes.byref(corr.st) if corr is not None else None
    lib().oracle_direct_convection_range(n, i0, i1, _ptr(pos), _ptr(forces), float(eta_const),
                                         float(L), int(bool(periodic)), cp, _ptr(V))
    return V


# ---------------------------------------------------------------- hash


def hash_pairs(pos, radii, L, periodic, table_size) -> tuple[bool, np.ndarray]:
    """(is_collision, unique sorted (K,2) int32 pairs) as detect() returns."""
    pos = _c(pos).reshape(-1, 3)
    radii = _c(radii)
    n = pos.shape[0]
    cap = max(16, 4 * n)
    while True:
        out = np.zeros((cap, 2), dtype=np.int32)
        k = lib().oracle_hash_pairs(n, _ptr(pos), _ptr(radii), float(L), int(bool(periodic)),
                                    int(table_size), _ptr(out), cap)
        if k < 0:
            return False, np.empty((0, 2), dtype=np.int32)
        if k <= cap:
            return k > 0, out[:k].copy()
        cap = k


def brute_pairs(pos, radii, L, periodic) -> np.ndarray:
    pos = _c(pos).reshape(-1, 3)
    radii = _c(radii)
    n = pos.shape[0]
    cap = max(16, 4 * n)
    while True:
        out = np.zeros((cap, 2), dtype=np.int32)
        k = lib().oracle_brute_pairs(n, _ptr(pos), _ptr(radii), float(L), int(bool(periodic)),
                                     _ptr(out), cap)
        if k <= cap:
            return out[:k].copy()
        cap = k


def direct_collision_counts(pos, radii, L, periodic, bins, n_bins) -> np.ndarray:
    """Class-pair collision counts over all pairs i < j (DirectCollisionFrequency,
    pbm/collision_frequency/direct_collision_freq.py:15-41); int32 (n_bins, n_bins)."""
    pos = _c(pos).reshape(-1, 3)
    radii = _c(radii)
    bins = _c(bins, np.int32)
    out = np.zeros((n_bins, n_bins), dtype=np.int32)
    lib().oracle_direct_collision_counts(_ptr(pos), _ptr(radii), pos.shape[0], float(L),
                                         int(bool(periodic)), _ptr(bins), int(n_bins), _ptr(out))
    return out


# ---------------------------------------------------------------- merge / Euler


def merge(positions, radii, pairs):
    """Greedy volume-conserving merge (droplet_solution.py:208-287).

    Returns (new_positions, new_radii, resulting_indices).  Pairs are ordered
    by their first index; ties keep lexicographic order (the reference's
    non-stable argsort leaves tie order to numpy's sort)."""
    pairs = np.asarray(pairs).reshape(-1, 2)
    order = np.argsort(pairs[:, 0], kind="stable")
    pairs = pairs[order]
    new_r = np.array(radii, dtype=np.float64, copy=True)
    new_p = np.array(positions, dtype=np.float64, copy=True)
    used: set[int] = set()
    add_r, add_p = [], []
    for a, b in pairs:
        a, b = int(a), int(b)
        if a in used or b in used:
            continue
        r1c = new_r[a] ** 3
        r2c = new_r[b] ** 3
        add_r.append((r1c + r2c) ** (1 / 3))
        add_p.append((new_p[a] * r1c + new_p[b] * r2c) / (r1c + r2c))
        used.add(a)
        used.add(b)
    drop = sorted(used, reverse=True)
    new_r = np.delete(new_r, drop)
    new_p = np.delete(new_p, drop, axis=0)
    new_r = np.append(new_r, add_r)
    new_p = np.vstack([new_p, add_p]) if add_p else new_p
    res = list(range(len(new_r) - len(add_r), len(new_r)))
    return new_p, new_r, res


def euler_step(positions, forces, conv, radii, stokes_factor, dt, L, periodic):
    """One direct-path Euler step (euler_droplet_solver.py:124-131)."""
    x = np.array(positions, dtype=np.float64, copy=True)
    mig = forces / (stokes_factor * radii[:, np.newaxis])
    x += (mig + conv) * dt
    if periodic:
        x %= L
    return x


def euler_step_total(positions, vel, dt, L, periodic):
    """Tree-path Euler step (euler_droplet_solver.py:119, :128-131)."""
    x = np.array(positions, dtype=np.float64, copy=True)
    x += vel * dt
    if periodic:
        x %= L
    return x


# ---------------------------------------------------------------- octree


class _Tree(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int), ("max_nodes", ctypes.c_int),
        ("pos", ctypes.c_void_p), ("r", ctypes.c_void_p),
        ("L", ctypes.c_double), ("periodic", ctypes.c_int), ("mpl", ctypes.c_int),
        ("bmin", ctypes.c_void_p), ("bmax", ctypes.c_void_p), ("r3sum", ctypes.c_void_p),
        ("r3c", ctypes.c_void_p), ("fsum", ctypes.c_void_p), ("fmom", ctypes.c_void_p),
        ("maxr", ctypes.c_void_p), ("first_child", ctypes.c_void_p), ("next", ctypes.c_void_p),
        ("leaf_start", ctypes.c_void_p), ("count", ctypes.c_void_p),
        ("parents", ctypes.c_void_p), ("pfc", ctypes.c_void_p),
        ("leaf_indices", ctypes.c_void_p),
        ("level_offsets", ctypes.c_int * 64), ("num_levels", ctypes.c_int),
        ("node_count", ctypes.c_int), ("parent_count", ctypes.c_int),
        ("leaf_slot_count", ctypes.c_int),
    ]


def max_nodes_for(num_particles: int, mpl: int) -> int:
    """FlatOctree capacity rule (flat_tree.py:63-67)."""
    n = max(num_particles, 64)
    return max(4 * n // max(mpl, 1), 1024)


class OracleTree:
    """The reference's level-by-level octree, restated serially in C
    (oracle/dem_octree_oracle.c), with its force / Stokeslet / total-velocity walks."""

    def __init__(self, pos, radii, L, periodic=False, mpl=1, max_nodes=None):
        self.pos = _c(pos).reshape(-1, 3)
        self.r = _c(radii)
        n = self.pos.shape[0]
        M = int(max_nodes if max_nodes is not None else max_nodes_for(n, mpl))
        self.arr = {
            "bmin": np.zeros(3 * M), "bmax": np.zeros(3 * M), "r3sum": np.zeros(M),
            "r3c": np.zeros(3 * M), "fsum": np.zeros(3 * M), "fmom": np.zeros(9 * M),
            "maxr": np.zeros(M),
            "first_child": np.full(M, -1, np.int32), "next": np.full(M, -1, np.int32),
            "leaf_start": np.full(M, -1, np.int32), "count": np.zeros(M, np.int32),
            "parents": np.zeros(M, np.int32), "pfc": np.zeros(M, np.int32),
            "leaf_indices": np.zeros(max(n, 1), np.int32),
        }
        a = self.arr
        self.st = _Tree(n, M, _ptr(self.pos), _ptr(self.r), float(L), int(bool(periodic)),
                        int(mpl), *(_ptr(a[k]) for k in (
                            "bmin", "bmax", "r3sum", "r3c", "fsum", "fmom", "maxr",
                            "first_child", "next", "leaf_start", "count", "parents", "pfc",
                            "leaf_indices")))
        lib().oracle_tree_build(ctypes.byref(self.st))
        self.node_count = self.st.node_count
        self.parent_count = self.st.parent_count
        self.leaf_slot_count = self.st.leaf_slot_count
        self.level_offsets = [self.st.level_offsets[k] for k in range(self.st.num_levels)]

    @property
    def n(self) -> int:
        return self.pos.shape[0]

    def forces(self, theta, m_const, targets=None) -> np.ndarray:
        i0, i1 = targets if targets is not None else (0, self.n)
        F = np.zeros((self.n, 3))
        lib().oracle_tree_forces(ctypes.byref(self.st), float(theta), float(m_const), _ptr(F),
                                 i0, i1)
        return F

    def _stokes(self, mode, theta, F, eta, stokes, corr, targets):
        F = _c(F).reshape(-1, 3)
        lib().oracle_tree_propagate_forces(ctypes.byref(self.st), _ptr(F))
        i0, i1 = targets if targets is not None else (0, self.n)
        V = np.zeros((self.n, 3))
        cp = ctypes.byref(corr.st) if corr is not None else None
        lib().oracle_tree_stokes(ctypes.byref(self.st), mode, float(theta), _ptr(F), float(eta),
                                 float(stokes), cp, _ptr(V), i0, i1)
        return V

    def stokeslet(self, theta, F, eta, corr=None, targets=None) -> np.ndarray:
        return self._stokes(0, theta, F, eta, 0.0, corr, targets)

    def total_velocity(self, theta, F, eta, stokes, corr=None, targets=None) -> np.ndarray:
        return self._stokes(1, theta, F, eta, stokes, corr, targets)

    def collision_counts(self, bins, n_bins) -> np.ndarray:
        """Class-pair collision counts by the pruned walk
        (TreeCollisionFrequency._count_collisions,
        pbm/collision_frequency/tree_collision_freq.py:62-175); int32 (n_bins, n_bins)."""
        bins = _c(bins, np.int32)
        out = np.zeros((n_bins, n_bins), dtype=np.int32)
        lib().oracle_tree_collision_counts(ctypes.byref(self.st), _ptr(bins), int(n_bins),
                                           _ptr(out))
        return out


def stokes_factor(eta_oil, eta_water) -> float:
    """euler_droplet_solver.py:37-41."""
    return 2 * np.pi * eta_oil * (2 * eta_oil + 3 * eta_water) / (eta_oil + eta_water)


EPS0 = 8.85418781762039e-12


def m_const(eps_oil, E) -> float:
    """direct_droplet_force_calculator.py:30."""
    return 12 * np.pi * EPS0 * eps_oil * E ** 2


def eta_const(eta_oil) -> float:
    """direct_droplet_force_calculator.py:31."""
    return 1 / (8 * np.pi * eta_oil)


__all__ = [
    "build", "lib", "num_threads", "set_threads", "direct_forces", "direct_convection",
    "Correction", "hash_pairs", "brute_pairs", "merge", "euler_step", "euler_step_total",
    "stokes_factor", "m_const", "eta_const", "EPS0", "OracleTree", "max_nodes_for",
]

if os.environ.get("DEM_ORACLE_EAGER_BUILD"):
    build()
