"""Flat Barnes-Hut octree, built and walked on the GPU.

Drop-in for the reference's ``FlatOctree`` (dem/octree/flat_tree.py:35-1630):
same constructor (theta, mpl, num_particles, correction_grid_resolution),
the same capacity rule max_nodes = max(4*max(N,64)//max(mpl,1), 1024)
(:63-67), the same public methods, and Taichi-field-like views of the node
arrays (``nodes.count.to_numpy()``, ``node_count[None]``, ``theta_sq[None]``,
``_parent_level_offsets``, ...) that tests, ``tree_stats``, the visualiser and
the PBM counter read.  Build and walks are the libdemhip kernels
(``dem_octree_*``): the node numbering, leaf slots, next pointers and
aggregates follow the reference's definitions (so the exposed arrays match
it), and the opening test of every walk is evaluated in float64 exactly as
the reference writes it.

Deliberate differences: node splitting stops at depth 21 (the reference's
loop allows 40 levels; only coincident particles get that deep), and the
particles of a leaf are listed in Morton order (the reference fills leaves
with an atomic race).
"""

from __future__ import annotations

import ctypes
import math

import numpy as np

from .. import _device, _native
from .._correction import DeviceCorrection
from .._native import DemTreeInfo
from .octree_node import DeviceField, OctreeNode, ScalarField

# Force unit of the tree's internal FP32 records: m_const of the reference's
# default physics (12 pi eps0 * 2.85 * (3e5)^2); any value of this order works.
_FORCE_UNIT_HINT = 12 * np.pi * 8.85418781762039e-12 * 2.85 * (3e5) ** 2


class _ParticleField:
    """Capacity-sized view of the particle arrays last given to ``build``."""

    def __init__(self, tree, key, width):
        self._tree, self._key, self._width = tree, key, width

    def to_numpy(self) -> np.ndarray:
        cap = self._tree._capacity
        out = np.zeros((cap, self._width) if self._width > 1 else cap)
        src = self._tree._host_particles.get(self._key)
        if src is not None:
            out[: len(src)] = src
        return out

    def __getitem__(self, idx):
        return self.to_numpy()[idx]


class FlatOctree:
    def __init__(self, theta: float = 0.5, mpl: int = 1, num_particles: int = 10000,
                 correction_grid_resolution: int = 0):
        self.theta = theta
        self.mpl = mpl
        self.theta_sq = ScalarField(theta * theta, float)
        self.mpl_field = ScalarField(mpl, int)
        n = max(num_particles, 64)
        self.max_nodes = max(4 * n // max(mpl, 1), 1024)
        self.max_leaf_slots = n
        self._capacity = n
        self.nodes = OctreeNode(self, self.max_nodes)
        self.leaf_indices = DeviceField(self, "leaf_indices", np.int32, length="capacity")
        self.parents = DeviceField(self, "parents", np.int32)
        self.parent_first_child = DeviceField(self, "parent_first_child", np.int32)
        self.particle_positions = _ParticleField(self, "positions", 3)
        self.particle_radii = _ParticleField(self, "radii", 1)
        self.particle_forces = DeviceField(self, "particle_forces", np.float64, 3,
                                           length="capacity")
        self.node_count = ScalarField(0, int)
        self.parent_count = ScalarField(0, int)
        self.leaf_slot_count = ScalarField(0, int)
        self.num_particles = ScalarField(0, int)
        self.L = ScalarField(0.0, float)
        self.periodic = ScalarField(0, int)
        self._parent_level_offsets: list = []
        self.correction_grid_resolution = correction_grid_resolution
        self.corr_enabled = ScalarField(0, int)
        self._corr: DeviceCorrection | None = None
        self._info = DemTreeInfo()
        self._info.n_capacity = n
        self._info.max_nodes = self.max_nodes
        self._ws = None
        self._host_particles: dict = {}
        self._dev = {}

    # ------------------------------------------------------------------ plumbing

    def _workspace(self):
        if self._ws is None:
            t = _device.torch()
            nbytes = _native.load().dem_octree_workspace_bytes(self._capacity, self.max_nodes)
            self._ws = t.zeros(max(int(nbytes), 256), dtype=t.uint8, device=_device.device())
        return self._ws

    def _field_length(self, which: str) -> int:
        return self._capacity if which == "capacity" else self.max_nodes

    def _read_field(self, name: str, dtype) -> np.ndarray:
        t = _device.torch()
        if self._ws is None:
            width = {"min": 3, "max": 3, "R3_c": 3, "force_sum": 3, "fmom": 9,
                     "particle_forces": 3}.get(name, 1)
            n = self._capacity if name in ("leaf_indices", "perm", "particle_forces") else self.max_nodes
            fill = -1 if name in ("first_child", "next", "leaf_start") else 0
            return np.full(n * width, fill, dtype=dtype)
        off, nb = ctypes.c_size_t(0), ctypes.c_size_t(0)
        _native.call("dem_octree_field", ctypes.byref(self._info), name.encode(), ctypes.byref(off),
                     ctypes.byref(nb))
        tt = {np.dtype(np.float64): t.float64, np.dtype(np.int32): t.int32}[np.dtype(dtype)]
        view = self._ws[off.value: off.value + nb.value].view(tt)
        return view.cpu().numpy()

    def _call(self, name, *args):
        ws = self._workspace()
        return _native.call(name, *args, _device.ptr(ws), ws.numel(), _device.stream())

    # ------------------------------------------------------------------ parameters

    def load_periodic_correction(self, correction, L_sim: float, eta_sim: float = None):
        """Copy the COMSOL grids to the device (reference :141-167).  As in the
        reference, eta_ratio is eta_comsol / eta_sim, or 1 when eta_sim is None."""
        if self.correction_grid_resolution == 0:
            raise RuntimeError("correction_grid_resolution=0: no correction storage was reserved.")
        gd = correction.get_grid_data()
        eta_ratio = gd["eta_comsol"] / eta_sim if eta_sim is not None else 1.0
        self._corr = DeviceCorrection(correction, L_sim, eta_ratio)
        self.corr_enabled[None] = 1

    def update_params(self, theta: float, mpl: int) -> None:
        self.theta = theta
        self.mpl = mpl
        self.theta_sq[None] = theta * theta
        self.mpl_field[None] = mpl

    # ------------------------------------------------------------------ build

    def build(self, positions: np.ndarray, radii: np.ndarray, L: float, periodic: bool = False):
        """Build the tree of [0, L]^3 (reference :180-246)."""
        n = positions.shape[0]
        if n > self._capacity:
            raise ValueError(
                f"Number of particles ({n}) exceeds allocated capacity ({self._capacity}). "
                f"Recreate FlatOctree with larger num_particles."
            )
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
        rad = np.ascontiguousarray(radii, dtype=np.float64)
        self._host_particles = {"positions": pos, "radii": rad}
        self.build_device(_device.to_device(pos), _device.to_device(rad), L, periodic)

    def build_device(self, pos_d, radii_d, L: float, periodic: bool = False):
        n = int(pos_d.shape[0])
        if n > self._capacity:
            raise ValueError(
                f"Number of particles ({n}) exceeds allocated capacity ({self._capacity}). "
                f"Recreate FlatOctree with larger num_particles."
            )
        self._dev = {"pos": pos_d, "radii": radii_d}
        info = self._info
        self._call("dem_octree_build", _device.ptr(pos_d), _device.ptr(radii_d), n,
                   ctypes.c_double(float(L)), 1 if periodic else 0, int(self.mpl_field[None]),
                   ctypes.c_double(_FORCE_UNIT_HINT), ctypes.byref(info))
        self.node_count[None] = info.node_count
        self.parent_count[None] = info.parent_count
        self.leaf_slot_count[None] = info.leaf_slot_count
        self.num_particles[None] = n
        self.L[None] = float(L)
        self.periodic[None] = 1 if periodic else 0
        self._parent_level_offsets = [int(info.level_offsets[k]) for k in range(info.num_levels)]

    # ------------------------------------------------------------------ device walks

    def forces_device(self, m_const: float):
        n = self.num_particles[None]
        F = _device.empty((n, 3))
        if n:
            self._call("dem_octree_forces", ctypes.byref(self._info),
                       ctypes.c_double(self.theta_sq[None]), ctypes.c_double(float(m_const)),
                       _device.ptr(F))
        return F

    def _stokes_device(self, F_d, eta_const: float, stokes_factor: float | None):
        n = self.num_particles[None]
        V = _device.empty((n, 3))
        if n == 0:
            return V
        corr = self._corr.ref() if self._corr is not None else None
        if stokes_factor is None:
            self._call("dem_octree_stokeslet", ctypes.byref(self._info),
                       ctypes.c_double(self.theta_sq[None]), _device.ptr(F_d),
                       ctypes.c_double(float(eta_const)), corr, _device.ptr(V))
        else:
            self._call("dem_octree_total_velocity", ctypes.byref(self._info),
                       ctypes.c_double(self.theta_sq[None]), _device.ptr(F_d),
                       ctypes.c_double(float(eta_const)), ctypes.c_double(float(stokes_factor)),
                       corr, _device.ptr(V))
        return V

    def stokeslet_device(self, F_d, eta_const: float):
        return self._stokes_device(F_d, eta_const, None)

    def total_velocity_device(self, F_d, eta_const: float, stokes_factor: float):
        return self._stokes_device(F_d, eta_const, stokes_factor)

    def collision_counts_device(self, bins, n_bins: int):
        """Class-pair collision counts by the pruned walk of this build
        (``dem_octree_collision_counts``): device int32 (n_bins, n_bins)."""
        t = _device.torch()
        n = self.num_particles[None]
        bins_d = bins if t.is_tensor(bins) else t.as_tensor(
            np.ascontiguousarray(bins, dtype=np.int32), device=_device.device())
        if int(bins_d.numel()) < n:
            raise ValueError(f"{int(bins_d.numel())} classes for {n} droplets")
        out = t.empty((n_bins, n_bins), dtype=t.int32, device=_device.device())
        self._call("dem_octree_collision_counts", ctypes.byref(self._info), _device.ptr(bins_d),
                   int(n_bins), _device.ptr(out))
        return out

    # ------------------------------------------------------------------ reference API

    def compute_forces(self, m_const: float) -> np.ndarray:
        """Dipole forces (reference :896-906), m_const = 12 pi eps0 eps_oil E^2."""
        return self.forces_device(m_const).cpu().numpy()

    def _forces_in(self, forces):
        n = self.num_particles[None]
        return _device.to_device(np.ascontiguousarray(np.asarray(forces)[:n], dtype=np.float64))

    def compute_stokeslet(self, forces: np.ndarray, eta_const: float) -> np.ndarray:
        """Stokeslet velocities from `forces` (reference :1282-1294)."""
        return self.stokeslet_device(self._forces_in(forces), eta_const).cpu().numpy()

    def compute_stokeslet_from_fields(self, forces: np.ndarray, eta_const: float) -> np.ndarray:
        """As compute_stokeslet, after compute_forces (reference :1296-1312)."""
        return self.compute_stokeslet(forces, eta_const)

    def compute_forces_and_stokeslet(self, m_const: float, eta_const: float) -> tuple:
        F = self.forces_device(m_const)
        V = self.stokeslet_device(F, eta_const)
        return F.cpu().numpy(), V.cpu().numpy()

    def compute_total_velocity(self, forces: np.ndarray, eta_const: float,
                               stokes_factor: float) -> np.ndarray:
        """Migration + convection in one walk (reference :1590-1604)."""
        return self.total_velocity_device(self._forces_in(forces), eta_const,
                                          stokes_factor).cpu().numpy()

    def compute_forces_and_total_velocity(self, m_const: float, eta_const: float,
                                          stokes_factor: float) -> tuple:
        F = self.forces_device(m_const)
        V = self.total_velocity_device(F, eta_const, stokes_factor)
        return F.cpu().numpy(), V.cpu().numpy()

    def get_tree_stats(self) -> dict:
        nc = self.node_count[None]
        return {
            "node_count": nc,
            "max_nodes": self.max_nodes,
            "node_utilization": nc / self.max_nodes if self.max_nodes > 0 else 0.0,
            "parent_count": self.parent_count[None],
            "num_particles": self.num_particles[None],
            "leaf_slot_count": self.leaf_slot_count[None],
        }

    @property
    def theta_value(self) -> float:
        return math.sqrt(self.theta_sq[None])
