"""Barnes-Hut force operator with the direct calculator's interface.

Drop-in for the reference's ``TreeDropletForceCalculator``
(dem/octree/force_tree.py:18-244): the same constructor keywords and
constants, the tree rebuilt on every call, the stale-tree bookkeeping of
``calculate`` / ``calculate_convection`` (:162-180), ``_ensure_octree_capacity``
re-creating a larger tree (and re-loading the correction) when N grows
(:92-106), and ``calculate_forces_and_total_velocity`` — the method whose
presence makes ``EulerDropletSolver`` take the one-walk total-velocity path
(euler_droplet_solver.py:113-119).
"""

from __future__ import annotations

import numpy as np

from .. import _device
from ..force_calculator.force_calculator_base import ForceCalculator
from .flat_tree import FlatOctree


class TreeDropletForceCalculator(ForceCalculator):
    def __init__(
        self,
        num_particles: int = 10000,
        theta: float = 0.5,
        mpl: int = 1,
        eps_oil: float = 2.85,
        eta_oil: float = 0.065,
        eta_water: float = 0.001,
        E: float = 3e5,
        L: float = 1.0,
        periodic: bool = False,
        correction_grid_resolution: int = 0,
    ):
        self.eps0 = 8.85418781762039e-12
        self.eps_oil = eps_oil
        self.eta_oil = eta_oil
        self.eta_water = eta_water
        self.E = E
        self.L = L
        self.periodic = periodic
        self.boundary_mode = "periodic" if periodic else "open"
        self.m_const = 12.0 * np.pi * self.eps0 * eps_oil * E**2
        self.eta_const = 1.0 / (8.0 * np.pi * eta_oil)
        self.num_particles = num_particles
        self.theta = theta
        self.mpl = mpl
        self.correction_grid_resolution = correction_grid_resolution
        self.octree = FlatOctree(theta=theta, mpl=mpl, num_particles=num_particles,
                                 correction_grid_resolution=correction_grid_resolution)
        self._tree_is_stale = True
        self._forces_in_fields = False
        self._correction = None
        self._correction_L_sim = None

    def update_params(self, theta: float, mpl: int) -> None:
        self.theta = theta
        self.mpl = mpl
        self.octree.update_params(theta, mpl)
        self._tree_is_stale = True

    @staticmethod
    def _to_contiguous(arr: np.ndarray, dtype=np.float64) -> np.ndarray:
        return np.ascontiguousarray(arr, dtype=dtype)

    def _ensure_octree_capacity(self, n: int) -> None:
        if n > self.num_particles:
            self.octree = FlatOctree(theta=self.theta, mpl=self.mpl, num_particles=n,
                                     correction_grid_resolution=self.correction_grid_resolution)
            self.num_particles = n
            if self._correction is not None:
                self.octree.load_periodic_correction(self._correction, self._correction_L_sim,
                                                     eta_sim=self.eta_oil)

    def load_periodic_correction(self, correction, L_sim: float):
        self._correction = correction
        self._correction_L_sim = L_sim
        self.octree.load_periodic_correction(correction, L_sim, eta_sim=self.eta_oil)

    def _resolve(self, L, periodic):
        return (L if L is not None else self.L,
                periodic if periodic is not None else self.periodic)

    def calculate(self, positions, radii, *, L=None, periodic=None) -> np.ndarray:
        positions = self._to_contiguous(positions)
        radii = self._to_contiguous(radii)
        use_L, use_periodic = self._resolve(L, periodic)
        self._ensure_octree_capacity(positions.shape[0])
        self.octree.build(positions, radii, use_L, use_periodic)
        self._tree_is_stale = False
        self._forces_in_fields = True
        return self.octree.compute_forces(self.m_const)

    def calculate_convection(self, positions, radii, forces, *, L=None, periodic=None) -> np.ndarray:
        forces = self._to_contiguous(forces)
        use_L, use_periodic = self._resolve(L, periodic)
        if self._tree_is_stale:
            positions = self._to_contiguous(positions)
            radii = self._to_contiguous(radii)
            self._ensure_octree_capacity(positions.shape[0])
            self.octree.build(positions, radii, use_L, use_periodic)
            self._tree_is_stale = False
        if self._forces_in_fields:
            velocities = self.octree.compute_stokeslet_from_fields(forces, self.eta_const)
        else:
            velocities = self.octree.compute_stokeslet(forces, self.eta_const)
        self._tree_is_stale = True
        self._forces_in_fields = False
        return velocities

    def calculate_forces_and_convection(self, positions, radii, *, L=None, periodic=None) -> tuple:
        positions = self._to_contiguous(positions)
        radii = self._to_contiguous(radii)
        use_L, use_periodic = self._resolve(L, periodic)
        self._ensure_octree_capacity(positions.shape[0])
        self.octree.build(positions, radii, use_L, use_periodic)
        self._tree_is_stale = True
        return self.octree.compute_forces_and_stokeslet(self.m_const, self.eta_const)

    def calculate_forces_and_total_velocity(self, positions, radii, stokes_factor, *, L=None,
                                            periodic=None) -> tuple:
        positions = self._to_contiguous(positions)
        radii = self._to_contiguous(radii)
        use_L, use_periodic = self._resolve(L, periodic)
        self._ensure_octree_capacity(positions.shape[0])
        self.octree.build(positions, radii, use_L, use_periodic)
        self._tree_is_stale = True
        return self.octree.compute_forces_and_total_velocity(self.m_const, self.eta_const,
                                                             stokes_factor)

    def forces_and_total_velocity_device(self, pos_d, radii_d, stokes_factor, *, L=None,
                                         periodic=None):
        """Device tensors in and out (no host copies): build + both walks."""
        use_L, use_periodic = self._resolve(L, periodic)
        self._ensure_octree_capacity(int(pos_d.shape[0]))
        self.octree.build_device(pos_d, radii_d, use_L, use_periodic)
        F = self.octree.forces_device(self.m_const)
        V = self.octree.total_velocity_device(F, self.eta_const, stokes_factor)
        return F, V


__all__ = ["TreeDropletForceCalculator", "_device"]
