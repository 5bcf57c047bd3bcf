"""Barnes-Hut force operator with the direct calculator's interface.

Stands in for the reference's ``TreeDropletForceCalculator``
(dem/octree/force_tree.py:18-244): same constructor keywords, constants and
public attributes (``octree``, ``m_const``, ``eta_const``, ``periodic``,
``boundary_mode``, ``theta``, ``mpl``, ``num_particles``), and the same
observable tree life cycle:

* every operator call builds a fresh tree of the positions it is given,
  except ``calculate_convection`` right after ``calculate``, which walks the
  tree ``calculate`` built (reference :162-180) — so a caller that moves the
  droplets between the two calls sees the old tree, as with the reference;
* a call with more droplets than the tree holds replaces it with a larger
  one and re-loads the COMSOL correction into it (:92-106);
* ``calculate_forces_and_total_velocity`` exists, which is what makes the
  solver take the one-walk total-velocity path (euler_droplet_solver.py:
  113-119).
The walks themselves run on the GPU (``FlatOctree``).
"""

from __future__ import annotations

import numpy as np

from ..force_calculator.force_calculator_base import ForceCalculator
from .flat_tree import FlatOctree

EPS0 = 8.85418781762039e-12


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


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
        self.eps0 = EPS0
        self.eps_oil, self.eta_oil, self.eta_water, self.E = eps_oil, eta_oil, eta_water, E
        self.L = L
        self.periodic = periodic
        self.boundary_mode = "periodic" if periodic else "open"
        self.m_const = 12.0 * np.pi * EPS0 * eps_oil * E**2  # dipole force constant
        self.eta_const = 1.0 / (8.0 * np.pi * eta_oil)      # Oseen prefactor
        self.theta, self.mpl = theta, mpl
        self.num_particles = num_particles
        self.correction_grid_resolution = correction_grid_resolution
        self.octree = self._new_tree(num_particles)
        self._reusable = False  # the tree was built by calculate() for calculate_convection
        self._correction = None  # (correction, L_sim) to load into a replacement tree

    def _new_tree(self, capacity: int) -> FlatOctree:
        return FlatOctree(theta=self.theta, mpl=self.mpl, num_particles=capacity,
                          correction_grid_resolution=self.correction_grid_resolution)

    def _fit(self, n: int) -> None:
        """Replace the tree by one that holds n droplets when it is too small."""
        if n <= self.num_particles:
            return
        self.octree = self._new_tree(n)
        self.num_particles = n
        if self._correction is not None:
            self.octree.load_periodic_correction(*self._correction, eta_sim=self.eta_oil)

    def _build(self, positions, radii, L, periodic) -> None:
        positions, radii = _f64(positions), _f64(radii)
        self._fit(positions.shape[0])
        self.octree.build(positions, radii, self.L if L is None else L,
                          self.periodic if periodic is None else periodic)

    def update_params(self, theta: float, mpl: int) -> None:
        self.theta, self.mpl = theta, mpl
        self.octree.update_params(theta, mpl)
        self._reusable = False

    def load_periodic_correction(self, correction, L_sim: float):
        self._correction = (correction, L_sim)
        self.octree.load_periodic_correction(correction, L_sim, eta_sim=self.eta_oil)

    # -- the reference's operator methods (numpy in, numpy out)

    def calculate(self, positions, radii, *, L=None, periodic=None) -> np.ndarray:
        self._build(positions, radii, L, periodic)
        self._reusable = True
        return self.octree.compute_forces(self.m_const)

    def calculate_convection(self, positions, radii, forces, *, L=None, periodic=None) -> np.ndarray:
        if not self._reusable:
            self._build(positions, radii, L, periodic)
        self._reusable = False
        return self.octree.compute_stokeslet(_f64(forces), self.eta_const)

    def calculate_forces_and_convection(self, positions, radii, *, L=None, periodic=None) -> tuple:
        self._build(positions, radii, L, periodic)
        self._reusable = False
        return self.octree.compute_forces_and_stokeslet(self.m_const, self.eta_const)

    def calculate_forces_and_total_velocity(self, positions, radii, stokes_factor, *, L=None,
                                            periodic=None) -> tuple:
        self._build(positions, radii, L, periodic)
        self._reusable = False
        return self.octree.compute_forces_and_total_velocity(self.m_const, self.eta_const,
                                                             stokes_factor)

    # -- device tensors in and out (no host copies)

    def forces_and_total_velocity_device(self, pos_d, radii_d, stokes_factor, *, L=None,
                                         periodic=None):
        """Build + force walk + total-velocity walk on device arrays."""
        self._fit(int(pos_d.shape[0]))
        self.octree.build_device(pos_d, radii_d, self.L if L is None else L,
                                 self.periodic if periodic is None else periodic)
        self._reusable = False
        F = self.octree.forces_device(self.m_const)
        return F, self.octree.total_velocity_device(F, self.eta_const, stokes_factor)


__all__ = ["TreeDropletForceCalculator"]
