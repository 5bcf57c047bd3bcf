"""Uniform droplet generator (host initialisation, not on the hot path).

Counterpart of the reference's ``UniformDropletGenerator``
(dem/particle_generator/uniform_droplet_generator.py:11-284) so that
``from dem.particle_generator import *`` in the reference's scripts works:
uniform radii in ``radii_range`` and positions in ``coord_range``^3, then
overlapping droplets (gap < minimum_distance) are re-seeded until none
overlap; gives up with ValueError after 1000 rounds (:77-80).  Re-seeding
uses numpy's generator (seedable) instead of the reference's ``ti.random``.
"""

from __future__ import annotations

import numpy as np

from ..particle_state import DropletState
from .particle_generator_base import ParticleGenerator


class UniformDropletGenerator(ParticleGenerator):
    def __init__(self, coord_range=(0, 10), radii_range=(1, 100), num_particles=1,
                 minimum_distance=0, seed=None):
        self.num_particles = num_particles
        self.coord_range = coord_range
        self.radii_range = radii_range
        self.minimum_distance = minimum_distance
        self.droplet_state = None
        self._rng = np.random.default_rng(seed)

    def _overlapping(self, pos, radii):
        """Indices of droplets closer than r_i + r_j + minimum_distance to an
        earlier droplet (open box)."""
        from scipy.spatial import cKDTree

        reach = 2 * float(np.max(radii)) + self.minimum_distance
        pairs = cKDTree(pos).query_pairs(reach, output_type="ndarray")
        if len(pairs) == 0:
            return np.empty(0, dtype=np.int64)
        d = np.linalg.norm(pos[pairs[:, 0]] - pos[pairs[:, 1]], axis=1)
        close = d < radii[pairs[:, 0]] + radii[pairs[:, 1]] + self.minimum_distance
        return np.unique(pairs[close].max(axis=1))

    def generate(self):
        lo, hi = self.coord_range
        n = self.num_particles
        radii = self._rng.uniform(self.radii_range[0], self.radii_range[1], size=n)
        pos = self._rng.uniform(lo, hi, size=(n, 3))
        for _ in range(1000):
            bad = self._overlapping(pos, radii)
            if len(bad) == 0:
                self.droplet_state = DropletState(pos, radii, 0)
                return self.droplet_state
            pos[bad] = self._rng.uniform(lo, hi, size=(len(bad), 3))
        raise ValueError("could not remove overlaps while generating droplets")

    def print(self):
        if self.droplet_state is None:
            print("no droplets generated yet")
        else:
            print(self.droplet_state)

    def plot(self):
        raise NotImplementedError("plotting is outside the accelerated hot path")
