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
        """Indices of droplets overlapping an earlier one (open box)."""
        n = len(radii)
        bad = np.zeros(n, dtype=bool)
        cell = max(2 * float(np.max(radii)) + self.minimum_distance, 1e-30)
        keys = np.floor((pos - self.coord_range[0]) / cell).astype(np.int64)
        buckets: dict[tuple, list[int]] = {}
        for i in range(n):
            k = tuple(keys[i])
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dz in (-1, 0, 1):
                        for j in buckets.get((k[0] + dx, k[1] + dy, k[2] + dz), ()):
                            if np.linalg.norm(pos[i] - pos[j]) < radii[i] + radii[j] + self.minimum_distance:
                                bad[i] = True
            buckets.setdefault(k, []).append(i)
        return np.nonzero(bad)[0]

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
