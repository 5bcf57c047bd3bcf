"""Tree-pruned collision counts between volume classes on the GPU octree
(reference pbm/collision_frequency/tree_collision_freq.py:15-175,
``dem_octree_collision_counts``).

For every droplet i a walk of the built ``FlatOctree`` skips subtrees whose
bounding box is farther than r_i + max_radius(node) and, at the leaves it
reaches, counts every j > i with |d|^2 <= (r_i + r_j)^2.  As in the
reference, the box distance under periodicity replaces a per-axis gap above
L/2 by L - gap, and the positions are those of the tree's last build (the
``positions`` argument of ``compute`` is not used for the walk).
"""

from __future__ import annotations

import numpy as np

from dem import _device

from ..volume_grid import VolumeGrid, sphere_volumes


class TreeCollisionFrequency:
    """Bound to one FlatOctree; ``compute`` uses its current build."""

    def __init__(self, flat_tree, grid: VolumeGrid, max_particles: int) -> None:
        self.tree = flat_tree
        self.grid = grid
        self.n_bins = grid.n_bins
        self.max_particles = int(max_particles)

    def compute_device(self, bins):
        """Device int32 (n_bins, n_bins) counts for host or device int32 classes
        of the tree's particles (original order)."""
        return self.tree.collision_counts_device(bins, self.n_bins)

    def compute(self, positions, radii) -> np.ndarray:
        radii = np.ascontiguousarray(radii, dtype=np.float64)
        n = radii.shape[0]
        if n > self.max_particles:
            raise ValueError(f"{n} droplets exceed max_particles={self.max_particles}")
        if n != self.tree.num_particles[None]:
            raise ValueError(f"{n} radii for a tree built on {self.tree.num_particles[None]} droplets")
        bins = self.grid.bin_indices(sphere_volumes(radii)).astype(np.int32)
        out = self.compute_device(bins)
        return out.cpu().numpy().astype(np.float64)
