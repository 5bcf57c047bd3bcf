"""Flat Barnes-Hut octree on the GPU (drop-in for the reference's dem.octree).

Exports the reference's names (dem/octree/__init__.py:16-34): OctreeNode,
FlatOctree, TreeDropletForceCalculator, TreeStatistics, compute_tree_stats,
print_tree_stats, visualize_tree.
"""

from .flat_tree import FlatOctree
from .force_tree import TreeDropletForceCalculator
from .octree_node import OctreeNode
from .tree_stats import TreeStatistics, compute_tree_stats, print_tree_stats
from .tree_visualizer import visualize_tree

__all__ = [
    "OctreeNode",
    "FlatOctree",
    "TreeDropletForceCalculator",
    "TreeStatistics",
    "compute_tree_stats",
    "print_tree_stats",
    "visualize_tree",
]
