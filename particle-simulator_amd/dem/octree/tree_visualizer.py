"""Octree wireframe view (outside the accelerated path).

The reference's ``visualize_tree`` (dem/octree/tree_visualizer.py:106-253)
draws the tree with PyVista and raises ImportError without it, which stops
``main_tree.py`` (it calls the function unconditionally, :130-133).  PyVista
is not part of this build: the function validates its arguments, reports what
it would draw, and returns None.
"""

from __future__ import annotations

import warnings

import numpy as np


def visualize_tree(octree, positions: np.ndarray | None = None, radii: np.ndarray | None = None,
                   color_by: str = "depth", max_level: int | None = None,
                   show_particles: bool = True, opacity: float = 0.3) -> None:
    if color_by not in ("depth", "fill"):
        raise ValueError(f"color_by must be 'depth' or 'fill', got {color_by!r}")
    warnings.warn(
        f"visualize_tree: no 3D viewer in this build; tree has {int(octree.node_count[None])} "
        f"nodes over {int(octree.num_particles[None])} particles",
        RuntimeWarning,
        stacklevel=2,
    )
    return None
