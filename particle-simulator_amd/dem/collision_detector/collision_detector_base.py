"""Base class of contact detectors (reference:
dem/collision_detector/collision_detector_base.py:4-25)."""

import numpy as np  # noqa: F401  (kept for parity with the reference module)


class CollisionDetector:
    """Abstract contact detector."""

    def detect(self, positions, radii, *, L=None, boundary_mode=None):
        """Return (is_collision, collided_pairs) with collided_pairs of shape (K, 2)."""
        raise NotImplementedError
