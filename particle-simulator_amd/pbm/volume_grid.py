"""Volume classes for the population balance (reference pbm/volume_grid.py:11-128).

Edges are ``n_bins + 1`` volumes spaced geometrically (``np.geomspace``),
logarithmically (powers of ten of ``np.linspace`` of the exponents) or
linearly; centres are geometric means for the first two and midpoints for the
linear grid.  A volume v belongs to class ``searchsorted(edges, v, 'right') - 1``
clamped to [0, n_bins - 1]: volumes at or above v_max land in the last class,
below v_min in the first, with one warning per grid for each side.  The edge
formulas are the reference's own (the classes are defined by them), so class
indices agree bit for bit.
"""

from __future__ import annotations

import logging

import numpy as np

logger = logging.getLogger(__name__)


def sphere_volumes(radii) -> np.ndarray:
    """(4/3) pi r^3 evaluated as the reference's collision counters do
    (pbm/collision_frequency/tree_collision_freq.py:51)."""
    radii = np.asarray(radii, dtype=np.float64)
    return (4.0 / 3.0) * np.pi * radii**3


class VolumeGrid:
    """Discretisation of droplet volume into ``n_bins`` classes."""

    def __init__(self, v_min: float, v_max: float, n_bins: int, spacing: str = "geometric") -> None:
        if v_min <= 0 or v_max <= v_min:
            raise ValueError("VolumeGrid requires 0 < v_min < v_max")
        if n_bins < 2:
            raise ValueError("VolumeGrid requires n_bins >= 2")
        self.n_bins = n_bins
        self.spacing = spacing
        if spacing == "geometric":
            self.edges = np.geomspace(v_min, v_max, n_bins + 1)
        elif spacing == "logarithmic":
            self.edges = 10 ** np.linspace(np.log10(v_min), np.log10(v_max), n_bins + 1)
        elif spacing == "linear":
            self.edges = np.linspace(v_min, v_max, n_bins + 1)
        else:
            raise ValueError(f"unknown spacing: {spacing}")
        lo, hi = self.edges[:-1], self.edges[1:]
        if spacing == "linear":
            self.centers = 0.5 * (lo + hi)
        else:
            self.centers = np.sqrt(lo * hi)
        self.widths = np.diff(self.edges)
        self._warned_low = False
        self._warned_high = False

    @classmethod
    def from_radii_range(cls, r_min: float, r_max: float, n_bins: int,
                         spacing: str = "geometric") -> "VolumeGrid":
        v = sphere_volumes([r_min, r_max])
        return cls(float(v[0]), float(v[1]), n_bins, spacing)

    def _warn(self, n_low: int, n_high: int, value: float | None = None) -> None:
        if n_high and not self._warned_high:
            logger.warning("VolumeGrid: %d volume(s) above v_max=%.3e clamped into class %d",
                           n_high, self.edges[-1], self.n_bins - 1)
            self._warned_high = True
        if n_low and not self._warned_low:
            logger.warning("VolumeGrid: %d volume(s) below v_min=%.3e clamped into class 0",
                           n_low, self.edges[0])
            self._warned_low = True

    def bin_index(self, volume: float) -> int:
        raw = int(np.searchsorted(self.edges, volume, side="right")) - 1
        self._warn(int(raw < 0), int(raw >= self.n_bins))
        return max(0, min(raw, self.n_bins - 1))

    def bin_indices(self, volumes) -> np.ndarray:
        raw = np.searchsorted(self.edges, volumes, side="right").astype(np.intp) - 1
        self._warn(int(np.sum(raw < 0)), int(np.sum(raw >= self.n_bins)))
        np.clip(raw, 0, self.n_bins - 1, out=raw)
        return raw

    def histogram(self, radii) -> np.ndarray:
        """Droplet count per class; volumes outside [v_min, v_max] go to the
        end classes (np.histogram alone would drop them)."""
        v = sphere_volumes(radii)
        counts, _ = np.histogram(v, bins=self.edges)
        n_high = int(np.sum(v > self.edges[-1]))
        n_low = int(np.sum(v < self.edges[0]))
        counts[-1] += n_high
        counts[0] += n_low
        self._warn(n_low, n_high)
        return counts.astype(np.float64)
