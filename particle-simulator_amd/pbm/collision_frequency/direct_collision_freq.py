"""All-pairs collision counts between volume classes on the GPU
(reference pbm/collision_frequency/direct_collision_freq.py:9-41, an O(N^2)
Python loop; here ``dem_direct_collision_counts``)."""

from __future__ import annotations

import ctypes

import numpy as np

from dem import _device, _native

from ..volume_grid import VolumeGrid, sphere_volumes


class DirectCollisionFrequency:
    """counts[a][b] = number of touching pairs (|d| <= r_i + r_j) between
    classes a and b, each pair counted in both [a][b] and [b][a]."""

    def __init__(self, grid: VolumeGrid) -> None:
        self.grid = grid

    def compute_device(self, pos_d, radii_d, bins, L: float = 0.0, periodic: bool = False):
        """Device int32 (n_bins, n_bins) counts for device positions / radii and
        host or device int32 classes."""
        t = _device.torch()
        n = int(radii_d.shape[0])
        nb = self.grid.n_bins
        bins_d = bins if t.is_tensor(bins) else t.as_tensor(
            np.ascontiguousarray(bins, dtype=np.int32), device=_device.device())
        out = t.empty((nb, nb), dtype=t.int32, device=_device.device())
        _native.call("dem_direct_collision_counts", _device.ptr(pos_d), _device.ptr(radii_d), n,
                     ctypes.c_double(float(L)), 1 if periodic else 0, _device.ptr(bins_d), nb,
                     _device.ptr(out), _device.stream())
        return out

    def compute(self, positions, radii, L: float = 0.0, periodic: bool = False) -> np.ndarray:
        positions = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
        radii = np.ascontiguousarray(radii, dtype=np.float64)
        bins = self.grid.bin_indices(sphere_volumes(radii)).astype(np.int32)
        out = self.compute_device(_device.to_device(positions), _device.to_device(radii), bins,
                                  L, periodic)
        return out.cpu().numpy().astype(np.float64)
