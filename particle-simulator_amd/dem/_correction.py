"""Device copy of the COMSOL periodic-image correction grid.

Mirrors what the reference's operators copy into Taichi fields in
``load_periodic_correction`` (dem/force_calculator/direct_droplet_force_calculator.py:74-120,
dem/octree/flat_tree.py:141-167): the three regular grids, their spacing and
origin, and the scalars Fz_inv, L_ratio, eta_ratio and the self coefficient
w(0,0,0)/Fz * eta_ratio * L_ratio.  The grids are interleaved as (u, v, w, 0)
per node so one 16-byte load fetches a node for all three components.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _device
from ._native import DemCorrection


class DeviceCorrection:
    def __init__(self, correction, L_sim: float, eta_ratio: float):
        gd = correction.get_grid_data()
        n = int(gd["grid_resolution"])
        inter = np.zeros((n, n, n, 4), dtype=np.float64)
        inter[..., 0] = gd["grid_u"]
        inter[..., 1] = gd["grid_v"]
        inter[..., 2] = gd["grid_w"]
        self.grid64 = _device.to_device(inter.reshape(-1), np.float64)
        self.grid32 = _device.to_device(inter.reshape(-1), np.float32)
        self.n = n
        self.grid_min = float(gd["grid_min"])
        self.grid_dx = float(gd["grid_dx"])
        self.Fz_comsol = float(gd["Fz_comsol"])
        self.L_ratio = float(gd["L_comsol"]) / float(L_sim)
        self.eta_ratio = float(eta_ratio)
        w_self = float(correction.evaluate(np.array([[0.0, 0.0, 0.0]]))[0, 2])
        self.self_coeff = w_self / self.Fz_comsol * self.eta_ratio * self.L_ratio
        self.struct = DemCorrection(
            ctypes.c_void_p(self.grid32.data_ptr()),
            ctypes.c_void_p(self.grid64.data_ptr()),
            n,
            1,
            self.grid_min,
            1.0 / self.grid_dx,
            1.0 / self.Fz_comsol,
            self.L_ratio,
            self.eta_ratio,
            self.self_coeff,
        )

    def ref(self):
        return ctypes.byref(self.struct)
