"""Direct O(N^2) dipole force + Oseen convection on gfx950.

Drop-in for the reference's ``DirectDropletForceCalculator``
(dem/force_calculator/direct_droplet_force_calculator.py:7-527): same
constructor keywords and defaults (:10-22), the same constants
m_const = 12 pi eps0 eps_oil E^2 and eta_const = 1/(8 pi eta_oil) (:23-31),
the same public methods and return shapes.  The arithmetic runs in the
fused two-phase HIP kernel of libdemhip (``dem_direct_forces``,
``dem_direct_convection``, ``dem_direct_forces_and_convection``): FP32 VALU on
32-bit fixed-point positions, FP64 accumulation across tiles, float64 in and
out.  Inputs are copied host->device and results device->host on every call,
exactly where the reference copies numpy <-> Taichi fields (:260-305).

Differences from the reference, by design:
* any N is accepted (the reference silently reads past its Taichi fields when
  N exceeds the constructor capacity, :39-48 / :261-272);
* coincident distinct particles contribute zero to the Oseen sum instead of
  inf/NaN (the reference divides by |x| = 0).
"""

from __future__ import annotations

import numpy as np

from .. import _device, _native
from .._correction import DeviceCorrection
from .._native import DemPhys
from .force_calculator_base import ForceCalculator


class DirectDropletForceCalculator(ForceCalculator):
    def __init__(
        self,
        num_particles=10000,
        eps_oil=2.85,
        eta_oil=0.065,
        eta_water=0.001,
        rho_water=1000,
        rho_oil=910,
        E=3e5,
        L=1.0,
        boundary_mode="periodic",
        correction_grid_resolution=0,
    ):
        self.eps0 = 8.85418781762039e-12
        self.eps_oil = eps_oil
        self.eta_oil = eta_oil
        self.eta_water = eta_water
        self.rho_water = rho_water
        self.rho_oil = rho_oil
        self.E = E
        self.m_const = 12 * np.pi * self.eps0 * eps_oil * E**2
        self.eta_const = 1 / (8 * np.pi * self.eta_oil)

        self.L = L
        self.boundary_mode = boundary_mode
        self.num_particles = num_particles
        self.correction_grid_resolution = correction_grid_resolution

        self._corr: DeviceCorrection | None = None
        self._ws = _device.Workspace()

    # ------------------------------------------------------------------ setup

    def _phys(self) -> DemPhys:
        return DemPhys(
            float(self.m_const),
            float(self.eta_const),
            float(self.L),
            1 if self.boundary_mode == "periodic" else 0,
            0,
        )

    def load_periodic_correction(self, correction, L_sim: float):
        """Copy the COMSOL correction grids to the device (reference :74-120)."""
        if self.correction_grid_resolution == 0:
            raise RuntimeError(
                "correction_grid_resolution=0: no correction storage was reserved; "
                "recreate DirectDropletForceCalculator with correction_grid_resolution > 0."
            )
        gd = correction.get_grid_data()
        eta_ratio = gd["eta_comsol"] / self.eta_oil
        self._corr = DeviceCorrection(correction, L_sim, eta_ratio)
        print(
            f"[DirectDropletForceCalculator] periodic correction loaded: "
            f"grid {gd['grid_resolution']}^3, L_ratio={self._corr.L_ratio:.6f}, "
            f"Fz_comsol={gd['Fz_comsol']:.4e}, eta_ratio={eta_ratio:.6f}, "
            f"self_coeff={self._corr.self_coeff:.6e}"
        )

    @property
    def corr_enabled(self) -> int:
        return 1 if self._corr is not None else 0

    def _corr_ref(self):
        return self._corr.ref() if self._corr is not None else None

    # ------------------------------------------------------------------ device entry points

    def forces_device(self, pos_d, radii_d):
        """Forces of device tensors (N,3)/(N,) float64 -> device (N,3) float64."""
        n = int(pos_d.shape[0])
        F = _device.empty((n, 3))
        if n == 0:
            return F
        ws = self._ws.get(_native.load().dem_direct_workspace_bytes(n))
        ph = self._phys()
        _native.call("dem_direct_forces", _device.ptr(pos_d), _device.ptr(radii_d), n,
                     _byref(ph), _device.ptr(F), _device.ptr(ws), ws.numel(), _device.stream())
        return F

    def forces_and_convection_device(self, pos_d, radii_d):
        n = int(pos_d.shape[0])
        F = _device.empty((n, 3))
        V = _device.empty((n, 3))
        if n == 0:
            return F, V
        ws = self._ws.get(_native.load().dem_direct_workspace_bytes(n))
        ph = self._phys()
        _native.call("dem_direct_forces_and_convection", _device.ptr(pos_d),
                     _device.ptr(radii_d), n, _byref(ph), self._corr_ref(), _device.ptr(F),
                     _device.ptr(V), _device.ptr(ws), ws.numel(), _device.stream())
        return F, V

    def convection_device(self, pos_d, radii_d, F_d):
        n = int(pos_d.shape[0])
        V = _device.empty((n, 3))
        if n == 0:
            return V
        ws = self._ws.get(_native.load().dem_direct_workspace_bytes(n))
        ph = self._phys()
        _native.call("dem_direct_convection", _device.ptr(pos_d), _device.ptr(radii_d),
                     _device.ptr(F_d), n, _byref(ph), self._corr_ref(), _device.ptr(V),
                     _device.ptr(ws), ws.numel(), _device.stream())
        return V

    def forces_and_convection_f64_device(self, pos_d, radii_d):
        """FP64 oracle kernel (checker): same outputs in float64 arithmetic."""
        n = int(pos_d.shape[0])
        F = _device.empty((n, 3))
        V = _device.empty((n, 3))
        if n == 0:
            return F, V
        ph = self._phys()
        _native.call("dem_direct_forces_and_convection_f64", _device.ptr(pos_d),
                     _device.ptr(radii_d), n, _byref(ph), self._corr_ref(), _device.ptr(F),
                     _device.ptr(V), _device.stream())
        return F, V

    def f64_targets_device(self, pos_d, radii_d, targets_d, sources_forces_d=None):
        """FP64 checker on a target subset: forces of the targets and, given the
        forces of all sources, their Oseen velocities (device tensors)."""
        nt = int(targets_d.shape[0])
        F = _device.empty((nt, 3))
        V = _device.empty((nt, 3)) if sources_forces_d is not None else None
        ph = self._phys()
        _native.call("dem_direct_f64_targets", _device.ptr(pos_d), _device.ptr(radii_d),
                     int(pos_d.shape[0]), _byref(ph), self._corr_ref(), _device.ptr(targets_d),
                     nt, _device.ptr(sources_forces_d), _device.ptr(F), _device.ptr(V),
                     _device.stream())
        return F, V

    # ------------------------------------------------------------------ reference API

    def calculate(self, positions: np.ndarray, radii: np.ndarray) -> np.ndarray:
        """Dipole forces (N,3) float64 (reference :260-305)."""
        self.num_particles = positions.shape[0]
        pos_d = _device.to_device(np.asarray(positions).reshape(-1, 3))
        r_d = _device.to_device(radii)
        F = self.forces_device(pos_d, r_d)
        return F.cpu().numpy()

    def calculate_convection(self, positions: np.ndarray, radii: np.ndarray,
                             forces: np.ndarray) -> np.ndarray:
        """Oseen convection (N,3) float64 driven by `forces` (reference :307-352)."""
        self.num_particles = positions.shape[0]
        pos_d = _device.to_device(np.asarray(positions).reshape(-1, 3))
        r_d = _device.to_device(radii)
        F_d = _device.to_device(np.asarray(forces).reshape(-1, 3))
        V = self.convection_device(pos_d, r_d, F_d)
        return V.cpu().numpy()

    def calculate_forces_and_convection(self, positions: np.ndarray, radii: np.ndarray) -> tuple:
        """(forces, velocities), both (N,3) float64 (reference :470-527)."""
        self.num_particles = positions.shape[0]
        pos_d = _device.to_device(np.asarray(positions).reshape(-1, 3))
        r_d = _device.to_device(radii)
        F, V = self.forces_and_convection_device(pos_d, r_d)
        return F.cpu().numpy(), V.cpu().numpy()


def _byref(s):
    import ctypes

    return ctypes.byref(s)
