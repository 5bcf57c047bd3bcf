"""COMSOL periodic-image correction of the Oseen tensor (host side).

Drop-in for the reference's ``COMSOLLatticeCorrection``
(dem/periodic_correction/comsol_lattice_correction.py:8-328).  The data
(U_lattice.txt: velocity of one droplet in a periodic cell minus the
free-space Stokeslet, sampled on an irregular quarter cell
[0, L/2]^2 x [-L/2, L/2]) is
1. mirrored to the full cell with the Stokeslet's parities for a force along
   z: u odd in x / even in y, v even in x / odd in y, w even in both
   (points on the x = 0 / y = 0 planes are not duplicated), :111-166;
2. resampled onto an n^3 regular grid over [-L/2, L/2]^3 with scipy's linear
   scattered-data interpolation, nearest-neighbour fill where that is
   undefined, :168-220;
3. evaluated with trilinear ``RegularGridInterpolator`` (zero outside), and
   for an arbitrary force direction by the cubic-lattice permutations
   G_x(x,y,z) = (w,v,u)(z,y,x), G_y(x,y,z) = (u,w,v)(x,z,y), :248-297.
The grids go to the GPU through ``load_periodic_correction`` of the force
operators; building them is a one-time host job and is not accelerated.
"""

from __future__ import annotations

import json
import os

import numpy as np
from scipy.interpolate import LinearNDInterpolator, NearestNDInterpolator, RegularGridInterpolator
from scipy.spatial import Delaunay


class COMSOLLatticeCorrection:
    _DEFAULT_DATA_DIR = os.path.join(os.path.dirname(__file__), "data")

    @classmethod
    def load_default(cls):
        return cls.from_data_dir(cls._DEFAULT_DATA_DIR)

    @classmethod
    def from_data_dir(cls, data_dir: str):
        with open(os.path.join(data_dir, "comsol_params.json")) as f:
            p = json.load(f)
        Fz = 4 * np.pi * p["R"] ** 2 * p["eps_0"] * p["eps_r"] * p["E0"] ** 2
        return cls(
            data_path=os.path.join(data_dir, p["velocity_file"]),
            L_comsol=p["L"],
            Fz_comsol=Fz,
            eta_comsol=p["eta"],
            grid_resolution=p.get("grid_resolution", 48),
        )

    def __init__(self, data_path: str, L_comsol: float, Fz_comsol: float,
                 eta_comsol: float = 0.065, grid_resolution: int = 48):
        self.L_comsol = L_comsol
        self.Fz_comsol = Fz_comsol
        self.eta_comsol = eta_comsol
        self.grid_resolution = grid_resolution
        pts, u, v, w = self._load_and_parse(data_path)
        fp, fu, fv, fw = self._mirror_to_full_cell(pts, u, v, w)
        self.grid_coords, self.grid_u, self.grid_v, self.grid_w = self._build_regular_grid(
            fp, fu, fv, fw, grid_resolution
        )
        self._interp_u, self._interp_v, self._interp_w = self._build_interpolators()
        print(
            f"[COMSOLLatticeCorrection] {len(pts)} points, mirrored to {len(fp)}, "
            f"resampled on a {grid_resolution}^3 grid"
        )

    @staticmethod
    def _load_and_parse(data_path: str):
        data = np.loadtxt(data_path, comments="%")
        pts = data[:, 0:3]
        if not (np.all(pts[:, 0] >= -1e-15) and np.all(pts[:, 1] >= -1e-15)):
            raise ValueError("expected quarter-cell data with x >= 0 and y >= 0")
        return pts, data[:, 3], data[:, 4], data[:, 5]

    @staticmethod
    def _mirror_to_full_cell(pts, u, v, w):
        eps = 1e-15
        px = pts[:, 0] > eps
        py = pts[:, 1] > eps
        # (sign of x, sign of y) -> parity factors of (u, v, w)
        blocks_p, blocks_u, blocks_v, blocks_w = [pts], [u], [v], [w]
        for mask, sx, sy in ((px, -1.0, 1.0), (py, 1.0, -1.0), (px & py, -1.0, -1.0)):
            if not np.any(mask):
                continue
            q = pts[mask].copy()
            q[:, 0] *= sx
            q[:, 1] *= sy
            blocks_p.append(q)
            blocks_u.append(sx * u[mask])
            blocks_v.append(sy * v[mask])
            blocks_w.append(w[mask])
        return (np.concatenate(blocks_p), np.concatenate(blocks_u), np.concatenate(blocks_v),
                np.concatenate(blocks_w))

    def _build_regular_grid(self, pts, u, v, w, n):
        h = self.L_comsol / 2.0
        axis = np.linspace(-h, h, n)
        gx, gy, gz = np.meshgrid(axis, axis, axis, indexing="ij")
        q = np.column_stack([gx.ravel(), gy.ravel(), gz.ravel()])
        # One triangulation and one simplex search serve all three components
        # (the barycentric weights are per point, so each column gets exactly
        # the values a single-column interpolator would give).
        tri = Delaunay(pts)
        vals = np.column_stack([u, v, w])
        lin = LinearNDInterpolator(tri, vals)(q)
        grids = []
        for k in range(3):
            g = lin[:, k].reshape(n, n, n).copy()
            bad = np.isnan(g)
            if np.any(bad):
                idx = np.argwhere(bad)
                g[bad] = NearestNDInterpolator(pts, vals[:, k])(
                    np.column_stack([axis[idx[:, 0]], axis[idx[:, 1]], axis[idx[:, 2]]])
                )
            grids.append(g)
        return axis, grids[0], grids[1], grids[2]

    def _build_interpolators(self):
        c = self.grid_coords
        mk = lambda g: RegularGridInterpolator(  # noqa: E731
            (c, c, c), g, method="linear", bounds_error=False, fill_value=0.0
        )
        return mk(self.grid_u), mk(self.grid_v), mk(self.grid_w)

    def evaluate(self, r_rel: np.ndarray, F_j: np.ndarray = None) -> np.ndarray:
        """Correction velocity at COMSOL-frame offsets r_rel (M,3); raw z-force
        column when F_j is None, else the full tensor applied to F_j (M,3)."""
        if F_j is None:
            return np.column_stack(
                [self._interp_u(r_rel), self._interp_v(r_rel), self._interp_w(r_rel)]
            )
        inv = 1.0 / self.Fz_comsol
        pz, px, py = r_rel, r_rel[:, [2, 1, 0]], r_rel[:, [0, 2, 1]]
        Gz = (self._interp_u(pz), self._interp_v(pz), self._interp_w(pz))
        Gx = (self._interp_w(px), self._interp_v(px), self._interp_u(px))
        Gy = (self._interp_u(py), self._interp_w(py), self._interp_v(py))
        s = (F_j[:, 0] * inv, F_j[:, 1] * inv, F_j[:, 2] * inv)
        return np.column_stack(
            [Gx[a] * s[0] + Gy[a] * s[1] + Gz[a] * s[2] for a in range(3)]
        )

    def get_grid_data(self) -> dict:
        c = self.grid_coords
        return {
            "grid_u": self.grid_u.copy(),
            "grid_v": self.grid_v.copy(),
            "grid_w": self.grid_w.copy(),
            "grid_min": c[0],
            "grid_max": c[-1],
            "grid_dx": c[1] - c[0],
            "grid_resolution": self.grid_resolution,
            "L_comsol": self.L_comsol,
            "Fz_comsol": self.Fz_comsol,
            "eta_comsol": self.eta_comsol,
        }
