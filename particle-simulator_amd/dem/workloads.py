"""Seeded synthetic droplet systems for benchmarks and parity tests.

Every system: radii r = clip(5 um * exp(0.3 z), 2 um, 20 um) with
z ~ N(0, 1) (log-normal, median 5 um, sigma 0.3); a periodic cube of edge
L = (sum 4/3 pi r^3 / phi)^(1/3) for volume fraction phi; positions uniform in
[0, L)^3 with overlapping droplets (minimum-image distance < r_i + r_j)
re-drawn until none overlap.  Physical constants: eps_oil = 2.85,
eta_oil = 0.05 Pa s, eta_water = 0.001 Pa s, E = 3e5 V/m along z unless an
ensemble draws its own.  Everything is drawn from numpy.random.default_rng.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

EPS0 = 8.85418781762039e-12


@dataclass
class System:
    positions: np.ndarray  # (N,3) float64, metres
    radii: np.ndarray      # (N,)  float64, metres
    L: float
    E: float = 3e5
    eps_oil: float = 2.85
    eta_oil: float = 0.05
    eta_water: float = 0.001
    phi: float = 0.02
    seed: int = 0
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.radii.shape[0])

    @property
    def m_const(self) -> float:
        return 12 * np.pi * EPS0 * self.eps_oil * self.E**2

    @property
    def eta_const(self) -> float:
        return 1 / (8 * np.pi * self.eta_oil)

    @property
    def stokes_factor(self) -> float:
        eo, ew = self.eta_oil, self.eta_water
        return 2 * np.pi * eo * (2 * eo + 3 * ew) / (eo + ew)


def lognormal_radii(rng: np.random.Generator, n: int) -> np.ndarray:
    z = rng.standard_normal(n)
    return np.clip(5e-6 * np.exp(0.3 * z), 2e-6, 20e-6)


def box_edge(radii: np.ndarray, phi: float) -> float:
    return float((np.sum(4.0 / 3.0 * np.pi * radii**3) / phi) ** (1.0 / 3.0))


def place_without_overlap(rng: np.random.Generator, radii: np.ndarray, L: float,
                          max_rounds: int = 10000) -> np.ndarray:
    """Uniform positions in [0,L)^3, re-drawing droplets that overlap under the
    minimum-image convention."""
    from scipy.spatial import cKDTree

    n = radii.shape[0]
    pos = rng.uniform(0.0, L, size=(n, 3))
    reach = 2.0 * float(np.max(radii))
    for _ in range(max_rounds):
        tree = cKDTree(pos, boxsize=L)
        pairs = tree.query_pairs(reach, output_type="ndarray")
        if len(pairs) == 0:
            return pos
        d = pos[pairs[:, 1]] - pos[pairs[:, 0]]
        d -= L * np.round(d / L)
        dist = np.sqrt(np.sum(d * d, axis=1))
        bad = pairs[dist < radii[pairs[:, 0]] + radii[pairs[:, 1]]]
        if len(bad) == 0:
            return pos
        redo = np.unique(bad[:, 1])
        pos[redo] = rng.uniform(0.0, L, size=(len(redo), 3))
        pos %= L
    raise ValueError("could not place droplets without overlap")


def make_system(n: int, seed: int = 0, phi: float = 0.02, E: float = 3e5) -> System:
    rng = np.random.default_rng(seed)
    radii = lognormal_radii(rng, n)
    L = box_edge(radii, phi)
    pos = place_without_overlap(rng, radii, L)
    return System(positions=pos, radii=radii, L=L, E=E, phi=phi, seed=seed)


def plain_small() -> System:
    """N = 2000, seed 0, phi = 0.02: the size the CPU reference path runs."""
    return make_system(2000, seed=0, phi=0.02)


def plain_at_scale() -> System:
    """N = 65536, seed 0, phi = 0.02 (L ~ 1.37 mm): the direct benchmark system."""
    return make_system(65536, seed=0, phi=0.02)


def tree_large() -> System:
    """N = 2^20, seed 0, phi = 0.02 (L ~ 3.45 mm): the Barnes-Hut system."""
    return make_system(1 << 20, seed=0, phi=0.02)


def differentiable() -> System:
    """N = 4096, seed 0, phi = 0.02: the system the adjoint is checked on."""
    return make_system(4096, seed=0, phi=0.02)


def differentiable_dt(s: System, E: float, mu: float, eps_oil: float = 2.85,
                      mu_water: float = 0.001) -> float:
    """dt of the differentiable workload: the fastest droplet moves 0.1 r_min
    in the first step (velocities of the FP32 direct path)."""
    from .force_calculator import DirectDropletForceCalculator

    fc = DirectDropletForceCalculator(num_particles=s.n, eps_oil=eps_oil, eta_oil=mu,
                                      eta_water=mu_water, E=E, L=s.L)
    F, V = fc.calculate_forces_and_convection(s.positions, s.radii)
    sf = 2 * np.pi * mu * (2 * mu + 3 * mu_water) / (mu + mu_water)
    v = np.linalg.norm(F / (sf * s.radii[:, None]) + V, axis=1)
    return float(0.1 * s.radii.min() / v.max())


def ensemble_specs(count: int = 256, base_seed: int = 0):
    """(N_e, E_e, phi_e) of the 256-ensemble batch: N_e ~ U{500..20000},
    E_e ~ U[1e5, 5e5] V/m, phi_e ~ U[0.01, 0.05], from default_rng(base_seed)."""
    rng = np.random.default_rng(base_seed)
    ns = rng.integers(500, 20001, size=count)
    Es = rng.uniform(1e5, 5e5, size=count)
    phis = rng.uniform(0.01, 0.05, size=count)
    return [(int(ns[e]), float(Es[e]), float(phis[e])) for e in range(count)]


def _make_ensemble(args) -> System:
    e, (n, E, phi) = args
    return make_system(n, seed=e, phi=phi, E=E)


def ensembles(count: int = 256, base_seed: int = 0, workers: int | None = None) -> list[System]:
    """The ensemble batch; system e is seeded with e, so the result does not
    depend on ``workers`` (forked processes used for the overlap-free
    placement; call before the GPU is initialised)."""
    import multiprocessing as mp
    import os

    jobs = list(enumerate(ensemble_specs(count, base_seed)))
    if workers is None:
        workers = min(16, os.cpu_count() or 1)
    if workers <= 1 or count < 4:
        return [_make_ensemble(j) for j in jobs]
    # close() + join() let the workers exit on their own; the pool's context
    # manager would terminate() them, a SIGTERM that profilers report as an abort
    pool = mp.get_context("fork").Pool(workers)
    try:
        out = pool.map(_make_ensemble, jobs, chunksize=1)
    finally:
        pool.close()
        pool.join()
    return out
