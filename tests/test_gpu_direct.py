"""GPU parity of the direct force + Oseen kernels (libdemhip
``dem_direct_*``) against the FP64 oracle and the committed golden fixture.

Accuracy contract of the FP32 kernels, against FP64 on the same float64
inputs, with e_i = |x_i - ref_i| / max(|ref_i|, 1e-3 * rms|ref|):
* forces:     max_i e_i <= 1e-5;
* velocities: field error ||x - ref|| / ||ref|| <= 1e-5, 99.9th percentile of
  e_i <= 1e-4 and max_i e_i <= 1e-3.  The Oseen sum decays as 1/r, so for
  particles whose velocity nearly cancels the FP32 rounding of each term is
  amplified by sum_j |t_ij| / |v_i| (up to ~3e5 at N = 65536); no FP32
  evaluation reaches a per-particle 1e-5 for those few particles.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

import oracle
from dem.force_calculator import DirectDropletForceCalculator
from dem.workloads import make_system

from ._errors import summary

pytestmark = pytest.mark.gpu

TOL_F = 1e-5
TOL_V_NORM = 1e-5
TOL_V_P999 = 1e-4
TOL_V_MAX = 1e-3
FLOOR = 1e-3


def _ok_forces(e):
    return e["max"] <= TOL_F


def _ok_velocities(e):
    return e["norm"] <= TOL_V_NORM and e["p999"] <= TOL_V_P999 and e["max"] <= TOL_V_MAX


GOLDEN = Path(__file__).parent / "golden" / "direct_n2000.npz"


def _calc(s, boundary="periodic", n=None, corr=0):
    return DirectDropletForceCalculator(num_particles=n or s.n, eps_oil=s.eps_oil,
                                        eta_oil=s.eta_oil, eta_water=s.eta_water, E=s.E, L=s.L,
                                        boundary_mode=boundary, correction_grid_resolution=corr)


# --- the reference's own tests, on the GPU operator (tests/force_calculator/...:22-95)


def test_single_particle_zero_force(gpu):
    fc = DirectDropletForceCalculator(num_particles=4, L=1.0, boundary_mode="open")
    forces = fc.calculate(np.array([[0.5, 0.5, 0.5]]), np.array([1e-5]))
    assert forces.shape == (1, 3)
    np.testing.assert_allclose(forces[0], np.zeros(3))


def test_pair_newton_third_law(gpu):
    fc = DirectDropletForceCalculator(num_particles=4, L=1.0, boundary_mode="open")
    pos = np.array([[0.40, 0.50, 0.50], [0.60, 0.50, 0.50]])
    forces = fc.calculate(pos, np.array([2e-5, 2e-5]))
    np.testing.assert_allclose(forces[0], -forces[1], rtol=1e-10)
    # and the closed form F_x = -m r^6 / d^4 (repulsion across the field)
    np.testing.assert_allclose(forces[0, 0], -fc.m_const * (2e-5) ** 6 / 0.2**4, rtol=1e-5)


def test_returns_correct_shape(gpu):
    fc = DirectDropletForceCalculator(num_particles=8, L=1.0, boundary_mode="open")
    rng = np.random.default_rng(0)
    pos = rng.uniform(0.1, 0.9, size=(5, 3))
    r = rng.uniform(1e-6, 5e-6, size=5)
    assert fc.calculate(pos, r).shape == (5, 3)
    F, V = fc.calculate_forces_and_convection(pos, r)
    assert F.shape == (5, 3) and V.shape == (5, 3)


def test_periodic_changes_force_vs_open(gpu):
    pos = np.array([[0.05, 0.50, 0.50], [0.95, 0.50, 0.50]])
    r = np.array([1e-5, 1e-5])
    f_open = DirectDropletForceCalculator(num_particles=4, L=1.0, boundary_mode="open").calculate(pos, r)
    f_per = DirectDropletForceCalculator(num_particles=4, L=1.0, boundary_mode="periodic").calculate(pos, r)
    assert f_open[0, 0] * f_per[0, 0] < 0


# --- parity against the FP64 oracle


@pytest.mark.parametrize("boundary", ["periodic", "open"])
def test_forces_and_convection_match_oracle_n2000(gpu, boundary):
    s = make_system(2000, seed=0, phi=0.02)
    per = boundary == "periodic"
    fc = _calc(s, boundary)
    F, V = fc.calculate_forces_and_convection(s.positions, s.radii)
    Fo = oracle.direct_forces(s.positions, s.radii, fc.m_const, s.L, per)
    Vo = oracle.direct_convection(s.positions, Fo, fc.eta_const, s.L, per)
    ef, ev = summary(F, Fo, FLOOR), summary(V, Vo, FLOOR)
    print(f"\n[{boundary}] force err {ef}\n[{boundary}] conv  err {ev}")
    assert _ok_forces(ef) and _ok_velocities(ev)


def test_convection_from_caller_forces(gpu):
    s = make_system(1500, seed=3, phi=0.03)
    fc = _calc(s)
    rng = np.random.default_rng(1)
    F = rng.normal(size=(s.n, 3)) * 1e-12
    V = fc.calculate_convection(s.positions, s.radii, F)
    Vo = oracle.direct_convection(s.positions, F, fc.eta_const, s.L, True)
    assert _ok_velocities(summary(V, Vo, FLOOR))


def test_golden_fixture(gpu):
    g = np.load(GOLDEN)
    fc = DirectDropletForceCalculator(num_particles=len(g["radii"]), eps_oil=float(g["eps_oil"]),
                                      eta_oil=float(g["eta_oil"]), E=float(g["E"]),
                                      L=float(g["L"]), boundary_mode="periodic")
    F, V = fc.calculate_forces_and_convection(g["positions"], g["radii"])
    assert _ok_forces(summary(F, g["forces"], FLOOR))
    assert _ok_velocities(summary(V, g["velocities"], FLOOR))


@pytest.fixture(scope="module")
def correction():
    from dem.periodic_correction import COMSOLLatticeCorrection

    return COMSOLLatticeCorrection.load_default()


def test_comsol_correction_matches_oracle(gpu, correction):
    s = make_system(1200, seed=5, phi=0.02)
    fc = _calc(s, corr=correction.grid_resolution)
    fc.load_periodic_correction(correction, L_sim=s.L)
    F, V = fc.calculate_forces_and_convection(s.positions, s.radii)
    gd = correction.get_grid_data()
    c = oracle.Correction(gd["grid_u"], gd["grid_v"], gd["grid_w"], gd["grid_min"], gd["grid_dx"],
                          gd["Fz_comsol"], gd["L_comsol"] / s.L, gd["eta_comsol"] / s.eta_oil,
                          fc._corr.self_coeff)
    Fo = oracle.direct_forces(s.positions, s.radii, fc.m_const, s.L, True)
    Vo = oracle.direct_convection(s.positions, Fo, fc.eta_const, s.L, True, c)
    Vn = oracle.direct_convection(s.positions, Fo, fc.eta_const, s.L, True)
    e = summary(V, Vo, FLOOR)
    print(f"\n[comsol] conv err {e}; correction/total = "
          f"{np.linalg.norm(Vo - Vn) / np.linalg.norm(Vo):.3e}")
    assert _ok_velocities(e)
    # the FP64 GPU oracle kernel agrees with the CPU oracle
    import torch

    from dem import _device

    F64, V64 = fc.forces_and_convection_f64_device(_device.to_device(s.positions),
                                                   _device.to_device(s.radii))
    torch.cuda.synchronize()
    np.testing.assert_allclose(F64.cpu().numpy(), Fo, rtol=1e-9, atol=1e-12 * np.abs(Fo).max())
    np.testing.assert_allclose(V64.cpu().numpy(), Vo, rtol=1e-9, atol=1e-12 * np.abs(Vo).max())


def test_fp64_oracle_kernel_matches_cpu_oracle(gpu):
    s = make_system(1000, seed=2, phi=0.02)
    fc = _calc(s)
    from dem import _device

    F64, V64 = fc.forces_and_convection_f64_device(_device.to_device(s.positions),
                                                   _device.to_device(s.radii))
    Fo = oracle.direct_forces(s.positions, s.radii, fc.m_const, s.L, True)
    Vo = oracle.direct_convection(s.positions, Fo, fc.eta_const, s.L, True)
    np.testing.assert_allclose(F64.cpu().numpy(), Fo, rtol=1e-10, atol=1e-13 * np.abs(Fo).max())
    np.testing.assert_allclose(V64.cpu().numpy(), Vo, rtol=1e-10, atol=1e-13 * np.abs(Vo).max())


def test_n65536_against_fp64_kernel(gpu):
    """N = 65536 (the at-scale system): FP32 path vs the FP64 oracle kernel."""
    s = make_system(65536, seed=0, phi=0.02)
    fc = _calc(s)
    from dem import _device

    pd, rd = _device.to_device(s.positions), _device.to_device(s.radii)
    F, V = fc.forces_and_convection_device(pd, rd)
    F64, V64 = fc.forces_and_convection_f64_device(pd, rd)
    ef = summary(F.cpu().numpy(), F64.cpu().numpy(), FLOOR)
    ev = summary(V.cpu().numpy(), V64.cpu().numpy(), FLOOR)
    print(f"\n[N=65536] force err {ef}\n[N=65536] conv  err {ev}")
    assert _ok_forces(ef) and _ok_velocities(ev)
    # spot-check the FP64 kernel against the CPU oracle on a slice of targets
    Fo = oracle.direct_forces(s.positions, s.radii, fc.m_const, s.L, True, targets=(0, 256))
    np.testing.assert_allclose(F64.cpu().numpy()[:256], Fo[:256], rtol=1e-9,
                               atol=1e-12 * np.abs(Fo[:256]).max())


@pytest.mark.parametrize("n,boundary", [(10240, "periodic"), (20000, "periodic"),
                                        (20000, "open"), (131072, "periodic")])
def test_pair_once_kernels_against_fp64_kernel(gpu, n, boundary):
    """Pair-once (Newton's third law) kernels, used for 10240 <= N <= 131072:
    smallest and largest N, a ragged last block of 256 (N = 20000: 79 blocks,
    an odd count) and both boundary modes.  Accuracy contract against the FP64
    kernel, for the pair-once and the one-sided evaluation alike; the two
    round differently, which shows that the pair-once kernels ran.  At
    N = 131072 the largest per-droplet force error reaches ~2.6e-5 for both
    evaluations (one droplet whose force nearly cancels: the FP32 rounding of
    its dominant pair terms, amplified sum_j |t_ij| / |F_i| times), so the
    force bar there is field error and 99.9th percentile <= 1e-5, max <= 1e-4."""
    from dem import _device, _native

    s = make_system(n, seed=1, phi=0.02)
    fc = _calc(s, boundary)
    pd, rd = _device.to_device(s.positions), _device.to_device(s.radii)
    F, V = fc.forces_and_convection_device(pd, rd)
    prev = _native.call("dem_direct_set_symmetric", 0)
    try:
        F1, V1 = fc.forces_and_convection_device(pd, rd)
    finally:
        _native.call("dem_direct_set_symmetric", prev)
    F64, V64 = fc.forces_and_convection_f64_device(pd, rd)
    F, V, F1, V1 = (t.cpu().numpy() for t in (F, V, F1, V1))
    F64, V64 = F64.cpu().numpy(), V64.cpu().numpy()
    ok = []
    for tag, (Fx, Vx) in (("pair-once", (F, V)), ("one-sided", (F1, V1))):
        ef = summary(Fx, F64, FLOOR)
        ev = summary(Vx, V64, FLOOR)
        print(f"\n[N={n} {boundary} {tag}] force err {ef}\n[N={n} {boundary} {tag}] conv  err {ev}")
        if n <= 65536:
            ok.append(_ok_forces(ef) and _ok_velocities(ev))
        else:
            ok.append(ef["norm"] <= TOL_F and ef["p999"] <= TOL_F and ef["max"] <= 1e-4
                      and _ok_velocities(ev))
    assert all(ok)
    assert not np.array_equal(F, F1) and not np.array_equal(V, V1)
    # repeated evaluation is bit-identical (fixed summation order)
    F2, V2 = fc.forces_and_convection_device(pd, rd)
    assert np.array_equal(F2.cpu().numpy(), F) and np.array_equal(V2.cpu().numpy(), V)


@pytest.mark.parametrize("n,split", [(300, 0), (2000, 1), (2000, 3), (5000, 7)])
def test_forced_pair_once_small_n(gpu, n, split):
    """Pair-once mode 2 (every N, explicit chunk count) under the accuracy
    contract: one block of 256 plus padding (N = 300), several blocks with
    1 / 3 / 7 chunks of the partner lists."""
    from dem import _device, _native

    s = make_system(n, seed=5, phi=0.04)
    fc = _calc(s)
    pd, rd = _device.to_device(s.positions), _device.to_device(s.radii)
    prev_sym = _native.call("dem_direct_set_symmetric", 2)
    prev = _native.call("dem_direct_set_tuning", split, 0)
    try:
        F, V = fc.forces_and_convection_device(pd, rd)
    finally:
        _native.call("dem_direct_set_tuning", prev >> 16, prev & 0xFFFF)
        _native.call("dem_direct_set_symmetric", prev_sym)
    F64, V64 = fc.forces_and_convection_f64_device(pd, rd)
    ef = summary(F.cpu().numpy(), F64.cpu().numpy(), FLOOR)
    ev = summary(V.cpu().numpy(), V64.cpu().numpy(), FLOOR)
    print(f"\n[pair-once mode 2 N={n} split={split}] force {ef}\n  conv {ev}")
    assert _ok_forces(ef) and _ok_velocities(ev)


# --- 32-bit Oseen separations: pairs at half the box (dem_direct_set_image32)


def _planted_half_box_system(dup: bool = False):
    """N = 10240 (pair-once range), periodic, box edge 2^-10 m so that the
    fixed-point ring words q = (x / L) 2^64 are exact.  Pairs are planted
    across half the box: droplet j sits at q_j = q_i + 2^63 -+ 2^11 on one axis,
    so the 32-bit high words differ by exactly 2^31.  With the borrow (-) the
    true separation is just below L/2 and the 32-bit difference names the other
    image; without it (+) both agree.  One pair is planted on two axes.  dup:
    a third droplet takes droplet 10's x, so the high word that droplet 5000
    looks up is shared and the call must fall back to 64-bit separations."""
    s = make_system(10240, seed=3, phi=0.02)
    L = 2.0**-10
    u = s.positions / s.L
    u -= np.floor(u)
    radii = s.radii * (L / s.L)
    plants = [(10, 5000, (0,), True), (20, 6000, (1,), True), (30, 7000, (2,), True),
              (40, 8000, (0,), False), (60, 9000, (0, 2), True)]
    for k, (i, j, axes, borrow) in enumerate(plants):
        for a in axes:
            qi = 2**62 + (k + 1) * 2**40 + (a + 1) * 2**20
            qj = qi + 2**63 + (-(2**11) if borrow else 2**11)
            u[i, a], u[j, a] = qi / 2.0**64, qj / 2.0**64
            assert int(u[i, a] * 2.0**64) == qi and int(u[j, a] * 2.0**64) == qj
    if dup:
        u[50, 0] = u[10, 0]
    return s, u * L, radii, L, plants


def _oseen_flip(pos, F, eta, L, i, j, axes):
    """Relative change of v_i if the pair (i, j) took the other image on
    `axes` (float64)."""
    d = pos[j] - pos[i]
    d -= L * np.round(d / L)
    df = d.copy()
    df[list(axes)] *= -1.0

    def G(dd):
        r = np.linalg.norm(dd)
        return eta * (F[j] / r + dd * np.dot(dd, F[j]) / r**3)

    return G(df) - G(d)


def test_image32_boundary_pairs_corrected(gpu):
    """The default pair-once Oseen sum (32-bit separations + boundary-pair
    correction) against the FP64 kernel and against 64-bit separations: the
    contract holds, forces are untouched (bit-identical), and every planted
    droplet whose velocity a wrong image would move by >= 1e-3 is within 5 % of
    that move and no worse than the 64-bit evaluation."""
    from dem import _device, _native

    s, pos, radii, L, plants = _planted_half_box_system()
    fc = DirectDropletForceCalculator(num_particles=s.n, eps_oil=s.eps_oil, eta_oil=s.eta_oil,
                                      eta_water=s.eta_water, E=s.E, L=L, boundary_mode="periodic")
    pd, rd = _device.to_device(pos), _device.to_device(radii)
    F, V = fc.forces_and_convection_device(pd, rd)
    prev = _native.call("dem_direct_set_image32", 0)
    try:
        F1, V1 = fc.forces_and_convection_device(pd, rd)
    finally:
        _native.call("dem_direct_set_image32", prev)
    F64, V64 = fc.forces_and_convection_f64_device(pd, rd)
    F, V, F1, V1, F64, V64 = (t.cpu().numpy() for t in (F, V, F1, V1, F64, V64))
    assert np.array_equal(F, F1)
    ev, ev1 = summary(V, V64, FLOOR), summary(V1, V64, FLOOR)
    print(f"\n[planted half-box pairs] image32 {ev}\n  64-bit {ev1}")
    assert _ok_forces(summary(F, F64, FLOOR)) and _ok_velocities(ev) and _ok_velocities(ev1)
    from ._errors import rel_err

    e, e1 = rel_err(V, V64, FLOOR), rel_err(V1, V64, FLOOR)
    rms = float(np.sqrt(np.mean(np.sum(V64**2, axis=1))))
    checked = 0
    for i, j, axes, _ in plants:
        for p, q in ((i, j), (j, i)):
            flip = np.linalg.norm(_oseen_flip(pos, F64, fc.eta_const, L, p, q, axes))
            rel = flip / max(np.linalg.norm(V64[p]), FLOOR * rms)
            print(f"  droplet {p}: wrong-image move {rel:.2e}, error {e[p]:.2e} (64-bit {e1[p]:.2e})")
            assert e[p] <= 2.0 * e1[p] + 1e-6
            if rel >= 1e-3:
                assert e[p] <= 0.05 * rel
                checked += 1
    assert checked >= 5


def test_image32_duplicate_word_falls_back(gpu):
    """A high word shared by two droplets, one of which has a partner across
    half the box, makes the call take 64-bit separations: the result equals the
    64-bit evaluation bit for bit."""
    from dem import _device, _native

    s, pos, radii, L, _ = _planted_half_box_system(dup=True)
    fc = DirectDropletForceCalculator(num_particles=s.n, eps_oil=s.eps_oil, eta_oil=s.eta_oil,
                                      eta_water=s.eta_water, E=s.E, L=L, boundary_mode="periodic")
    pd, rd = _device.to_device(pos), _device.to_device(radii)
    F, V = fc.forces_and_convection_device(pd, rd)
    prev = _native.call("dem_direct_set_image32", 0)
    try:
        F1, V1 = fc.forces_and_convection_device(pd, rd)
    finally:
        _native.call("dem_direct_set_image32", prev)
    assert np.array_equal(F.cpu().numpy(), F1.cpu().numpy())
    assert np.array_equal(V.cpu().numpy(), V1.cpu().numpy())
    # and the same positions without the duplicate take the 32-bit path
    s2, pos2, radii2, _, _ = _planted_half_box_system(dup=False)
    F2, V2 = fc.forces_and_convection_device(_device.to_device(pos2), _device.to_device(radii2))
    _native.call("dem_direct_set_image32", 0)
    try:
        F3, V3 = fc.forces_and_convection_device(_device.to_device(pos2), _device.to_device(radii2))
    finally:
        _native.call("dem_direct_set_image32", prev)
    assert not np.array_equal(V2.cpu().numpy(), V3.cpu().numpy())
