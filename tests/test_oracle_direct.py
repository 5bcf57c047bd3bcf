"""Pins the FP64 oracle's direct force and Oseen convection
(oracle/dem_oracle.c, restating
dem/force_calculator/direct_droplet_force_calculator.py:183-253, :354-468)
against closed forms, invariants and independent formulations."""

from __future__ import annotations

import numpy as np
import pytest

import oracle

M = oracle.m_const(2.85, 3e5)
ETA = oracle.eta_const(0.065)


def _pair(d, axis, ri=5e-6, rj=7e-6, L=1.0, periodic=False, offset=0.3):
    pos = np.full((2, 3), offset)
    pos[1, axis] += d
    return pos, np.array([ri, rj]), L, periodic


def test_closed_form_along_field():
    d, ri, rj = 40e-6, 5e-6, 7e-6
    pos, r, L, per = _pair(d, 2, ri, rj)
    d = pos[1, 2] - pos[0, 2]  # the separation actually represented
    F = oracle.direct_forces(pos, r, M, L, per)
    expect = 2 * M * ri**3 * rj**3 / d**4
    np.testing.assert_allclose(F[0], [0, 0, expect], rtol=1e-12, atol=0)
    np.testing.assert_allclose(F[1], [0, 0, -expect], rtol=1e-12, atol=0)


def test_closed_form_across_field():
    d, ri, rj = 35e-6, 6e-6, 3e-6
    pos, r, L, per = _pair(d, 0, ri, rj)
    d = pos[1, 0] - pos[0, 0]
    F = oracle.direct_forces(pos, r, M, L, per)
    expect = -M * ri**3 * rj**3 / d**4  # repulsion: i pushed to -x
    np.testing.assert_allclose(F[0], [expect, 0, 0], rtol=1e-12, atol=0)


def test_newton_third_law_and_zero_net():
    rng = np.random.default_rng(3)
    n = 40
    pos = rng.uniform(0, 4e-4, size=(n, 3))
    r = rng.uniform(2e-6, 6e-6, size=n)
    F = oracle.direct_forces(pos, r, M, 4e-4, True)
    scale = np.abs(F).sum()
    assert np.abs(F.sum(axis=0)).max() < 1e-12 * scale


def _potential(pos, r, i):
    """U_i = sum_j (m/3) r_i^3 r_j^3 (r^2 - 3 dz^2) / r^5 (open box)."""
    d = pos - pos[i]
    d = np.delete(d, i, axis=0)
    rj = np.delete(r, i)
    rr = np.sqrt(np.sum(d * d, axis=1))
    return np.sum(M / 3 * r[i] ** 3 * rj**3 * (rr**2 - 3 * d[:, 2] ** 2) / rr**5)


def test_force_is_minus_gradient_of_dipole_potential():
    rng = np.random.default_rng(7)
    n = 6
    pos = rng.uniform(0, 1e-4, size=(n, 3))
    r = rng.uniform(2e-6, 4e-6, size=n)
    F = oracle.direct_forces(pos, r, M, 1.0, False)
    h = 1e-10
    for i in range(n):
        g = np.zeros(3)
        for a in range(3):
            p1, p2 = pos.copy(), pos.copy()
            p1[i, a] += h
            p2[i, a] -= h
            g[a] = (_potential(p1, r, i) - _potential(p2, r, i)) / (2 * h)
        np.testing.assert_allclose(F[i], -g, rtol=1e-5, atol=1e-8 * np.abs(F[i]).max())


def test_minimum_image_flips_sign():
    """Mirrors the reference's test_periodic_changes_force_vs_open."""
    pos = np.array([[0.05, 0.5, 0.5], [0.95, 0.5, 0.5]])
    r = np.array([1e-5, 1e-5])
    f_open = oracle.direct_forces(pos, r, M, 1.0, False)
    f_per = oracle.direct_forces(pos, r, M, 1.0, True)
    assert f_open[0, 0] * f_per[0, 0] < 0
    # the periodic image sits at -0.1 (not +0.9): repulsion pushes particle 0 to +x
    d = (pos[1, 0] - pos[0, 0]) - 1.0
    np.testing.assert_allclose(f_per[0, 0], M * 1e-30 / d**4, rtol=1e-12)


def test_symmetric_lattice_has_zero_force():
    K, L = 5, 5e-4
    g = (np.arange(K) + 0.5) * L / K
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    r = np.full(K**3, 5e-6)
    F = oracle.direct_forces(pos, r, M, L, True)
    pair = M * (5e-6) ** 6 / (L / K) ** 4
    assert np.abs(F).max() < 1e-9 * pair


def test_scaling_laws():
    rng = np.random.default_rng(11)
    pos = rng.uniform(0, 2e-4, size=(10, 3))
    r = rng.uniform(2e-6, 5e-6, size=10)
    F1 = oracle.direct_forces(pos, r, M, 1.0, False)
    F2 = oracle.direct_forces(pos, r, 4 * M, 1.0, False)  # E -> 2E
    np.testing.assert_allclose(F2, 4 * F1, rtol=1e-13)
    s = 3.0
    F3 = oracle.direct_forces(pos * s, r * s, M, 1.0, False)
    np.testing.assert_allclose(F3, s**2 * F1, rtol=1e-12)


def test_contact_mask():
    """Touching / overlapping pairs contribute nothing (mask |d| > r_i + r_j)."""
    for d in (9e-6, 10e-6):  # overlapping, exactly touching
        pos, r, L, per = _pair(d, 2, 5e-6, 5e-6, offset=0.0)
        assert np.all(oracle.direct_forces(pos, r, M, L, per) == 0)
    pos, r, L, per = _pair(10.0001e-6, 2, 5e-6, 5e-6, offset=0.0)
    assert np.all(oracle.direct_forces(pos, r, M, L, per)[:, 2] != 0)


def _unit_vector_forces(pos, r, L, periodic):
    """Independent formulation with unit vectors: F = m r_i^3 r_j^3 / r^4 g(u)."""
    n = len(r)
    F = np.zeros((n, 3))
    for i in range(n):
        d = pos - pos[i]
        if periodic:
            d = np.where(np.abs(d) > L / 2, d - np.sign(d) * L, d)
        rr = np.linalg.norm(d, axis=1)
        ok = rr > r + r[i]
        ok[i] = False
        u = d[ok] / rr[ok, None]
        c = M * r[i] ** 3 * r[ok] ** 3 / rr[ok] ** 4
        g = np.stack([u[:, 0] * (4 * u[:, 2] ** 2 - u[:, 0] ** 2 - u[:, 1] ** 2),
                      u[:, 1] * (4 * u[:, 2] ** 2 - u[:, 0] ** 2 - u[:, 1] ** 2),
                      u[:, 2] * (2 * u[:, 2] ** 2 - 3 * u[:, 0] ** 2 - 3 * u[:, 1] ** 2)], axis=1)
        F[i] = (c[:, None] * g).sum(axis=0)
    return F


@pytest.mark.parametrize("periodic", [False, True])
def test_matches_independent_formulation(periodic):
    rng = np.random.default_rng(5)
    L = 3e-4
    pos = rng.uniform(0, L, size=(30, 3))
    r = rng.uniform(2e-6, 8e-6, size=30)
    np.testing.assert_allclose(oracle.direct_forces(pos, r, M, L, periodic),
                               _unit_vector_forces(pos, r, L, periodic), rtol=1e-10, atol=0)


# ---------------------------------------------------------------- Oseen


def test_oseen_two_body_closed_form():
    rng = np.random.default_rng(2)
    pos = np.array([[0.1, 0.2, 0.3], [0.1 + 3e-5, 0.2 - 1e-5, 0.3 + 2e-5]])
    F = rng.normal(size=(2, 3)) * 1e-12
    V = oracle.direct_convection(pos, F, ETA, 1.0, False)
    x = pos[0] - pos[1]
    rr = np.linalg.norm(x)
    u = x / rr
    expect = ETA / rr * (F[1] + np.dot(F[1], u) * u)
    np.testing.assert_allclose(V[0], expect, rtol=1e-12)


def test_oseen_linear_and_inverse_viscosity():
    rng = np.random.default_rng(4)
    pos = rng.uniform(0, 2e-4, size=(25, 3))
    F1 = rng.normal(size=(25, 3))
    F2 = rng.normal(size=(25, 3))
    V1 = oracle.direct_convection(pos, F1, ETA, 2e-4, True)
    V2 = oracle.direct_convection(pos, F2, ETA, 2e-4, True)
    V12 = oracle.direct_convection(pos, 2 * F1 - F2, ETA, 2e-4, True)
    np.testing.assert_allclose(V12, 2 * V1 - V2, rtol=1e-10, atol=1e-12 * np.abs(V1).max())
    V1h = oracle.direct_convection(pos, F1, ETA / 2, 2e-4, True)
    np.testing.assert_allclose(V1h, V1 / 2, rtol=1e-13)


def test_oseen_tensor_symmetry():
    """G(x) is symmetric: F_j . v_i(F_k) == F_k . v_i(F_j) for one source."""
    rng = np.random.default_rng(9)
    pos = np.array([[0.0, 0.0, 0.0], rng.uniform(1e-5, 5e-5, size=3)])
    a, b = rng.normal(size=3), rng.normal(size=3)
    va = oracle.direct_convection(pos, np.array([[0, 0, 0], a]), ETA, 1.0, False)[0]
    vb = oracle.direct_convection(pos, np.array([[0, 0, 0], b]), ETA, 1.0, False)[0]
    np.testing.assert_allclose(np.dot(b, va), np.dot(a, vb), rtol=1e-12)


def test_oseen_has_no_contact_mask():
    pos = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1e-6]])  # overlapping droplets
    F = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    V = oracle.direct_convection(pos, F, ETA, 1.0, False)
    np.testing.assert_allclose(V[0, 2], ETA * 2 / 1e-6, rtol=1e-12)
