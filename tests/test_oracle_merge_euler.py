"""Pins the oracle's merge (droplet_solution.py:208-287) and Euler step
(euler_droplet_solver.py:119-131) with hand-computed cases."""

from __future__ import annotations

import numpy as np

import oracle


def test_merge_two_of_three():
    """Mirrors the reference's test_collision_merges_two_drops."""
    pos = np.array([[0.0, 0.0, 0.0], [0.1, 0.0, 0.0], [0.5, 0.5, 0.5]])
    r = np.array([0.05, 0.05, 0.05])
    p, rr, res = oracle.merge(pos, r, np.array([[0, 1]]))
    assert len(rr) == 2 and res == [1]
    assert np.isclose(rr[-1] ** 3, 2 * 0.05**3)
    np.testing.assert_allclose(p[0], [0.5, 0.5, 0.5])  # survivor keeps its order
    np.testing.assert_allclose(p[1], [0.05, 0.0, 0.0])  # equal volumes: midpoint


def test_merge_greedy_skips_used_and_weights_by_volume():
    pos = np.array([[0.0, 0, 0], [1.0, 0, 0], [2.0, 0, 0], [3.0, 0, 0], [9.0, 9, 9]])
    r = np.array([1.0, 2.0, 1.0, 1.0, 1.0])
    pairs = np.array([[1, 2], [0, 1], [2, 3]])  # sorted by first index: (0,1) first
    p, rr, res = oracle.merge(pos, r, pairs)
    # (0,1) merges; (1,2) skipped (1 used); (2,3) merges
    assert len(rr) == 3 and res == [1, 2]
    np.testing.assert_allclose(rr, [1.0, 9.0 ** (1 / 3), 2.0 ** (1 / 3)])
    np.testing.assert_allclose(p[1], [(0 * 1 + 1 * 8) / 9, 0, 0])
    np.testing.assert_allclose(p[2], [2.5, 0, 0])
    # raw coordinates, no minimum image
    p2, _, _ = oracle.merge(np.array([[0.01, 0, 0], [0.99, 0, 0]]), np.array([1.0, 1.0]),
                            np.array([[0, 1]]))
    np.testing.assert_allclose(p2[0], [0.5, 0, 0])


def test_euler_step_hand_computed():
    pos = np.array([[0.1, 0.2, 0.3]])
    F = np.array([[1e-12, -2e-12, 3e-12]])
    conv = np.array([[1e-6, 0.0, -1e-6]])
    r = np.array([5e-6])
    sf = oracle.stokes_factor(0.065, 0.001)
    x = oracle.euler_step(pos, F, conv, r, sf, 0.5, 1.0, False)
    np.testing.assert_allclose(x, pos + (F / (sf * 5e-6) + conv) * 0.5, rtol=0, atol=0)
    assert sf == 2 * np.pi * 0.065 * (2 * 0.065 + 3 * 0.001) / (0.065 + 0.001)


def test_euler_periodic_wrap():
    pos = np.array([[0.999, 0.0005, 0.5]])
    v = np.array([[0.01, -0.01, 0.0]])
    x = oracle.euler_step_total(pos, v, 0.2, 1.0, True)
    np.testing.assert_allclose(x, [[0.001, 0.9985, 0.5]], atol=1e-14)
    assert np.all((x >= 0) & (x < 1.0))
