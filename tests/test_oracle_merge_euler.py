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


def _pair_lists(rng, n, k):
    """Random pair lists with chains (a droplet in several pairs) and repeats
    of a first index, canonicalised as detect() returns them."""
    a = rng.integers(0, n, k)
    b = (a + 1 + rng.integers(0, n - 1, k)) % n
    pairs = np.unique(np.sort(np.stack([a, b], 1), axis=1), axis=0).astype(np.int32)
    return pairs[pairs[:, 0] != pairs[:, 1]]


def test_product_merge_equals_oracle_merge_bit_for_bit():
    """merge_collided (the product's host merge, droplet_solution.py:208-271)
    equals the oracle's merge bit for bit: survivors, merged radii and
    centres, and their order.  The reference visits pairs in np.argsort order
    of the first index, which numpy does not keep stable for ties; the oracle
    visits them in stable order, so it is handed the pairs already in
    np.argsort order (and, for lists without repeated first indices, the raw
    list)."""
    from dem.solution.droplet_solution import merge_collided

    rng = np.random.default_rng(21)
    chained = 0
    for n, k in ((2, 1), (7, 5), (40, 30), (300, 250), (3000, 900)):
        for _ in range(5):
            pos = rng.random((n, 3)) * 1e-3
            r = rng.uniform(2e-6, 2e-5, n)
            pairs = _pair_lists(rng, n, k)
            p, rr, coll = merge_collided(pos, r, pairs)
            op, orr, res = oracle.merge(pos, r, pairs[np.argsort(pairs[:, 0])])
            assert np.array_equal(p, op) and np.array_equal(rr, orr)
            assert res == list(range(len(orr) - len(res), len(orr)))
            chained += len(pairs) - len(res)
            first = pairs[:, 0]
            if len(np.unique(first)) == len(first):
                op2, orr2, _ = oracle.merge(pos, r, pairs[::-1])
                assert np.array_equal(p, op2) and np.array_equal(rr, orr2)
    assert chained > 100, "the lists should contain skipped (chained) pairs"
