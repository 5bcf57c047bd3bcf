"""Pins the oracle's spatial-hash contact set (oracle/dem_oracle.c
``oracle_hash_pairs``, restating
dem/collision_detector/spatial_hash_collision_detector.py:73-316) with the
reference's own exact cases and with brute force."""

from __future__ import annotations

import numpy as np
import pytest

import oracle


def _table(cap):
    p = 1
    while p < 2 * cap:
        p <<= 1
    return p


# --- the reference's exact small cases (test_spatial_hash_collision_detector.py:32-113)


def test_no_overlap_no_collisions():
    pos = np.array([[0.10, 0.10, 0.10], [0.90, 0.90, 0.90]])
    is_coll, pairs = oracle.hash_pairs(pos, np.full(2, 0.05), 1.0, False, _table(8))
    assert is_coll is False and pairs.shape == (0, 2)


def test_overlap_returns_pair():
    pos = np.array([[0.50, 0.50, 0.50], [0.52, 0.50, 0.50]])
    is_coll, pairs = oracle.hash_pairs(pos, np.full(2, 0.05), 1.0, False, _table(8))
    assert is_coll is True
    assert pairs.tolist() == [[0, 1]]


def test_periodic_minimum_image_records_contact():
    pos = np.array([[0.01, 0.50, 0.50], [0.99, 0.50, 0.50]])
    r = np.full(2, 0.05)
    assert oracle.hash_pairs(pos, r, 1.0, False, _table(4))[0] is False
    is_yes, pairs = oracle.hash_pairs(pos, r, 1.0, True, _table(4))
    assert is_yes is True and pairs.tolist() == [[0, 1]]


def test_zero_radii_returns_no_collision():
    is_coll, pairs = oracle.hash_pairs(np.zeros((2, 3)), np.zeros(2), 1.0, False, _table(4))
    assert is_coll is False and pairs.shape == (0, 2)


def test_pairs_unique_after_dedup():
    pos = np.array([[0.50, 0.50, 0.50], [0.51, 0.50, 0.50], [0.52, 0.50, 0.50]])
    is_coll, pairs = oracle.hash_pairs(pos, np.full(3, 0.02), 1.0, False, _table(8))
    assert is_coll is True
    assert pairs.tolist() == [[0, 1], [0, 2], [1, 2]]


# --- brute force


def _cells_and_buckets(pos, r, L, periodic, T):
    """Cells, buckets and neighbour-bucket sets B(i) of every particle."""
    cell = 2 * r.max()
    g = max(1, int(np.ceil(L / cell)))
    c = np.floor(pos / cell).astype(np.int64)
    c = np.mod(c, g) if periodic else np.clip(c, 0, g - 1)

    def h(ix, iy, iz):
        v = ((ix * 73856093) & 0xFFFFFFFF) ^ ((iy * 19349663) & 0xFFFFFFFF) ^ ((iz * 83492791) & 0xFFFFFFFF)
        return v & (T - 1)

    buckets = np.array([h(*ci) for ci in c])
    B = []
    for ci in c:
        s = set()
        for d in np.array(np.meshgrid([-1, 0, 1], [-1, 0, 1], [-1, 0, 1])).reshape(3, -1).T:
            nb = ci + d
            if periodic:
                nb = np.mod(nb, g)
            elif np.any(nb < 0) or np.any(nb >= g):
                continue
            s.add(h(*nb))
        B.append(s)
    return buckets, B


@pytest.mark.parametrize("periodic", [True, False])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_hash_equals_brute_force_through_stencil(periodic, seed):
    rng = np.random.default_rng(seed)
    n, L = 300, 1e-3
    pos = rng.uniform(0, L, size=(n, 3))
    r = rng.uniform(10e-6, 45e-6, size=n)
    T = _table(n)
    _, pairs = oracle.hash_pairs(pos, r, L, periodic, T)
    brute = oracle.brute_pairs(pos, r, L, periodic)
    buckets, B = _cells_and_buckets(pos, r, L, periodic, T)
    expect = [tuple(p) for p in brute if buckets[p[1]] in B[p[0]]]
    assert [tuple(p) for p in pairs] == expect
    assert len(brute) > 20  # the case exercises contacts


def test_periodic_miss_across_narrow_last_cell():
    """L/cell = 4.5: cell 0 and cell 3 are not stencil neighbours although the
    minimum-image distance through the narrow cell 4 is below r_i + r_j."""
    r = np.full(2, 1 / 9)
    pos = np.array([[0.01, 0.5, 0.5], [0.87, 0.5, 0.5]])
    T = _table(1024)
    buckets, B = _cells_and_buckets(pos, r, 1.0, True, T)
    assert buckets[1] not in B[0]
    assert oracle.brute_pairs(pos, r, 1.0, True).tolist() == [[0, 1]]
    assert oracle.hash_pairs(pos, r, 1.0, True, T)[0] is False
