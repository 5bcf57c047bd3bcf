"""Pins the oracle's octree (oracle/dem_octree_oracle.c, restating
dem/octree/flat_tree.py) with the reference's own tests
(tests/octree/test_flat_tree.py, test_force_tree.py, test_tree_stats.py), the
theta = 0 identity with the direct sum, and structural invariants."""

from __future__ import annotations

import numpy as np
import pytest

import oracle

M = oracle.m_const(2.85, 3e5)
ETA = oracle.eta_const(0.065)


def _rel(ref, x, eps=1e-10):
    d = np.linalg.norm(ref - x, axis=1)
    r = np.linalg.norm(ref, axis=1)
    return d / (r + np.mean(r) * eps)


def test_root_contains_all_particles():
    rng = np.random.default_rng(42)
    pos = rng.random((50, 3))
    t = oracle.OracleTree(pos, rng.random(50) * 0.005 + 0.005, 1.0, False, mpl=1)
    assert t.arr["count"][0] == 50


def test_node_count_grows_with_smaller_mpl():
    rng = np.random.default_rng(0)
    pos = rng.random((80, 3))
    counts = [oracle.OracleTree(pos, np.full(80, 0.005), 1.0, False, mpl=m).node_count
              for m in (8, 1)]
    assert counts[1] > counts[0]


@pytest.mark.parametrize("mpl", [1, 3, 16])
def test_structural_invariants(mpl):
    rng = np.random.default_rng(mpl)
    n = 700
    pos = rng.random((n, 3))
    t = oracle.OracleTree(pos, np.full(n, 1e-3), 1.0, True, mpl=mpl)
    a, nc = t.arr, t.node_count
    fc, cnt, ls = a["first_child"][:nc], a["count"][:nc], a["leaf_start"][:nc]
    leaves = fc < 0
    assert cnt[leaves].sum() == n == t.leaf_slot_count
    assert np.all(cnt[leaves] <= mpl)
    assert sorted(a["leaf_indices"][:n].tolist()) == list(range(n))
    # every non-empty leaf's slot range is contiguous and its particles lie in its box
    for i in np.nonzero(leaves & (cnt > 0))[0]:
        idx = a["leaf_indices"][ls[i]: ls[i] + cnt[i]]
        lo, hi = a["bmin"][3 * i: 3 * i + 3], a["bmax"][3 * i: 3 * i + 3]
        assert np.all(pos[idx] >= lo) and np.all(pos[idx] <= hi)
    # nodes are created 8 at a time after the root
    assert (nc - 1) % 8 == 0 and t.parent_count == (nc - 1) // 8
    # the stackless walk from the root visits every non-empty leaf exactly once
    seen, node = [], 0
    while node >= 0:
        if a["count"][node] == 0:
            node = a["next"][node]
        elif a["first_child"][node] < 0:
            seen.append(node)
            node = a["next"][node]
        else:
            node = a["first_child"][node]
    assert sorted(seen) == sorted(np.nonzero(leaves & (cnt > 0))[0].tolist())
    # R^3 aggregates of the root
    np.testing.assert_allclose(a["r3sum"][0], n * 1e-9, rtol=1e-12)
    np.testing.assert_allclose(a["r3c"][:3], pos.sum(0) * 1e-9, rtol=1e-12)


def test_capacity_clamp_stops_splitting():
    rng = np.random.default_rng(3)
    n = 500
    pos = rng.random((n, 3))
    full = oracle.OracleTree(pos, np.full(n, 1e-3), 1.0, False, mpl=1)
    small = oracle.OracleTree(pos, np.full(n, 1e-3), 1.0, False, mpl=1, max_nodes=201)
    assert small.node_count <= 201 < full.node_count
    a = small.arr
    leaves = a["first_child"][: small.node_count] < 0
    assert a["count"][: small.node_count][leaves].sum() == n


@pytest.mark.parametrize("periodic", [False, True])
def test_theta_zero_equals_direct(periodic):
    rng = np.random.default_rng(11)
    n, L = 400, 1e-3
    pos = rng.random((n, 3)) * L
    r = np.full(n, 3e-6)
    t = oracle.OracleTree(pos, r, L, periodic, mpl=4)
    Fd = oracle.direct_forces(pos, r, M, L, periodic)
    np.testing.assert_allclose(t.forces(0.0, M), Fd, rtol=1e-9, atol=1e-12 * np.abs(Fd).max())
    Vd = oracle.direct_convection(pos, Fd, ETA, L, periodic)
    np.testing.assert_allclose(t.stokeslet(0.0, Fd, ETA), Vd, rtol=1e-9,
                               atol=1e-12 * np.abs(Vd).max())
    sf = oracle.stokes_factor(0.065, 0.001)
    np.testing.assert_allclose(t.total_velocity(0.0, Fd, ETA, sf), Fd / (sf * r[:, None]) + Vd,
                               rtol=1e-9, atol=1e-12 * np.abs(Vd).max())


def test_reference_error_bands_and_theta_monotonicity():
    """test_force_tree.py:24-86: N=30, theta=0.3 -> mean err < 5 % (forces),
    < 10 % (convection); error does not grow as theta shrinks."""
    rng = np.random.default_rng(42)
    n, L = 30, 0.01
    pos = rng.random((n, 3)) * L
    r = rng.random(n) * 30e-6 + 20e-6
    Fd = oracle.direct_forces(pos, r, M, L, False)
    Vd = oracle.direct_convection(pos, Fd, ETA, L, False)
    t = oracle.OracleTree(pos, r, L, False, mpl=1)
    F = t.forces(0.3, M)
    assert _rel(Fd, F).mean() < 0.05
    assert _rel(Vd, t.stokeslet(0.3, F, ETA)).mean() < 0.1
    prev = np.inf
    for th in (0.8, 0.5, 0.3):
        e = _rel(Fd, t.forces(th, M)).mean()
        assert e <= prev + 1e-4
        prev = e


def test_dipole_correction_improves_monopole():
    """The dipole term of accepted nodes brings the Stokeslet sum closer to
    the direct sum than the monopole alone would (flat_tree.py:1215-1265)."""
    rng = np.random.default_rng(5)
    n, L = 2000, 1e-3
    pos = rng.random((n, 3)) * L
    r = np.full(n, 2e-6)
    F = rng.normal(size=(n, 3)) * 1e-12
    Vd = oracle.direct_convection(pos, F, ETA, L, False)
    t = oracle.OracleTree(pos, r, L, False, mpl=8)
    e_small = _rel(Vd, t.stokeslet(0.3, F, ETA)).mean()
    e_big = _rel(Vd, t.stokeslet(0.7, F, ETA)).mean()
    assert e_small < e_big < 0.5
