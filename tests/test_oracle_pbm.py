"""Pins the oracle's population-balance collision counts
(oracle.direct_collision_counts, oracle.OracleTree.collision_counts, restating
pbm/collision_frequency/direct_collision_freq.py:15-41 and
tree_collision_freq.py:62-175) and the host-side VolumeGrid.

* the reference's own cases (tests/pbm/collision_frequency/
  test_tree_collision_freq.py:16-92, test_direct_collision_freq.py:11-38,
  tests/pbm/test_volume_grid.py);
* a pure-Python restatement of both loops on small systems;
* brute force: for open boxes the pruned walk counts exactly the pairs with
  |d|^2 <= (r_i + r_j)^2; with periodicity it can only miss pairs (the
  reference's box distance overestimates wrapped gaps), and a constructed
  case shows such a miss.
"""

from __future__ import annotations

import logging

import numpy as np
import pytest

import oracle
from pbm import VolumeGrid
from pbm.volume_grid import sphere_volumes


def _bins(grid, radii):
    return grid.bin_indices(sphere_volumes(radii)).astype(np.int32)


def _tree_counts(pos, radii, L, periodic, grid, mpl=1):
    t = oracle.OracleTree(pos, radii, L, periodic, mpl=mpl)
    return t.collision_counts(_bins(grid, radii), grid.n_bins)


# ---------------------------------------------------------------- the reference's cases


def test_tree_symmetric_matrix_with_contact():
    pos = np.array([[0.10, 0.10, 0.10], [0.12, 0.10, 0.10], [0.50, 0.50, 0.50],
                    [0.80, 0.10, 0.10], [0.10, 0.80, 0.10], [0.10, 0.10, 0.80]])
    radii = np.full(6, 0.025)
    g = VolumeGrid.from_radii_range(0.01, 0.05, 4)
    M = _tree_counts(pos, radii, 1.0, False, g)
    assert M.shape == (4, 4)
    np.testing.assert_array_equal(M, M.T)
    assert M.sum() == 2  # exactly the pair (0, 1), in [a][b] and [b][a]


def test_tree_no_overlap_zero_matrix():
    pos = np.array([[0.10, 0.10, 0.10], [0.90, 0.10, 0.10], [0.10, 0.90, 0.10],
                    [0.10, 0.10, 0.90]])
    g = VolumeGrid.from_radii_range(0.005, 0.05, 4)
    assert _tree_counts(pos, np.full(4, 0.01), 1.0, False, g).sum() == 0


def test_tree_periodic_boundary_records_contact():
    pos = np.array([[0.01, 0.5, 0.5], [0.99, 0.5, 0.5]])
    radii = np.full(2, 0.05)
    g = VolumeGrid.from_radii_range(0.02, 0.1, 3)
    assert _tree_counts(pos, radii, 1.0, False, g).sum() == 0
    assert _tree_counts(pos, radii, 1.0, True, g).sum() > 0


def test_direct_reference_cases():
    g = VolumeGrid.from_radii_range(1e-6, 5e-6, 5)
    far = oracle.direct_collision_counts(np.array([[0.0, 0, 0], [1.0, 0, 0]]),
                                         np.array([1e-6, 1e-6]), 0.0, False,
                                         _bins(g, [1e-6, 1e-6]), 5)
    assert far.sum() == 0
    r = np.array([2e-6, 2e-6])
    near = oracle.direct_collision_counts(np.array([[0.0, 0, 0], [3e-6, 0, 0]]), r, 0.0, False,
                                          _bins(g, r), 5)
    assert near.sum() == 2
    np.testing.assert_array_equal(near, near.T)
    L = 1e-5
    pos = np.array([[1e-7, 0.0, 0.0], [L - 1e-7, 0.0, 0.0]])
    r = np.array([3e-7, 3e-7])
    assert oracle.direct_collision_counts(pos, r, L, False, _bins(g, r), 5).sum() == 0
    assert oracle.direct_collision_counts(pos, r, L, True, _bins(g, r), 5).sum() == 2


# ---------------------------------------------------------------- pure-Python restatements


def _py_direct(pos, radii, L, periodic, bins, nb):
    M = np.zeros((nb, nb), np.int64)
    for i in range(len(radii)):
        for j in range(i + 1, len(radii)):
            d = pos[j] - pos[i]
            if periodic and L > 0:
                d = d - L * np.round(d / L)
            if np.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) <= radii[i] + radii[j]:
                M[bins[i], bins[j]] += 1
                M[bins[j], bins[i]] += 1
    return M


def _py_tree(t, pos, radii, L, periodic, bins, nb):
    a = t.arr
    M = np.zeros((nb, nb), np.int64)

    def gap(node, p):
        s = 0.0
        for ax in range(3):
            lo, hi = a["bmin"][3 * node + ax], a["bmax"][3 * node + ax]
            d = lo - p[ax] if p[ax] < lo else (p[ax] - hi if p[ax] > hi else 0.0)
            if periodic and L > 0 and d > L * 0.5:
                d = L - d
            s += d * d
        return s

    for i in range(len(radii)):
        node = 0
        while node >= 0:
            cnt = a["count"][node]
            if cnt == 0:
                node = a["next"][node]
                continue
            if a["first_child"][node] < 0:
                ls = a["leaf_start"][node]
                for k in range(cnt):
                    j = a["leaf_indices"][ls + k]
                    if j > i:
                        d = pos[j] - pos[i]
                        if periodic:
                            d = np.where(np.abs(d) > L * 0.5, d - np.sign(d) * L, d)
                        if d[0] * d[0] + d[1] * d[1] + d[2] * d[2] <= (radii[i] + radii[j]) ** 2:
                            M[bins[i], bins[j]] += 1
                            M[bins[j], bins[i]] += 1
                node = a["next"][node]
                continue
            mc = radii[i] + a["maxr"][node]
            node = a["next"][node] if gap(node, pos[i]) > mc * mc else a["first_child"][node]
    return M


def _dense_system(n, seed, L=1.0):
    rng = np.random.default_rng(seed)
    pos = rng.random((n, 3)) * L
    radii = rng.uniform(0.02, 0.06, n) * L
    return pos, radii


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("mpl", [1, 4])
def test_oracle_matches_python_loops(periodic, mpl):
    pos, radii = _dense_system(70, seed=3 + mpl)
    g = VolumeGrid.from_radii_range(0.02, 0.06, 5)
    b = _bins(g, radii)
    t = oracle.OracleTree(pos, radii, 1.0, periodic, mpl=mpl)
    np.testing.assert_array_equal(t.collision_counts(b, 5),
                                  _py_tree(t, pos, radii, 1.0, periodic, b, 5))
    np.testing.assert_array_equal(oracle.direct_collision_counts(pos, radii, 1.0, periodic, b, 5),
                                  _py_direct(pos, radii, 1.0, periodic, b, 5))


# ---------------------------------------------------------------- brute force


def _brute(pos, radii, L, periodic, bins, nb, squared):
    d = pos[None, :, :] - pos[:, None, :]
    if periodic:
        d = np.where(np.abs(d) > L * 0.5, d - np.sign(d) * L, d)
    d2 = (d * d).sum(-1)
    c = radii[:, None] + radii[None, :]
    hit = (d2 <= c * c) if squared else (np.sqrt(d2) <= c)
    i, j = np.nonzero(np.triu(hit, 1))
    M = np.zeros((nb, nb), np.int64)
    np.add.at(M, (bins[i], bins[j]), 1)
    np.add.at(M, (bins[j], bins[i]), 1)
    return M


@pytest.mark.parametrize("mpl", [1, 2, 8])
def test_open_tree_counts_equal_brute_force(mpl):
    pos, radii = _dense_system(900, seed=mpl)
    g = VolumeGrid.from_radii_range(0.02, 0.06, 7)
    b = _bins(g, radii)
    M = _tree_counts(pos, radii, 1.0, False, g, mpl=mpl)
    B = _brute(pos, radii, 1.0, False, b, 7, squared=True)
    assert B.sum() > 100
    np.testing.assert_array_equal(M, B)
    np.testing.assert_array_equal(oracle.direct_collision_counts(pos, radii, 1.0, False, b, 7),
                                  _brute(pos, radii, 1.0, False, b, 7, squared=False))


def test_periodic_tree_counts_never_exceed_brute_force():
    pos, radii = _dense_system(900, seed=11)
    g = VolumeGrid.from_radii_range(0.02, 0.06, 7)
    b = _bins(g, radii)
    M = _tree_counts(pos, radii, 1.0, True, g, mpl=2)
    B = _brute(pos, radii, 1.0, True, b, 7, squared=True)
    assert (M <= B).all() and M.sum() > 0


def test_periodic_box_distance_rule_misses_a_wrapped_contact():
    # droplets 0 and 1 touch across x = 0 (MIC distance 0.02 < 0.1), but
    # droplet 1 sits in an internal node (it shares the octant with droplet 2)
    # whose box gap from droplet 0 is 0.49 <= L/2, so the walk prunes it
    pos = np.array([[0.01, 0.25, 0.25], [0.99, 0.25, 0.25], [0.60, 0.25, 0.25]])
    radii = np.full(3, 0.05)
    g = VolumeGrid.from_radii_range(0.02, 0.1, 3)
    b = _bins(g, radii)
    assert _tree_counts(pos, radii, 1.0, True, g, mpl=1).sum() == 0
    assert _brute(pos, radii, 1.0, True, b, 3, squared=True).sum() == 2
    assert oracle.direct_collision_counts(pos, radii, 1.0, True, b, 3).sum() == 2


# ---------------------------------------------------------------- VolumeGrid (host)


def test_volume_grid_reference_cases(caplog):
    g = VolumeGrid(1e-18, 1e-12, 30)
    np.testing.assert_allclose(g.centers, np.sqrt(g.edges[:-1] * g.edges[1:]))
    gl = VolumeGrid(1.0, 10.0, 9, spacing="linear")
    np.testing.assert_allclose(gl.centers, 0.5 * (gl.edges[:-1] + gl.edges[1:]))
    np.testing.assert_allclose(VolumeGrid(1.0, 1000.0, 6, "logarithmic").edges,
                               VolumeGrid(1.0, 1000.0, 6, "geometric").edges)
    g2 = VolumeGrid.from_radii_range(2.5e-6, 7.5e-6, 20)
    assert g2.edges[0] == pytest.approx((4 / 3) * np.pi * 2.5e-6**3)
    assert g2.edges[-1] == pytest.approx((4 / 3) * np.pi * 7.5e-6**3)
    for args in ((-1.0, 1.0, 5), (1.0, 1.0, 5)):
        with pytest.raises(ValueError, match="0 < v_min < v_max"):
            VolumeGrid(*args)
    with pytest.raises(ValueError, match="n_bins >= 2"):
        VolumeGrid(1.0, 10.0, 1)
    with pytest.raises(ValueError, match="spacing"):
        VolumeGrid(1.0, 10.0, 5, spacing="unsupported")
    g3 = VolumeGrid(1e-18, 1e-12, 10)
    assert g3.bin_index(1e-20) == 0 and g3.bin_index(1e-10) == 9
    g4 = VolumeGrid(1.0, 10.0, 5)
    caplog.clear()
    with caplog.at_level(logging.WARNING, logger="pbm.volume_grid"):
        idx = g4.bin_indices(np.array([0.5, 5.0, 50.0]))
        g4.bin_index(0.1)
        g4.bin_index(100.0)
    assert idx[0] == 0 and idx[2] == 4
    msgs = [r.getMessage() for r in caplog.records]
    assert sum("v_max" in m for m in msgs) == 1 and sum("v_min" in m for m in msgs) == 1
    radii = np.random.default_rng(0).uniform(2.5e-6, 7.5e-6, 1000)
    assert VolumeGrid.from_radii_range(2.5e-6, 7.5e-6, 30).histogram(radii).sum() == 1000
    # classes are the right-closed searchsorted intervals of the edges
    e = g2.edges
    np.testing.assert_array_equal(g2.bin_indices(e[:-1]), np.arange(20))
