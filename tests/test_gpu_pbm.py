"""GPU collision counts between volume classes (``dem_octree_collision_counts``,
``dem_direct_collision_counts`` behind pbm.collision_frequency) against the
oracle: the reference's own cases, the periodic box-distance miss, random
dense systems (open / periodic, several leaf sizes) bit-exact, and at
N = 65536 the tree walk against the all-pairs kernel (open box: equal;
periodic: never more) and against the oracle's walk."""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from dem.octree.flat_tree import FlatOctree
from pbm import VolumeGrid
from pbm.collision_frequency import DirectCollisionFrequency, TreeCollisionFrequency
from pbm.volume_grid import sphere_volumes

pytestmark = pytest.mark.gpu


def _bins(g, radii):
    return g.bin_indices(sphere_volumes(radii)).astype(np.int32)


def _tree(pos, radii, L, periodic, mpl=1):
    t = FlatOctree(theta=0.5, mpl=mpl, num_particles=len(radii))
    t.build(pos, radii, L=L, periodic=periodic)
    return t


def test_reference_tree_cases(gpu):
    pos = np.array([[0.10, 0.10, 0.10], [0.12, 0.10, 0.10], [0.50, 0.50, 0.50],
                    [0.80, 0.10, 0.10], [0.10, 0.80, 0.10], [0.10, 0.10, 0.80]])
    radii = np.full(6, 0.025)
    g = VolumeGrid.from_radii_range(0.01, 0.05, 4)
    M = TreeCollisionFrequency(_tree(pos, radii, 1.0, False), g, 6).compute(pos, radii)
    assert M.shape == (4, 4) and M.dtype == np.float64
    np.testing.assert_array_equal(M, M.T)
    assert M.sum() == 2

    pos = np.array([[0.10, 0.10, 0.10], [0.90, 0.10, 0.10], [0.10, 0.90, 0.10],
                    [0.10, 0.10, 0.90]])
    g = VolumeGrid.from_radii_range(0.005, 0.05, 4)
    r = np.full(4, 0.01)
    assert TreeCollisionFrequency(_tree(pos, r, 1.0, False), g, 4).compute(pos, r).sum() == 0

    pos = np.array([[0.01, 0.5, 0.5], [0.99, 0.5, 0.5]])
    r = np.full(2, 0.05)
    g = VolumeGrid.from_radii_range(0.02, 0.1, 3)
    assert TreeCollisionFrequency(_tree(pos, r, 1.0, False), g, 2).compute(pos, r).sum() == 0
    assert TreeCollisionFrequency(_tree(pos, r, 1.0, True), g, 2).compute(pos, r).sum() > 0


def test_reference_direct_cases(gpu):
    g = VolumeGrid.from_radii_range(1e-6, 5e-6, 5)
    cf = DirectCollisionFrequency(g)
    assert cf.compute(np.array([[0.0, 0, 0], [1.0, 0, 0]]), np.array([1e-6, 1e-6])).sum() == 0
    M = cf.compute(np.array([[0.0, 0, 0], [3e-6, 0, 0]]), np.array([2e-6, 2e-6]))
    assert M.sum() == 2
    np.testing.assert_array_equal(M, M.T)
    L = 1e-5
    pos = np.array([[1e-7, 0.0, 0.0], [L - 1e-7, 0.0, 0.0]])
    r = np.array([3e-7, 3e-7])
    assert cf.compute(pos, r, L=L, periodic=False).sum() == 0
    assert cf.compute(pos, r, L=L, periodic=True).sum() == 2


def test_periodic_box_distance_miss_reproduced(gpu):
    pos = np.array([[0.01, 0.25, 0.25], [0.99, 0.25, 0.25], [0.60, 0.25, 0.25]])
    r = np.full(3, 0.05)
    g = VolumeGrid.from_radii_range(0.02, 0.1, 3)
    assert TreeCollisionFrequency(_tree(pos, r, 1.0, True), g, 3).compute(pos, r).sum() == 0
    assert DirectCollisionFrequency(g).compute(pos, r, L=1.0, periodic=True).sum() == 2


def _dense(n, seed, L=1.0, lo=0.02, hi=0.06):
    rng = np.random.default_rng(seed)
    return rng.random((n, 3)) * L, rng.uniform(lo, hi, n) * L


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("mpl", [1, 4, 16])
def test_random_systems_bit_exact(gpu, periodic, mpl):
    pos, r = _dense(3000, seed=20 + mpl + (7 if periodic else 0), lo=0.005, hi=0.02)
    g = VolumeGrid.from_radii_range(0.005, 0.02, 9)
    b = _bins(g, r)
    M = TreeCollisionFrequency(_tree(pos, r, 1.0, periodic, mpl), g, 3000).compute(pos, r)
    ref = oracle.OracleTree(pos, r, 1.0, periodic, mpl=mpl).collision_counts(b, 9)
    assert ref.sum() > 100
    np.testing.assert_array_equal(M.astype(np.int64), ref)
    D = DirectCollisionFrequency(g).compute(pos, r, L=1.0, periodic=periodic)
    np.testing.assert_array_equal(D.astype(np.int64),
                                  oracle.direct_collision_counts(pos, r, 1.0, periodic, b, 9))


@pytest.mark.parametrize("periodic", [False, True])
def test_n65536_tree_against_all_pairs_and_oracle(gpu, periodic):
    from dem.workloads import make_system

    s = make_system(65536, seed=0, phi=0.02)
    # grow the radii so that many droplets touch (the generator keeps them apart)
    r = s.radii * 1.6
    g = VolumeGrid.from_radii_range(float(r.min()), float(r.max()), 12)
    t = _tree(s.positions, r, s.L, periodic, mpl=16)
    M = TreeCollisionFrequency(t, g, s.n).compute(s.positions, r).astype(np.int64)
    D = DirectCollisionFrequency(g).compute(s.positions, r, L=s.L,
                                            periodic=periodic).astype(np.int64)
    print(f"\n[N=65536 periodic={periodic}] tree pairs {M.sum() // 2}, all-pairs {D.sum() // 2}")
    assert D.sum() > 1000
    if periodic:
        assert (M <= D).all()
    else:
        np.testing.assert_array_equal(M, D)
    ref = oracle.OracleTree(s.positions, r, s.L, periodic, mpl=16).collision_counts(_bins(g, r), 12)
    np.testing.assert_array_equal(M, ref)
