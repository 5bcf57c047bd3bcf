"""Far-field order of the oracle's Barnes-Hut walks.

A target sees a cluster of four sources of unequal radii from afar; the walk
takes the cluster's node as one far-field interaction.  Shrinking the
cluster about its R^3 centroid by 2 must cut the error against the direct
sum by 4: both far fields are exact to first order in the cluster size
(``_compute_forces_kernel``: monopole of R^3_sum at the R^3 centroid,
dem/octree/flat_tree.py:871-886; ``_compute_stokeslet_kernel``: monopole of
sum F at the centroid plus the dipole D = P - cm (x) F, :1154-1168 and
:1215-1265).  A dropped or wrong dipole term, such as the trace term s r,
leaves a first-order error (a factor of 2 per halving); so would a monopole
placed at the geometric centre.  The same cases run on the GPU walks in
tests/test_gpu_octree.py.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle

OFFSETS = np.array([[1.0, 0.3, -0.2], [-0.6, 1.0, 0.4], [0.2, -0.7, 1.0], [-0.5, -0.4, -0.9]])
RADII = np.array([0.004, 0.006, 0.009, 0.005, 0.007])
SRC_F = OFFSETS + np.array([0.3, -0.2, 0.5])  # radial part: a large dipole trace


def cluster_case(h: float, periodic: bool):
    """(positions, radii, forces): target first, then the cluster of size h.
    Open box: the cluster's accepted node is the level-1 cell [0.5, 1]^3.
    Periodic: every source's minimum image lies across the boundary, as does
    the centroid's; the accepted node is the level-3 cell [0.75, 0.875]^3."""
    tgt = np.array([[0.06, 0.09, 0.05]])
    c = np.array([0.78, 0.81, 0.83])
    pos = np.vstack([tgt, c + h * OFFSETS])
    F = np.vstack([np.zeros((1, 3)), SRC_F])
    return pos, RADII.copy(), F


def far_field_errors(h: float, periodic: bool, theta: float = 0.5):
    pos, r, F = cluster_case(h, periodic)
    t = oracle.OracleTree(pos, r, 1.0, periodic, mpl=1)
    Fd = oracle.direct_forces(pos, r, 1.0, 1.0, periodic)
    Ft = t.forces(theta, 1.0)
    Vd = oracle.direct_convection(pos, F, 1.0, 1.0, periodic)
    Vs = t.stokeslet(theta, F, 1.0)
    Vt = t.total_velocity(theta, F, 1.0, 1.0)
    # the migration term F/(sf r) of the target is zero: F_target = 0
    rel = lambda a, b: float(np.linalg.norm(a[0] - b[0]) / np.linalg.norm(b[0]))  # noqa: E731
    return rel(Ft, Fd), rel(Vs, Vd), rel(Vt, Vd)


@pytest.mark.parametrize("periodic", [False, True])
def test_far_field_error_is_second_order(periodic):
    hs = (0.02, 0.01, 0.005)
    errs = np.array([far_field_errors(h, periodic) for h in hs])
    # the cluster really is approximated (not opened down to the leaves)
    assert np.all(errs[0] > 1e-5)
    ratios = errs[:-1] / errs[1:]
    assert np.all((ratios > 3.6) & (ratios < 4.5)), ratios


def test_cluster_is_one_accepted_node():
    """At theta = 0 every node opens and the walk equals the direct sum, so
    the errors above come from the far field alone."""
    for periodic in (False, True):
        pos, r, F = cluster_case(0.02, periodic)
        t = oracle.OracleTree(pos, r, 1.0, periodic, mpl=1)
        Fd = oracle.direct_forces(pos, r, 1.0, 1.0, periodic)
        np.testing.assert_allclose(t.forces(0.0, 1.0)[0], Fd[0], rtol=1e-12)
        Vd = oracle.direct_convection(pos, F, 1.0, 1.0, periodic)
        np.testing.assert_allclose(t.stokeslet(0.0, F, 1.0)[0], Vd[0], rtol=1e-12)


@pytest.mark.parametrize("periodic", [False, True])
def test_planted_far_field_mistakes_fail_the_order_test(periodic):
    """Mistakes planted in the oracle walks leave a first-order (or worse)
    far-field error, which the order test above rejects; the untouched field
    keeps its second order."""
    hs = (0.02, 0.01, 0.005)

    def ratios(plant):
        oracle.set_plant(plant)
        try:
            e = np.array([far_field_errors(h, periodic) for h in hs])
        finally:
            oracle.set_plant(0)
        return e[:-1] / e[1:]

    second = lambda r: bool(np.all((r > 3.6) & (r < 4.5)))  # noqa: E731
    r = ratios(1)  # trace term of the Stokeslet dipole dropped
    assert second(r[:, 0]) and not second(r[:, 1]) and not second(r[:, 2]), r
    r = ratios(2)  # Stokeslet dipole dropped
    assert second(r[:, 0]) and not second(r[:, 1]) and not second(r[:, 2]), r
    r = ratios(4)  # force monopole at the geometric node centre
    assert not second(r[:, 0]) and second(r[:, 1]), r
    assert second(ratios(0)), "the oracle is restored"
