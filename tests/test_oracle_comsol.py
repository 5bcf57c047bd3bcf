"""Pins the COMSOL periodic-image correction: the host grid build
(``COMSOLLatticeCorrection``, restating
dem/periodic_correction/comsol_lattice_correction.py:59-297) and the oracle's
trilinear / permuted lookups (oracle/dem_oracle.c ``corr_pair``, restating
dem/force_calculator/direct_droplet_force_calculator.py:139-181, :413-464).

What holds them, independently of each other:
* the reference's own tests (tests/periodic_correction/
  test_comsol_lattice_correction.py:19-142), ported as they are;
* the Stokeslet parities on the whole grid (a wrong mirror sign or a
  transposed grid axis breaks them by O(1));
* the grid reproduces the COMSOL samples it was resampled from, to
  interpolation accuracy (a transposed axis reads other samples);
* the oracle's trilinear lookup equals scipy's ``RegularGridInterpolator``
  (a library routine) at grid nodes and at random points;
* the oracle's three permuted lookups, read back as a 3x3 tensor through the
  oracle's Oseen sum, equal the scipy evaluation of the reference's
  ``evaluate`` (:248-297) and are symmetric, G_ab = G_ba, as a mobility
  tensor must be (a wrong permutation breaks that by O(1)).
"""

from __future__ import annotations

import numpy as np
import pytest
from scipy.interpolate import RegularGridInterpolator

import oracle
from dem.periodic_correction import COMSOLLatticeCorrection


@pytest.fixture(scope="module")
def correction() -> COMSOLLatticeCorrection:
    return COMSOLLatticeCorrection.load_default()


def _oracle_corr(c: COMSOLLatticeCorrection, L_sim: float, eta_oil: float) -> oracle.Correction:
    gd = c.get_grid_data()
    L_ratio = gd["L_comsol"] / L_sim
    eta_ratio = gd["eta_comsol"] / eta_oil
    w_self = float(c.evaluate(np.zeros((1, 3)))[0, 2])
    return oracle.Correction(gd["grid_u"], gd["grid_v"], gd["grid_w"], gd["grid_min"],
                             gd["grid_dx"], gd["Fz_comsol"], L_ratio, eta_ratio,
                             w_self / gd["Fz_comsol"] * eta_ratio * L_ratio)


# --- the reference's tests (test_comsol_lattice_correction.py:19-142) ----------


def test_load_default(correction):
    assert isinstance(correction, COMSOLLatticeCorrection)
    assert correction.grid_resolution > 0
    assert correction.L_comsol > 0 and correction.Fz_comsol > 0


def test_grid_shapes_and_no_nan(correction):
    n = correction.grid_resolution
    for g in (correction.grid_u, correction.grid_v, correction.grid_w):
        assert g.shape == (n, n, n)
        assert not np.any(np.isnan(g))


def test_evaluate_shapes_and_full_convolution(correction):
    r = np.array([[0.0, 0.0, 0.0], [correction.L_comsol * 0.1, 0.0, 0.0]])
    assert correction.evaluate(r).shape == (2, 3)
    r = np.array([[correction.L_comsol * 0.1, 0.0, 0.0]])
    F = np.array([[0.0, 0.0, correction.Fz_comsol]])
    np.testing.assert_allclose(correction.evaluate(r, F), correction.evaluate(r), atol=1e-12)


def test_evaluate_outside_grid_is_zero(correction):
    far = correction.L_comsol * 10
    np.testing.assert_allclose(correction.evaluate(np.array([[far, far, far]])), 0.0)


def test_grid_data(correction):
    d = correction.get_grid_data()
    for key in ("grid_u", "grid_v", "grid_w", "grid_min", "grid_max", "grid_dx",
                "grid_resolution", "L_comsol", "Fz_comsol", "eta_comsol"):
        assert key in d
    np.testing.assert_allclose(d["grid_min"], -d["grid_max"])
    n = d["grid_resolution"]
    np.testing.assert_allclose(d["grid_dx"], (d["grid_max"] - d["grid_min"]) / (n - 1), rtol=1e-12)


def test_mirror_symmetry_points(correction):
    x = correction.L_comsol * 0.15
    e = correction.evaluate
    np.testing.assert_allclose(e(np.array([[x, 0, 0]]))[0, 0], -e(np.array([[-x, 0, 0]]))[0, 0],
                               atol=1e-10)
    np.testing.assert_allclose(e(np.array([[0, x, 0]]))[0, 1], -e(np.array([[0, -x, 0]]))[0, 1],
                               atol=1e-10)
    np.testing.assert_allclose(e(np.array([[x, x, 0]]))[0, 2], e(np.array([[-x, -x, 0]]))[0, 2],
                               atol=1e-10)


# --- the grid as a whole ---------------------------------------------------------


def test_grid_parities_hold_on_every_node(correction):
    """u odd in x / even in y, v even in x / odd in y, w even in both.  The
    grid axis is symmetric about 0, so reversing an axis mirrors the nodes.
    scipy's triangulation of the mirrored samples is not itself mirror
    symmetric, so u keeps a few-percent asymmetry near the droplet; a wrong
    mirror sign or a transposed axis gives O(1)."""
    u, v, w = correction.grid_u, correction.grid_v, correction.grid_w
    q = lambda a, g: np.quantile(np.abs(a) / np.abs(g).max(), 0.99)  # noqa: E731
    assert q(u + u[::-1], u) < 0.15 and q(u - u[:, ::-1], u) < 0.15
    assert q(v - v[::-1], v) < 1e-3 and q(v + v[:, ::-1], v) < 1e-3
    assert np.abs(w - w[::-1]).max() < 1e-12 * np.abs(w).max()
    assert np.abs(w - w[:, ::-1]).max() < 1e-12 * np.abs(w).max()


def test_grid_reproduces_comsol_samples(correction):
    """Trilinear lookups of the resampled grid at the COMSOL sample points
    give the sampled velocities to interpolation accuracy (relative to each
    component's largest sample): median < 6e-4, 90th percentile < 2e-2.
    With the grid's x and y axes swapped the figures are 1.0e-3 … 1.5e-3 and
    3.4e-2 … 4.3e-2."""
    data = np.loadtxt(COMSOLLatticeCorrection._DEFAULT_DATA_DIR + "/U_lattice.txt",
                      comments="%")
    pts, raw = data[:, :3], data[:, 3:6]
    err = np.abs(correction.evaluate(pts) - raw) / np.abs(raw).max(axis=0)
    assert np.all(np.median(err, axis=0) < 6e-4)
    assert np.all(np.quantile(err, 0.9, axis=0) < 2e-2)


# --- the oracle's lookups ------------------------------------------------------


def test_oracle_trilinear_matches_regular_grid_interpolator(correction):
    gd = correction.get_grid_data()
    oc = _oracle_corr(correction, gd["L_comsol"], gd["eta_comsol"])
    axis = correction.grid_coords
    n = len(axis)
    rng = np.random.default_rng(0)
    nodes = rng.integers(0, n, size=(500, 3))
    pts = axis[nodes]
    ref = np.stack([correction.grid_u[tuple(nodes.T)], correction.grid_v[tuple(nodes.T)],
                    correction.grid_w[tuple(nodes.T)]], axis=1)
    scale = np.abs(ref).max()
    np.testing.assert_allclose(oc.eval(pts), ref, rtol=1e-12, atol=1e-13 * scale)
    inside = (rng.random((4000, 3)) - 0.5) * (axis[-1] - axis[0])
    rgi = [RegularGridInterpolator((axis, axis, axis), g, method="linear")
           for g in (correction.grid_u, correction.grid_v, correction.grid_w)]
    ref = np.stack([f(inside) for f in rgi], axis=1)
    np.testing.assert_allclose(oc.eval(inside), ref, rtol=1e-10, atol=1e-12 * scale)


def _oracle_tensor(oc, pts, L_sim, eta_oil):
    """3x3 correction tensor at offsets pts (simulation frame, |pts| < L/2)
    read back through the oracle's Oseen sum: a source at the origin carrying
    a unit force along b, targets at pts with zero force; the sum with the
    correction minus the sum without it is column b."""
    pos = np.vstack([np.full((1, 3), L_sim / 2), L_sim / 2 + pts])
    cols = []
    for b in range(3):
        F = np.zeros_like(pos)
        F[0, b] = 1.0
        eta = oracle.eta_const(eta_oil)
        with_c = oracle.direct_convection(pos, F, eta, L_sim, True, oc)
        without = oracle.direct_convection(pos, F, eta, L_sim, True)
        cols.append((with_c - without)[1:])
    return np.stack(cols, axis=2)  # [k, a, b]: velocity component a from force b


def test_oracle_permuted_lookups_equal_reference_evaluate(correction):
    L_sim, eta_oil = 4.2e-4, 0.05
    oc = _oracle_corr(correction, L_sim, eta_oil)
    gd = correction.get_grid_data()
    L_ratio, eta_ratio = gd["L_comsol"] / L_sim, gd["eta_comsol"] / eta_oil
    rng = np.random.default_rng(1)
    pts = (rng.random((300, 3)) - 0.5) * 0.98 * L_sim
    T = _oracle_tensor(oc, pts, L_sim, eta_oil)
    # the reference's evaluate(r_rel, F_j) scaled as its kernel scales the
    # lookups (direct...py:449-457): times eta_ratio * L_ratio
    for b in range(3):
        F = np.zeros((len(pts), 3))
        F[:, b] = 1.0
        ref = correction.evaluate(pts * L_ratio, F) * eta_ratio * L_ratio
        np.testing.assert_allclose(T[:, :, b], ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())


def test_oracle_correction_tensor_is_symmetric(correction):
    L_sim, eta_oil = 4.2e-4, 0.05
    oc = _oracle_corr(correction, L_sim, eta_oil)
    rng = np.random.default_rng(2)
    pts = (rng.random((1000, 3)) - 0.5) * 0.98 * L_sim
    T = _oracle_tensor(oc, pts, L_sim, eta_oil)
    m = np.abs(T).max()
    for a, b in ((0, 1), (0, 2), (1, 2)):
        # measured 0.012-0.015; a permutation that keeps (u, v, w) unswapped: 0.78
        assert np.median(np.abs(T[:, a, b] - T[:, b, a])) < 0.05 * m


def test_oracle_self_term(correction):
    """The self term adds w(0,0,0)/Fz * eta_ratio * L_ratio times the droplet's
    own force (direct...py:101-108, :459-464): a lone droplet's velocity."""
    L_sim, eta_oil = 4.2e-4, 0.05
    oc = _oracle_corr(correction, L_sim, eta_oil)
    gd = correction.get_grid_data()
    w0 = float(correction.evaluate(np.zeros((1, 3)))[0, 2])
    coeff = w0 / gd["Fz_comsol"] * gd["eta_comsol"] / eta_oil * gd["L_comsol"] / L_sim
    F = np.array([[1e-12, -2e-12, 3e-12]])
    V = oracle.direct_convection(np.full((1, 3), 1e-4), F, oracle.eta_const(eta_oil), L_sim, True,
                                 oc)
    np.testing.assert_allclose(V, coeff * F, rtol=1e-14)
