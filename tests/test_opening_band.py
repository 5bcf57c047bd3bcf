"""The walks' FP32 opening test decides only outside an error band, and then
must agree with the reference's float64 test (flat_tree.py:871: s^2 <
theta^2 D^2 and D^2 > 0).  This restates the device test in numpy float32
(fixed-point centre and target quantized to the nearest unit, int32
differences, float32 squares and sums) and checks, on separations crowded
around the opening boundary with random quantization offsets, that every
decided case equals the float64 decision and that the band stays narrow
(the float64 fallback is rare away from the boundary)."""

from __future__ import annotations

import numpy as np

F = np.float32
HALF_BAND = F(2147481600.0)  # 2^31 - 2^11


def fast_test(Rx, Ry, Rz, th2f, band_s, s2f, sf, rel=2.0**-17, const=8.0):
    """1 take, 0 open, -1 undecided (float32 arithmetic, as the walks';
    band = rel s^2 + band_s s + const)."""
    d2 = (Rx * Rx + Ry * Ry).astype(F) + (Rz * Rz).astype(F)
    m = (th2f * d2).astype(F) - s2f
    band = (F(rel) * s2f + (band_s * sf + F(const))).astype(F)
    amax = np.maximum(np.abs(Rx), np.maximum(np.abs(Ry), np.abs(Rz)))
    out = np.where(m > 0, 1, 0)
    return np.where((np.abs(m) <= band) | (amax >= HALF_BAND), -1, out)


def _case(rng, n, theta, s_units, spread):
    """Targets at separations |D| = s/theta * (1 + spread*u) in random
    directions; true (unquantized) centre and target in units of the box
    (2^32 per side, periodic)."""
    u = rng.uniform(-1, 1, n)
    dist = s_units / theta * (1.0 + spread * u)
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    c = rng.uniform(0, 2.0**32, (n, 3))
    x = c - v * dist[:, None]
    return c, x


def _decide(c, x, s_units, theta):
    L = 2.0**32
    # float64 reference test (the minimum image of centre - target)
    d = c - x
    d = np.where(np.abs(d) > L / 2, d - np.sign(d) * L, d)
    d2 = (d * d).sum(1)
    exact = ((s_units * s_units < theta * theta * d2) & (d2 > 0)).astype(int)
    # fixed point: nearest unit, wrapped into the box, int32 difference
    qc = np.mod(np.rint(c), L).astype(np.uint64).astype(np.uint32)
    qx = np.mod(np.rint(x), L).astype(np.uint64).astype(np.uint32)
    R = (qc - qx).view(np.int32).astype(F)
    th2f = F(theta * theta)
    band_s = F(8.0 * theta * (1 + 2.0**-20))
    fast = fast_test(R[:, 0], R[:, 1], R[:, 2], th2f, band_s, F(s_units * s_units),
                     F(s_units))
    return exact, fast


def test_decided_cases_equal_float64():
    rng = np.random.default_rng(0)
    total = undecided = 0
    for theta in (0.35, 0.5, 0.7, 1.0):
        # node sizes from the smallest fast-path node (2^22 units) to the box
        for s_units in (2.0**22, 3.1e6 * 2, 2.0**25, 1.3e9, 2.0**31):
            for spread in (1e-3, 1e-5, 1e-7):
                c, x = _case(rng, 20000, theta, s_units, spread)
                exact, fast = _decide(c, x, s_units, theta)
                dec = fast >= 0
                assert np.array_equal(fast[dec], exact[dec]), (theta, s_units, spread)
                total += len(fast)
                undecided += int((~dec).sum())
    # crowded within 1e-7..1e-3 of the boundary: the band leaves some open
    assert undecided < total


def test_band_is_narrow_away_from_the_boundary():
    rng = np.random.default_rng(1)
    for theta in (0.5, 0.7):
        for s_units in (2.0**22, 2.0**26, 2.0**30):
            c, x = _case(rng, 20000, theta, s_units, 0.5)
            exact, fast = _decide(c, x, s_units, theta)
            dec = fast >= 0
            assert np.array_equal(fast[dec], exact[dec])
            # undecided: within the band of the boundary or near half the box
            assert (~dec).mean() < 2e-3, (theta, s_units, (~dec).mean())


def test_half_box_components_are_undecided():
    # a component within 2^11 units of half the box may take the other image
    R = np.array([2.0**31 - 1024, 2.0**31 - 4096, -(2.0**31)], dtype=F)
    z = np.zeros(3, dtype=F)
    out = fast_test(R, z, z, F(0.25), F(4.0), F(1e10), F(1e5))
    assert list(out) == [-1, 1, -1]


def test_planted_missing_band_is_caught():
    """Without the band the FP32 test decides cases the float64 test decides
    the other way: the comparison above would see such a mistake."""
    rng = np.random.default_rng(2)
    theta, s_units = 0.5, 2.0**22
    c, x = _case(rng, 200000, theta, s_units, 1e-6)
    L = 2.0**32
    d = c - x
    d2 = (d * d).sum(1)
    exact = (s_units * s_units < theta * theta * d2).astype(int)
    qc = np.mod(np.rint(c), L).astype(np.uint64).astype(np.uint32)
    qx = np.mod(np.rint(x), L).astype(np.uint64).astype(np.uint32)
    R = (qc - qx).view(np.int32).astype(F)
    fast = fast_test(R[:, 0], R[:, 1], R[:, 2], F(0.25), F(0.0), F(s_units**2), F(0.0),
                     rel=0.0, const=0.0)
    dec = fast >= 0
    assert not np.array_equal(fast[dec], exact[dec])
