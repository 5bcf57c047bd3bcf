"""``dem_merge_plan`` (the host-side greedy merge planner of the batched
engine) against ``merge_plan_rows`` — the reference's greedy merge
(droplet_solution.py:208-262) restated on Python floats — bit for bit: chains,
droplets in several pairs (ties in the first column, visited in np.argsort
order) and single pairs.  Host only: no GPU needed."""

from __future__ import annotations

import ctypes

import numpy as np

from dem import _native
from dem.solution.droplet_solution import merge_plan_rows


def _random_systems(rng, count):
    out = []
    for _ in range(count):
        c = int(rng.integers(1, 40))
        n = int(rng.integers(4, 60))
        i = rng.integers(0, n - 1, c)
        j = i + 1 + rng.integers(0, n - 1 - i.max() + 1, c) % (n - i)
        pairs = np.unique(np.stack([i, np.minimum(j, n - 1)], 1), axis=0)
        pairs = pairs[pairs[:, 0] < pairs[:, 1]].astype(np.int32)
        if len(pairs) == 0:
            pairs = np.array([[0, 1]], dtype=np.int32)
        # per-droplet values (x, y, z, r), shared by every pair naming it
        vals_d = np.column_stack([rng.random((n, 3)) * 1e-3, rng.uniform(1e-6, 2e-5, n)])
        vals = np.concatenate([vals_d[pairs[:, 0]], vals_d[pairs[:, 1]]], axis=1)
        out.append((pairs, vals))
    return out


def test_native_plan_equals_merge_plan_rows():
    rng = np.random.default_rng(5)
    systems = _random_systems(rng, 300)
    off = np.concatenate([[0], np.cumsum([len(p) for p, _ in systems])]).astype(np.int32)
    pairs = np.ascontiguousarray(np.concatenate([p for p, _ in systems]), dtype=np.int32)
    vals = np.ascontiguousarray(np.concatenate([v for _, v in systems]), dtype=np.float64)
    order = np.concatenate([np.argsort(p[:, 0]) for p, _ in systems]).astype(np.int32)
    total = int(off[-1])
    ro = np.zeros(len(systems) + 1, np.int32)
    mo = np.zeros(len(systems) + 1, np.int32)
    rem = np.zeros(2 * total, np.int32)
    mg = np.zeros(4 * total, np.float64)
    vp = ctypes.c_void_p
    _native.call("dem_merge_plan", len(systems), off.ctypes.data_as(vp),
                 pairs.ctypes.data_as(vp), vals.ctypes.data_as(vp), order.ctypes.data_as(vp),
                 ro.ctypes.data_as(vp), rem.ctypes.data_as(vp), mo.ctypes.data_as(vp),
                 mg.ctypes.data_as(vp))
    ties = 0
    for k, (p, v) in enumerate(systems):
        ties += int(np.any(np.diff(p[:, 0]) == 0))
        r_ref, m_ref = merge_plan_rows(p, v)
        assert rem[ro[k]: ro[k + 1]].tolist() == list(r_ref), f"system {k} removed"
        got = mg[4 * mo[k]: 4 * mo[k + 1]].reshape(-1, 4)
        assert np.array_equal(got, np.asarray(m_ref, dtype=np.float64).reshape(-1, 4)), k
    assert ties > 50, "the systems should contain droplets in several pairs"


def test_native_plan_rejects_bad_input():
    vp = ctypes.c_void_p
    lib = _native.load()
    assert lib.dem_merge_plan(-1, None, None, None, None, None, None, None, None) == \
        _native.DEM_ERR_INVALID
    off = np.array([0, 1], np.int32)
    pairs = np.array([[0, 1]], np.int32)
    vals = np.ones((1, 8))
    order = np.array([3], np.int32)  # outside the system's segment
    out_i = np.zeros(4, np.int32)
    out_d = np.zeros(8)
    rc = lib.dem_merge_plan(1, off.ctypes.data_as(vp), pairs.ctypes.data_as(vp),
                            vals.ctypes.data_as(vp), order.ctypes.data_as(vp),
                            out_i.ctypes.data_as(vp), out_i.ctypes.data_as(vp),
                            out_i.ctypes.data_as(vp), out_d.ctypes.data_as(vp))
    assert rc == _native.DEM_ERR_INVALID and "order" in _native.last_error()
