"""CPU checks of the C ABI: libdemhip.so loads and exports every function
declared in include/dem_hip.h; host-only entry points behave (no kernels are
launched here)."""

from __future__ import annotations

import ctypes

import pytest

from dem import _native


def test_library_loads_and_exports_header():
    lib = _native.load()
    names = _native.header_functions()
    assert "dem_direct_forces_and_convection" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []
    # every exported entry point also has a ctypes signature
    assert set(names) <= set(_native.SIGNATURES)


def test_version_and_error_string():
    lib = _native.load()
    assert lib.dem_version() >= 1
    assert isinstance(_native.last_error(), str)


def test_workspace_sizes_are_host_only_and_monotone():
    lib = _native.load()
    a = lib.dem_direct_workspace_bytes(1000)
    b = lib.dem_direct_workspace_bytes(100000)
    assert 0 < a < b and a % 256 == 0
    assert lib.dem_hash_workspace_bytes(1000, 2048) < lib.dem_hash_workspace_bytes(1000, 1 << 20)


def test_invalid_arguments_are_rejected_before_any_launch():
    lib = _native.load()
    ph = _native.DemPhys(1.0, 1.0, -1.0, 1, 0)  # periodic with L < 0
    rc = lib.dem_direct_forces(ctypes.c_void_p(8), ctypes.c_void_p(8), 10, ctypes.byref(ph),
                               ctypes.c_void_p(8), ctypes.c_void_p(8), 1 << 30, None)
    assert rc == _native.DEM_ERR_INVALID
    assert "L > 0" in _native.last_error()
    rc = lib.dem_hash_pairs(None, None, 4, ctypes.c_double(1.0), 1, 3, None, 0,
                            ctypes.c_void_p(8), None, 0, None)
    assert rc == _native.DEM_ERR_INVALID  # table size not a power of two
    ph = _native.DemPhys(1.0, 1.0, 1.0, 1, 0)
    rc = lib.dem_direct_forces(ctypes.c_void_p(8), ctypes.c_void_p(8), 10, ctypes.byref(ph),
                               ctypes.c_void_p(8), ctypes.c_void_p(8), 16, None)
    assert rc == _native.DEM_ERR_WORKSPACE
    with pytest.raises(_native.NativeError):
        _native.call("dem_euler_step", None, None, None, None, 5, 1.0, 1.0, 1.0, 0, None)
    # merge compaction: more removed than droplets, negative counts, null outputs
    for args in ((4, 5, 0), (-1, 0, 0), (4, 0, -1)):
        rc = lib.dem_merge_compact(ctypes.c_void_p(8), ctypes.c_void_p(8), args[0],
                                   ctypes.c_void_p(8), args[1], ctypes.c_void_p(8), args[2],
                                   ctypes.c_void_p(8), ctypes.c_void_p(8), None)
        assert rc == _native.DEM_ERR_INVALID and "bad counts" in _native.last_error()
    rc = lib.dem_merge_compact(ctypes.c_void_p(8), ctypes.c_void_p(8), 4, ctypes.c_void_p(8), 2,
                               ctypes.c_void_p(8), 1, None, None, None)
    assert rc == _native.DEM_ERR_INVALID and "null" in _native.last_error()
    rc = lib.dem_gather_droplets(ctypes.c_void_p(8), ctypes.c_void_p(8), 4, None, 3,
                                 ctypes.c_void_p(8), None)
    assert rc == _native.DEM_ERR_INVALID and "null" in _native.last_error()
    rc = lib.dem_gather_droplets(None, None, -1, None, 0, None, None)
    assert rc == _native.DEM_ERR_INVALID and "bad counts" in _native.last_error()


def test_zero_particles_is_a_no_op():
    lib = _native.load()
    ph = _native.DemPhys(1.0, 1.0, 1.0, 1, 0)
    assert lib.dem_direct_forces(None, None, 0, ctypes.byref(ph), None, None, 0, None) == 0
    assert lib.dem_euler_step(None, None, None, None, 0, 1.0, 1.0, 1.0, 1, None) == 0
    assert lib.dem_merge_compact(None, None, 0, None, 0, None, 0, None, None, None) == 0


def _split_max(n):
    return max(1, min(64, (n + 63) // 64, (1 << 21) // max(n, 1)))


def test_batch_split_rule_is_host_only():
    lib = _native.load()
    for n in (1, 63, 300, 2048, 20000, 65536):
        for jlen in (0, 256, 1024, 4096):
            want = 1 if jlen <= 0 else min(-(-n // jlen), _split_max(n))
            assert lib.dem_batch_split(n, jlen) == want
    # the split times n never grows when a system shrinks (capacity bounds)
    prev = 0
    for n in range(1, 5000, 37):
        cur = lib.dem_batch_split(n, 256) * n
        assert cur >= prev
        prev = cur


def _ensembles(specs):
    arr = (_native.DemBatchEnsemble * len(specs))()
    for e, (off, n, cap) in enumerate(specs):
        arr[e] = _native.DemBatchEnsemble(off, n, cap, 10, 1e-3, 1e-3, 1.0, 1.0, 1.0)
    return arr


def test_batch_workspace_and_argument_checks():
    lib = _native.load()
    tune = _native.DemBatchTuning(256, 1024, 0, 0, 1, 0)
    ok = _ensembles([(0, 500, 500), (500, 800, 1000)])
    small = lib.dem_batch_workspace_bytes(ok, 2, ctypes.byref(tune))
    big = lib.dem_batch_workspace_bytes(_ensembles([(0, 500, 500), (500, 800, 20000)]), 2,
                                        ctypes.byref(tune))
    assert 0 < small < big and small % 256 == 0
    out = ctypes.c_void_p()
    # overlapping reserved ranges
    bad = _ensembles([(0, 500, 500), (400, 100, 100)])
    rc = lib.dem_batch_create(bad, 2, ctypes.byref(tune), ctypes.c_void_p(8), ctypes.c_void_p(8),
                              ctypes.c_void_p(8), ctypes.c_void_p(8), ctypes.c_void_p(8), 1 << 40,
                              None, ctypes.byref(out))
    assert rc == _native.DEM_ERR_INVALID and "overlap" in _native.last_error()
    # n above capacity
    bad = _ensembles([(0, 600, 500)])
    assert lib.dem_batch_workspace_bytes(bad, 1, ctypes.byref(tune)) == 0
    # unsupported block size
    t2 = _native.DemBatchTuning(96, 1024, 0, 0, 1, 0)
    rc = lib.dem_batch_create(ok, 2, ctypes.byref(t2), ctypes.c_void_p(8), ctypes.c_void_p(8),
                              ctypes.c_void_p(8), ctypes.c_void_p(8), ctypes.c_void_p(8), 1 << 40,
                              None, ctypes.byref(out))
    assert rc == _native.DEM_ERR_INVALID and "block" in _native.last_error()
    # workspace too small
    rc = lib.dem_batch_create(ok, 2, ctypes.byref(tune), ctypes.c_void_p(8), ctypes.c_void_p(8),
                              ctypes.c_void_p(8), ctypes.c_void_p(8), ctypes.c_void_p(8), 16,
                              None, ctypes.byref(out))
    assert rc == _native.DEM_ERR_WORKSPACE
    assert out.value is None
    assert lib.dem_batch_destroy(None) == 0
    # the incremental re-check switch on a null handle
    rc = lib.dem_batch_set_incremental(None, 1)
    assert rc == _native.DEM_ERR_INVALID and "null handle" in _native.last_error()
