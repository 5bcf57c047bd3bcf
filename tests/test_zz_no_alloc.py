"""The allocation-free contact pass (``dem_hash_pairs_noalloc``) against the
workspace entry point (``dem_hash_pairs``) through the C ABI: bit-identical
pair lists and counts on two calls in a row, the second with a smaller
system (the shared scratch keeps the first call's contents), and the size
limit rejected before anything runs."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from dem import _device, _native
from dem.workloads import make_system

pytestmark = pytest.mark.gpu


def _next_pow2(n: int) -> int:
    return 1 << max(0, int(n - 1).bit_length())


def _both(pos, r, L, periodic):
    import torch

    n = len(r)
    T = _next_pow2(2 * n)
    pos_d, r_d = _device.to_device(pos), _device.to_device(r)
    cap = max(1024, 8 * n)
    out = {}
    for name in ("dem_hash_pairs", "dem_hash_pairs_noalloc"):
        pairs = torch.full((cap, 2), -7, dtype=torch.int32, device="cuda")
        count = torch.full((1,), -7, dtype=torch.int32, device="cuda")
        if name == "dem_hash_pairs":
            ws = torch.empty(_native.load().dem_hash_workspace_bytes(n, T), dtype=torch.uint8,
                             device="cuda")
            _native.call(name, _device.ptr(pos_d), _device.ptr(r_d), n, ctypes.c_double(L),
                         periodic, T, _device.ptr(pairs), cap, _device.ptr(count),
                         _device.ptr(ws), ws.numel(), _device.stream())
        else:
            _native.call(name, _device.ptr(pos_d), _device.ptr(r_d), n, ctypes.c_double(L),
                         periodic, T, _device.ptr(pairs), cap, _device.ptr(count),
                         _device.stream())
        k = int(count.item())
        assert 0 <= k <= cap
        out[name] = (k, pairs[:k].cpu().numpy())
    return out


def test_noalloc_equals_workspace_entry_point_twice(gpu):
    big = make_system(20000, seed=3, phi=0.03)
    small = make_system(3000, seed=4, phi=0.05)
    for s, periodic in ((big, 1), (small, 0)):
        print(f"\n[noalloc hash] n={s.n} periodic={periodic}", flush=True)
        got = _both(s.positions, s.radii * 1.6, s.L, periodic)
        (k0, p0), (k1, p1) = got["dem_hash_pairs"], got["dem_hash_pairs_noalloc"]
        assert k0 > 0, "the case should have contacts"
        assert k1 == k0 and np.array_equal(p1, p0)


def test_noalloc_rejects_sizes_beyond_the_scratch(gpu):
    n, T = 500000, 1 << 20
    assert _native.load().dem_hash_workspace_bytes(n, T) > (32 << 20)
    x = np.zeros(4)  # never read: the size check comes first
    rc = _native.load().dem_hash_pairs_noalloc(x.ctypes.data, x.ctypes.data, n, 1.0, 1, T,
                                               x.ctypes.data, 1, x.ctypes.data, None)
    assert rc == _native.DEM_ERR_CAPACITY
    assert "scratch" in _native.last_error()
