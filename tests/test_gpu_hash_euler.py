"""GPU parity of the spatial-hash contact pass (bit-exact pair sets) and of
the float64 Euler step (bit-exact with numpy)."""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from dem import _device, _native
from dem.collision_detector import SpatialHashCollisionDetector
from dem.collision_detector.spatial_hash_collision_detector import _next_power_of_2
from dem.workloads import make_system

pytestmark = pytest.mark.gpu


# --- the reference's own tests on the GPU operator (test_spatial_hash...py:24-113)


@pytest.mark.parametrize("n,expected", [(1, 1), (2, 2), (3, 4), (5, 8), (16, 16), (17, 32)])
def test_pow2_ceiling(n, expected):
    assert _next_power_of_2(n) == expected


def test_no_overlap_no_collisions(gpu):
    det = SpatialHashCollisionDetector(num_particles=8, L=1.0, boundary_mode="open")
    pos = np.array([[0.10, 0.10, 0.10], [0.90, 0.90, 0.90]])
    is_coll, pairs = det.detect(pos, np.full(2, 0.05), L=1.0, boundary_mode="open")
    assert is_coll is False and pairs.shape == (0, 2)


def test_overlap_returns_pair(gpu):
    det = SpatialHashCollisionDetector(num_particles=8, L=1.0, boundary_mode="open")
    pos = np.array([[0.50, 0.50, 0.50], [0.52, 0.50, 0.50]])
    is_coll, pairs = det.detect(pos, np.full(2, 0.05), L=1.0, boundary_mode="open")
    assert is_coll is True and pairs.tolist() == [[0, 1]]
    assert pairs.dtype == np.int32


def test_periodic_minimum_image_records_contact(gpu):
    det = SpatialHashCollisionDetector(num_particles=4, L=1.0, boundary_mode="periodic")
    pos = np.array([[0.01, 0.50, 0.50], [0.99, 0.50, 0.50]])
    r = np.full(2, 0.05)
    assert det.detect(pos, r, L=1.0, boundary_mode="open")[0] is False
    is_yes, pairs = det.detect(pos, r, L=1.0, boundary_mode="periodic")
    assert is_yes is True and len(pairs) >= 1


def test_zero_radii_returns_no_collision(gpu):
    det = SpatialHashCollisionDetector(num_particles=4, L=1.0, boundary_mode="open")
    is_coll, pairs = det.detect(np.zeros((2, 3)), np.zeros(2))
    assert is_coll is False and pairs.shape == (0, 2)


def test_pairs_unique_after_dedup(gpu):
    det = SpatialHashCollisionDetector(num_particles=8, L=1.0, boundary_mode="open")
    pos = np.array([[0.50, 0.50, 0.50], [0.51, 0.50, 0.50], [0.52, 0.50, 0.50]])
    is_coll, pairs = det.detect(pos, np.full(3, 0.02))
    assert is_coll is True and pairs.tolist() == [[0, 1], [0, 2], [1, 2]]


# --- bit-exact pair sets against the oracle


@pytest.mark.parametrize("periodic", [True, False])
@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_dense_random_sets_bit_exact(gpu, periodic, seed):
    rng = np.random.default_rng(seed)
    n, L = 3000, 1e-3
    pos = rng.uniform(0, L, size=(n, 3))
    r = rng.uniform(5e-6, 25e-6, size=n)
    cap = n + seed * 1000  # different table sizes
    det = SpatialHashCollisionDetector(num_particles=cap, L=L,
                                       boundary_mode="periodic" if periodic else "open")
    is_coll, pairs = det.detect(pos, r)
    ok, ref = oracle.hash_pairs(pos, r, L, periodic, _next_power_of_2(2 * cap))
    assert is_coll is ok
    assert pairs.shape == ref.shape and np.array_equal(pairs, ref)
    assert len(ref) > 50


def test_positions_outside_the_box_bit_exact(gpu):
    """Periodic pass with 30 % of the droplets shifted by +-L along one axis
    (unwrapped positions, as a caller may pass them).  L / cell = 16 exactly,
    so a shifted droplet keeps its cell and bucket and the reference's single
    image correction (|d| > L/2: d -= sign(d) L) brings it into contact with
    its neighbours; the search's shortcut for groups away from the seam
    (no minimum image) is valid only for positions inside the box and must
    stand aside here."""
    rng = np.random.default_rng(11)
    n, L = 3000, 1.0
    pos = rng.uniform(0, L, size=(n, 3))
    r = rng.uniform(0.005, 0.03125, size=n)
    r[0] = 0.03125  # cell = 2 r_max = 1/16
    sel = rng.random(n) < 0.3
    axis = rng.integers(0, 3, size=n)
    sign = rng.choice([-1.0, 1.0], size=n)
    pos[sel, axis[sel]] += sign[sel] * L
    det = SpatialHashCollisionDetector(num_particles=n, L=L, boundary_mode="periodic")
    is_coll, pairs = det.detect(pos, r)
    ok, ref = oracle.hash_pairs(pos, r, L, True, _next_power_of_2(2 * n))
    assert is_coll is ok and np.array_equal(pairs, ref) and len(ref) > 100
    # the shifted droplets take part in contacts
    assert np.any(sel[ref[:, 0]] | sel[ref[:, 1]])


@pytest.mark.parametrize("periodic", [True, False])
@pytest.mark.parametrize("capacity", [16, 200])
def test_aliased_buckets_bit_exact(gpu, periodic, capacity):
    """A table far smaller than the cell grid (32 or 512 buckets for ~8000
    cells): every bucket holds droplets of many cells, whose neighbourhoods
    differ, and neighbour cells alias onto shared buckets."""
    rng = np.random.default_rng(40 + capacity)
    n, L = 3000, 1e-3
    pos = rng.uniform(0, L, size=(n, 3))
    r = rng.uniform(5e-6, 25e-6, size=n)
    det = SpatialHashCollisionDetector(num_particles=capacity, L=L,
                                       boundary_mode="periodic" if periodic else "open")
    is_coll, pairs = det.detect(pos, r)
    ok, ref = oracle.hash_pairs(pos, r, L, periodic, _next_power_of_2(2 * capacity))
    assert is_coll is ok and np.array_equal(pairs, ref) and len(ref) > 50


def test_crowded_droplets_beyond_the_stash_bit_exact(gpu):
    """Droplets with many partners (more than the search keeps per droplet)
    are rescanned by the emit step: same sorted segments."""
    rng = np.random.default_rng(8)
    n, L = 800, 1.2e-4
    pos = rng.uniform(0, L, size=(n, 3))
    r = rng.uniform(6e-6, 14e-6, size=n)
    det = SpatialHashCollisionDetector(num_particles=n, L=L, boundary_mode="periodic")
    _, pairs = det.detect(pos, r)
    _, ref = oracle.hash_pairs(pos, r, L, True, _next_power_of_2(2 * n))
    per_target = np.bincount(ref[:, 0], minlength=n)
    assert per_target.max() > 16 and (per_target > 16).sum() > 10
    assert np.array_equal(pairs, ref)


def test_periodic_miss_reproduced(gpu):
    r = np.full(2, 1 / 9)
    pos = np.array([[0.01, 0.5, 0.5], [0.87, 0.5, 0.5]])
    det = SpatialHashCollisionDetector(num_particles=1024, L=1.0, boundary_mode="periodic")
    assert det.detect(pos, r)[0] is False  # the reference misses this contact too
    assert oracle.brute_pairs(pos, r, 1.0, True).tolist() == [[0, 1]]


def test_n65536_scale_bit_exact(gpu):
    """N = 65536 at phi = 0.02 with radii inflated 1.6x to create contacts."""
    s = make_system(65536, seed=0, phi=0.02)
    r = s.radii * 1.6
    det = SpatialHashCollisionDetector(num_particles=s.n, L=s.L, boundary_mode="periodic")
    is_coll, pairs = det.detect(s.positions, r)
    ok, ref = oracle.hash_pairs(s.positions, r, s.L, True, _next_power_of_2(2 * s.n))
    assert is_coll is ok and np.array_equal(pairs, ref)
    assert len(ref) > 100
    # every pair is a true contact and the set is closed under the brute force
    # restricted to the rows it reports
    g = max(1, int(np.ceil(s.L / (2 * r.max()))))
    print(f"\n[N=65536] g={g} L/cell={s.L / (2 * r.max()):.3f} pairs={len(ref)}")


def _hash_through_abi(pos, r, L, periodic, table_size):
    """dem_hash_pairs with an explicit table size, pairs as a host array."""
    import ctypes

    t = _device.torch()
    n = pos.shape[0]
    pd, rd = _device.to_device(pos), _device.to_device(r)
    lib = _native.load()
    ws = t.empty(lib.dem_hash_workspace_bytes(n, table_size), dtype=t.uint8, device="cuda")
    cap = 4 * n
    pairs = t.empty((cap, 2), dtype=t.int32, device="cuda")
    cnt = t.zeros(1, dtype=t.int32, device="cuda")
    _native.call("dem_hash_pairs", _device.ptr(pd), _device.ptr(rd), n, ctypes.c_double(L),
                 int(periodic), table_size, _device.ptr(pairs), cap, _device.ptr(cnt),
                 _device.ptr(ws), ws.numel(), _device.stream())
    k = int(cnt.item())
    assert k <= cap
    return pairs[:k].cpu().numpy()


@pytest.mark.parametrize("n,log2_table,periodic", [
    (300_000, 23, True),     # bucket-count scan in 1024-entry tiles
    (300_000, 25, False),    # 4096-entry tiles (the largest)
    (2_200_000, 23, True),   # pair-count scan in 512-entry tiles
])
def test_scan_tile_sizes_bit_exact(gpu, n, log2_table, periodic):
    """The counts are scanned in tiles sized to the input (256 entries up to
    2^21 entries, then 512, 1024, ... 4096): large tables and n > 2^21 take
    the wider tiles.  Uniform droplets with radii for ~0.03-0.04 contacts
    each; the oracle uses the same table size."""
    rng = np.random.default_rng(n + log2_table)
    L = 1e-3
    pos = rng.uniform(0.0, L, size=(n, 3))
    rmax = 2e-6 if n < 1_000_000 else 1e-6
    r = rng.uniform(0.5, 1.0, size=n) * rmax
    T = 1 << log2_table
    got = _hash_through_abi(pos, r, L, periodic, T)
    ok, ref = oracle.hash_pairs(pos, r, L, periodic, T)
    assert ok and len(ref) > 1000
    assert np.array_equal(got, ref)
    print(f"\n[n={n} T=2^{log2_table}] pairs={len(ref)}")


def test_unsynced_capacity_growth(gpu):
    """More pairs than the initial buffer: the operator grows it, nothing is lost."""
    rng = np.random.default_rng(9)
    n, L = 1500, 2e-4
    pos = rng.uniform(0, L, size=(n, 3))
    r = np.full(n, 12e-6)
    det = SpatialHashCollisionDetector(num_particles=n, L=L, boundary_mode="periodic")
    _, pairs = det.detect(pos, r)
    _, ref = oracle.hash_pairs(pos, r, L, True, _next_power_of_2(2 * n))
    assert len(ref) > 1024 and np.array_equal(pairs, ref)


# --- Euler step


@pytest.mark.parametrize("periodic", [True, False])
def test_euler_step_bit_exact(gpu, periodic):
    rng = np.random.default_rng(4)
    n, L = 5000, 7e-4
    pos = rng.uniform(-0.1 * L, 1.1 * L, size=(n, 3))
    F = rng.normal(size=(n, 3)) * 1e-11
    conv = rng.normal(size=(n, 3)) * 1e-5
    r = rng.uniform(2e-6, 20e-6, size=n)
    sf = oracle.stokes_factor(0.05, 0.001)
    dt = 0.013
    ref = oracle.euler_step(pos, F, conv, r, sf, dt, L, periodic)
    # keep every device tensor referenced until the kernel has run
    pd, Fd, cd, rd = (_device.to_device(a) for a in (pos, F, conv, r))
    _native.call("dem_euler_step", _device.ptr(pd), _device.ptr(Fd), _device.ptr(cd),
                 _device.ptr(rd), n, sf, dt, L, int(periodic), _device.stream())
    assert np.array_equal(pd.cpu().numpy(), ref)
    v = rng.normal(size=(n, 3)) * 1e-4
    ref2 = oracle.euler_step_total(pos, v, dt, L, periodic)
    pd, vd = _device.to_device(pos), _device.to_device(v)
    _native.call("dem_euler_step_total", _device.ptr(pd), _device.ptr(vd), n, dt, L,
                 int(periodic), _device.stream())
    assert np.array_equal(pd.cpu().numpy(), ref2)
