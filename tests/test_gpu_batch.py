"""Batched ensembles (``dem_batch_*``) on the GPU.

* Bit-for-bit: uneven periodic systems stepped together (contacts, merge
  rounds, systems finishing at different step counts) end in exactly the
  states the single-system path reaches system after system with the same
  source split; also without the contact pass.
* Pair lists of the batched contact pass equal the CPU oracle's restatement
  of the reference hash (bucket aliases and periodic misses included), and a
  system whose pairs overflow its slot still merges exactly.
* Against the FP64 oracle: the first batched step's forces, velocities and
  displacements, under the direct kernels' accuracy contract.
* Graph replay equals eager launches; merge rounds overlapped with the
  advance on a second stream equal serial rounds; finished systems stay put;
  resizing beyond the capacity is rejected.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

import oracle
from dem import _native
from dem.batch import BatchedEnsembles, SequentialEnsembles, Tuning, autotune
from dem.force_calculator import DirectDropletForceCalculator
from dem.workloads import make_system

from ._errors import summary

pytestmark = pytest.mark.gpu

TOL_F = 1e-5
TOL_V_NORM = 1e-5
TOL_V_P999 = 1e-4
TOL_V_MAX = 1e-3
FLOOR = 1e-3

# (n, phi, E): uneven sizes, dense enough that merges happen within a few steps
SPECS = [(300, 0.06, 4e5), (777, 0.05, 3e5), (1500, 0.04, 5e5), (1024, 0.06, 2e5),
         (513, 0.08, 4.5e5), (2100, 0.03, 3.5e5)]
STEPS = [15, 10, 15, 12, 15, 8]


def _systems(specs=SPECS):
    return [make_system(n, seed=10 + e, phi=phi, E=E) for e, (n, phi, E) in enumerate(specs)]


def _dts(systems, frac=0.5):
    """dt moving the fastest droplet frac * r_min in the first step."""
    out = []
    for s in systems:
        fc = DirectDropletForceCalculator(num_particles=s.n, eps_oil=s.eps_oil, eta_oil=s.eta_oil,
                                          eta_water=s.eta_water, E=s.E, L=s.L)
        F, V = fc.calculate_forces_and_convection(s.positions, s.radii)
        vmax = np.max(np.linalg.norm(F / (s.stokes_factor * s.radii[:, None]) + V, axis=1))
        out.append(frac * s.radii.min() / vmax)
    return out


@pytest.fixture(scope="module")
def setup():
    systems = _systems()
    return systems, _dts(systems)


@pytest.mark.parametrize("block,jlen", [(64, 256), (256, 512), (128, 100000)])
def test_batched_equals_sequential_bit_for_bit(gpu, setup, block, jlen):
    systems, dts = setup
    b = BatchedEnsembles(systems, dts, STEPS, Tuning(block, jlen), capture=True)
    g = b.run_until_done(max_global_steps=200)
    s = SequentialEnsembles(systems, dts, STEPS, jlen=jlen)
    s.run_until_done()
    assert b.done()
    assert np.array_equal(b.steps_done.cpu().numpy(), np.array(STEPS))
    assert b.merges == s.merges and b.merges > 0, "the run should contain merges"
    assert g == max(STEPS), "every active system advances once per global step"
    for e in range(len(systems)):
        assert np.array_equal(b.radii_of(e), s.radii_of(e)), f"system {e} radii"
        assert np.array_equal(b.positions(e), s.positions(e)), f"system {e} positions"
        n = len(b.radii_of(e))
        np.testing.assert_array_equal(b.forces(e), s.F[e].cpu().numpy()[:n])
    # volume is conserved through the merges
    for e, sy in enumerate(systems):
        np.testing.assert_allclose(np.sum(b.radii_of(e) ** 3), np.sum(sy.radii ** 3), rtol=1e-12)
    b.close()


@pytest.mark.parametrize("jlen", [256, 1024, 100000])
def test_pair_once_batched_equals_sequential_bit_for_bit(gpu, setup, jlen):
    """Pair-once sums (blocks of 256, 1 / 4 / all partner blocks per work
    item) against the single-system path in pair-once mode 2 with the same
    chunks, through contacts, merges and uneven step budgets."""
    systems, dts = setup
    b = BatchedEnsembles(systems, dts, STEPS, Tuning(256, jlen, pair_once=True), capture=True)
    b.run_until_done(max_global_steps=200)
    s = SequentialEnsembles(systems, dts, STEPS, jlen=jlen, pair_once=True)
    s.run_until_done()
    assert b.done() and b.merges == s.merges and b.merges > 0
    for e in range(len(systems)):
        assert np.array_equal(b.radii_of(e), s.radii_of(e)), f"system {e} radii"
        assert np.array_equal(b.positions(e), s.positions(e)), f"system {e} positions"
        n = len(b.radii_of(e))
        np.testing.assert_array_equal(b.forces(e), s.F[e].cpu().numpy()[:n])
    b.close()


@pytest.mark.parametrize("pair_once,merge_first", [(False, False), (True, False), (False, True),
                                                  (True, True)])
def test_overlapped_merge_rounds_equal_serial_rounds(gpu, setup, pair_once, merge_first):
    """Merge rounds on a second stream beside the advance of the contact-free
    systems (dem_batch_wait_go, dem_batch_merge_async, tables rebuilt before
    the second advance), or all merge rounds first and then one advance of
    every system, end in exactly the states of the serial rounds."""
    systems, dts = setup
    tune = Tuning(256, 1024, pair_once=pair_once)
    a = BatchedEnsembles(systems, dts, STEPS, tune, capture=True)
    a.merge_first = merge_first
    c = BatchedEnsembles(systems, dts, STEPS, tune, capture=True)
    c.overlap = False
    a.run_until_done(200)
    c.run_until_done(200)
    assert a.merges == c.merges and a.merges > 0 and a.merge_rounds == c.merge_rounds
    for e in range(len(systems)):
        assert np.array_equal(a.radii_of(e), c.radii_of(e)), f"system {e} radii"
        assert np.array_equal(a.positions(e), c.positions(e)), f"system {e} positions"
    assert np.array_equal(a.time.cpu().numpy(), c.time.cpu().numpy())
    a.close()
    c.close()


def test_wait_go_needs_an_advance(gpu, setup):
    systems, dts = setup
    b = BatchedEnsembles(systems[:2], dts[:2], [2, 2], Tuning(64, 256), capture=False)
    rc = _native.load().dem_batch_wait_go(b._h, None)
    assert rc == _native.DEM_ERR_INVALID and "advance" in _native.last_error()
    b.close()


@pytest.mark.parametrize("pair_once", [False, True])
def test_first_step_against_fp64_oracle(gpu, setup, pair_once):
    systems, dts = setup
    sub = [systems[0], systems[2], systems[5]]
    sdt = [dts[0], dts[2], dts[5]]
    b = BatchedEnsembles(sub, sdt, [1, 1, 1], Tuning(256, 512, pair_once=pair_once),
                         capture=False)
    b.run(1, merge=False)
    assert np.array_equal(b.steps_done.cpu().numpy(), np.array([1, 1, 1]))
    for e, sy in enumerate(sub):
        F_ref = oracle.direct_forces(sy.positions, sy.radii, sy.m_const, sy.L, True)
        V_ref = oracle.direct_convection(sy.positions, F_ref, sy.eta_const, sy.L, True)
        ef = summary(b.forces(e), F_ref, FLOOR)
        ev = summary(b.velocities(e), V_ref, FLOOR)
        assert ef["max"] <= TOL_F, ef
        assert ev["norm"] <= TOL_V_NORM and ev["p999"] <= TOL_V_P999 and ev["max"] <= TOL_V_MAX, ev
        x_ref = oracle.euler_step(sy.positions, F_ref, V_ref, sy.radii, sy.stokes_factor, sdt[e],
                                  sy.L, True)
        d = b.positions(e) - sy.positions
        d_ref = x_ref - sy.positions
        d -= sy.L * np.round(d / sy.L)
        d_ref -= sy.L * np.round(d_ref / sy.L)
        ed = summary(d, d_ref, FLOOR)
        assert ed["norm"] <= TOL_V_NORM and ed["max"] <= TOL_V_MAX, ed
    b.close()


@pytest.mark.parametrize("jlen", [1024, 8192])
def test_pair_once_first_step_at_ensemble_sizes_against_fp64_oracle(gpu, jlen):
    """Systems of the 256-ensemble workload's sizes (10k-20k droplets, 40-79
    blocks of 256, own seed, E and volume fraction) through the pair-once
    batched step: forces, velocities and displacements of the first step
    against the FP64 oracle under the direct kernels' accuracy contract."""
    specs = [(10240, 0.012, 1.3e5), (15001, 0.035, 4.1e5), (19999, 0.049, 2.6e5)]
    sub = [make_system(n, seed=300 + e, phi=phi, E=E) for e, (n, phi, E) in enumerate(specs)]
    sdt = _dts(sub, frac=0.1)
    b = BatchedEnsembles(sub, sdt, [1] * len(sub), Tuning(256, jlen, pair_once=True),
                         capture=True)
    b.run(1, merge=False)
    assert np.array_equal(b.steps_done.cpu().numpy(), np.ones(len(sub), dtype=np.int32))
    for e, sy in enumerate(sub):
        F_ref = oracle.direct_forces(sy.positions, sy.radii, sy.m_const, sy.L, True)
        V_ref = oracle.direct_convection(sy.positions, F_ref, sy.eta_const, sy.L, True)
        ef = summary(b.forces(e), F_ref, FLOOR)
        ev = summary(b.velocities(e), V_ref, FLOOR)
        print(f"\n[batched pair-once n={sy.n} jlen={jlen}] force {ef}\n  velocity {ev}")
        assert ef["max"] <= TOL_F, ef
        assert ev["norm"] <= TOL_V_NORM and ev["p999"] <= TOL_V_P999 and ev["max"] <= TOL_V_MAX, ev
        x_ref = oracle.euler_step(sy.positions, F_ref, V_ref, sy.radii, sy.stokes_factor, sdt[e],
                                  sy.L, True)
        d = b.positions(e) - sy.positions
        d_ref = x_ref - sy.positions
        d -= sy.L * np.round(d / sy.L)
        d_ref -= sy.L * np.round(d_ref / sy.L)
        ed = summary(d, d_ref, FLOOR)
        assert ed["norm"] <= TOL_V_NORM and ed["max"] <= TOL_V_MAX, ed
    b.close()


def test_graph_replay_equals_eager_launches(gpu, setup):
    systems, dts = setup
    a = BatchedEnsembles(systems[:3], dts[:3], [6, 6, 6], Tuning(128, 1024), capture=True)
    c = BatchedEnsembles(systems[:3], dts[:3], [6, 6, 6], Tuning(128, 1024), capture=False)
    a.run_until_done(100)
    c.run_until_done(100)
    assert a.captured and not c.captured
    for e in range(3):
        assert np.array_equal(a.positions(e), c.positions(e))
        assert np.array_equal(a.radii_of(e), c.radii_of(e))
    a.close()
    c.close()


def test_finished_and_empty_step_budgets(gpu, setup):
    systems, dts = setup
    b = BatchedEnsembles(systems[:2], dts[:2], [0, 3], Tuning(64, 256))
    p0 = b.positions(0).copy()
    b.run(5)
    assert np.array_equal(b.positions(0), p0)
    assert list(b.steps_done.cpu().numpy()) == [0, 3]
    np.testing.assert_allclose(b.time.cpu().numpy()[1], 3 * dts[1], rtol=1e-15)
    es = np.array([1], dtype=np.int32)
    ns = np.array([int(b.capacity[1]) + 1], dtype=np.int32)
    with pytest.raises(_native.NativeError):
        _native.call("dem_batch_resize", b._h, 1, es.ctypes.data_as(ctypes.c_void_p),
                     ns.ctypes.data_as(ctypes.c_void_p), None)
    b.close()


def _overlapping(n, seed, phi, L_scale=1.0):
    """Uniform positions without overlap rejection: many contacts."""
    from dem.workloads import System, box_edge, lognormal_radii

    rng = np.random.default_rng(seed)
    r = lognormal_radii(rng, n)
    L = box_edge(r, phi) * L_scale
    return System(positions=rng.uniform(0, L, size=(n, 3)), radii=r, L=L, phi=phi, seed=seed)


def test_pair_lists_equal_the_oracle_hash(gpu):
    systems = [_overlapping(n, 50 + k, phi) for k, (n, phi) in
               enumerate([(300, 0.1), (1200, 0.05), (64, 0.2), (2500, 0.03), (700, 0.08)])]
    b = BatchedEnsembles(systems, [1e-3] * len(systems), [1] * len(systems), Tuning(128, 512))
    b.check()
    cnt = b.pair_count.cpu().numpy()
    assert cnt.sum() <= b.pair_capacity
    base = b.pair_base.cpu().numpy()
    vals = b.pair_values.cpu().numpy()
    for e, sy in enumerate(systems):
        T = 1 << int(np.ceil(np.log2(2 * sy.n)))
        hit, ref = oracle.hash_pairs(sy.positions, sy.radii, sy.L, True, T)
        assert cnt[e] == len(ref) and hit == (len(ref) > 0)
        got = b.contact_pairs(e)
        assert np.array_equal(got, ref.astype(np.int32)), f"system {e}"
        # the packed values are the droplets' own coordinates and radii
        v = vals[base[e]: base[e] + cnt[e]]
        np.testing.assert_array_equal(v[:, 0:3], sy.positions[got[:, 0]])
        np.testing.assert_array_equal(v[:, 3], sy.radii[got[:, 0]])
        np.testing.assert_array_equal(v[:, 4:7], sy.positions[got[:, 1]])
        np.testing.assert_array_equal(v[:, 7], sy.radii[got[:, 1]])
    b.close()


def test_overflowing_pair_buffer_merges_exactly(gpu):
    # 80 droplets of 8 um in a 40 um box: ~850 pairs against a packed buffer
    # of 480 (sum of max(64, capacity))
    from dem.workloads import System

    rng = np.random.default_rng(3)
    r = np.full(80, 8e-6)
    L = 4e-5
    crowd = System(positions=rng.uniform(0, L, size=(80, 3)), radii=r, L=L)
    calm = _systems([(400, 0.02, 3e5)])[0]
    systems = [crowd, calm]
    dts = [1e-4, _dts([calm])[0]]
    b = BatchedEnsembles(systems, dts, [3, 3], Tuning(64, 256))
    b.check()
    assert b.pair_count.cpu().numpy().sum() > b.pair_capacity
    b.run_until_done(20)
    s = SequentialEnsembles(systems, dts, [3, 3], jlen=256)
    s.run_until_done()
    for e in range(2):
        assert np.array_equal(b.radii_of(e), s.radii_of(e))
        assert np.array_equal(b.positions(e), s.positions(e))
    b.close()


def test_without_contact_pass_equals_sequential_without_detector(gpu, setup):
    systems, dts = setup
    steps = [4, 2, 3]
    b = BatchedEnsembles(systems[:3], dts[:3], steps, Tuning(256, 1024, contacts=False))
    b.run_until_done(10)
    s = SequentialEnsembles(systems[:3], dts[:3], steps, jlen=1024, contacts=False)
    s.run_until_done()
    assert b.merges == 0 and s.merges == 0
    for e in range(3):
        assert np.array_equal(b.positions(e), s.positions(e))
    b.close()


def test_autotune_returns_a_timed_candidate(gpu, setup):
    systems, dts = setup
    best, times = autotune(systems[:3], dts[:3], blocks=(64, 256), jlens=(512, 4096),
                           pair_once_jlens=(1024,), reps=1)
    key = ("pair_once", best.jlen) if best.pair_once else (best.block, best.jlen)
    assert key in times and len(times) == 5 and ("pair_once", 1024) in times
    assert times[key] == min(times.values())
    assert all(t > 0 for t in times.values())


def _recheck_rounds(b, systems, check_oracle=True):
    """Merge rounds of one global step, comparing every check's pair lists
    with the oracle's restatement of the reference hash on the systems'
    current droplets.  Returns (rounds, incremental re-checks, full re-checks)."""
    caps = [int(c) for c in b.capacity]
    rounds = inc = full = 0
    b.check(new_step=True)
    while True:
        b._stage()
        b._staged.synchronize()
        cnt = b.pair_count.cpu().numpy()
        if rounds:
            act = b.active.cpu().numpy()
            flag = b.incremental_rechecks.cpu().numpy()
            inc += int(((act == 1) & (flag == 1)).sum())
            full += int(((act == 1) & (flag == 0)).sum())
        if check_oracle:
            for e, sy in enumerate(systems):
                T = 1 << int(np.ceil(np.log2(2 * caps[e])))
                _, ref = oracle.hash_pairs(b.positions(e), b.radii_of(e), sy.L, True, T)
                got = b.contact_pairs(e)
                assert np.array_equal(got, ref.astype(np.int32)), f"round {rounds}, system {e}"
        if b._merge_staged(defer_tables=True) == 0:
            return rounds, inc, full
        rounds += 1
        b.check(new_step=False)


def test_incremental_rechecks_equal_the_oracle_hash(gpu):
    """Re-checks after complete greedy merges search only around the merged
    droplets; every round's pair lists still equal the oracle hash of the
    current droplets.  Dense systems give chains of rounds; a system whose
    largest droplet merges (the cell grows) and one with g <= 5 take the full
    path or the wrapped neighbourhood."""
    from dem.workloads import System

    systems = [_overlapping(n, 70 + k, phi) for k, (n, phi) in
               enumerate([(900, 0.12), (2500, 0.06), (400, 0.2), (1500, 0.08)])]
    rng = np.random.default_rng(5)
    # few large droplets in a small box: g <= 5, the largest ones touch
    r = np.concatenate([np.full(6, 12e-6), np.full(60, 4e-6)])
    L = 1.1e-4
    systems.append(System(positions=rng.uniform(0, L, size=(len(r), 3)), radii=r, L=L))
    b = BatchedEnsembles(systems, [1e-3] * len(systems), [1] * len(systems), Tuning(128, 512))
    rounds, inc, full = _recheck_rounds(b, systems)
    assert rounds >= 3, rounds
    assert inc > 0 and full > 0, (inc, full)
    b.close()


def test_incremental_rechecks_equal_full_rechecks_bit_for_bit(gpu, setup):
    """The same systems stepped with incremental re-checks and with full ones
    end in identical states (and the dense setup does take incremental ones)."""
    systems, dts = setup
    dense = [_overlapping(n, 90 + k, phi) for k, (n, phi) in enumerate([(800, 0.1), (1600, 0.07)])]
    sys_all = list(systems) + dense
    dts_all = list(dts) + _dts(dense)
    steps = STEPS + [6, 6]
    runs = []
    for on in (1, 0):
        b = BatchedEnsembles(sys_all, dts_all, steps, Tuning(256, 512, pair_once=True))
        _native.call("dem_batch_set_incremental", b._h, on)
        b.run_until_done(60)
        runs.append(b)
    a, c = runs
    assert a.merges == c.merges and a.merges > 0 and a.merge_rounds == c.merge_rounds
    for e in range(len(sys_all)):
        assert np.array_equal(a.radii_of(e), c.radii_of(e)), e
        assert np.array_equal(a.positions(e), c.positions(e)), e
    for b in runs:
        b.close()
