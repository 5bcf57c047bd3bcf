"""CPU tests of the host-side drop-in API (no kernels): solver / solution /
state semantics mirroring the reference's tests, and objects that must be
constructible without a GPU."""

from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path

import numpy as np

from dem.collision_detector.spatial_hash_collision_detector import _next_power_of_2
from dem.octree import FlatOctree
from dem.particle_state import DropletState
from dem.solution import DropletSolution
from dem.solver import EulerDropletSolver


@dataclass
class StubForceCalculator:
    """Constant force (0,0,F), no convection (reference test_euler...py:22-37)."""

    F: float = 1e-12
    L: float = 1.0
    eta_oil: float = 0.065
    eta_water: float = 0.001
    boundary_mode: str = "open"

    def calculate_forces_and_convection(self, positions, radii):
        n = len(radii)
        forces = np.zeros((n, 3))
        forces[:, 2] = self.F
        return forces, np.zeros((n, 3))


class StubPostProcessor:
    def __init__(self):
        self.stop_simulation = False

    def update_solution(self, s):
        self.solution = s

    def update_live_plot(self):
        pass

    def live_plot(self):
        pass

    def finalize_live_plot(self):
        pass


class StubDetector:
    """Reports one contact (0, 1) on the first call, none afterwards."""

    def __init__(self):
        self.calls = 0

    def detect(self, positions, radii):
        self.calls += 1
        if self.calls == 1:
            return True, np.array([[0, 1]], dtype=np.int32)
        return False, np.empty((0, 2), dtype=np.int32)


def _solver(F=1e-12, n=2, det=None):
    state = DropletState(np.zeros((n, 3)), np.full(n, 50e-6), time=0.0)
    sol = DropletSolution(initial_droplet_state=state, length=20)
    return EulerDropletSolver(force_calculator=StubForceCalculator(F=F), solution=sol,
                              post_processor=StubPostProcessor(), collision_detector=det)


def test_drops_move_in_z():
    s = _solver()
    s.solve(dt=1e-3, total_time=1e-2)
    p = s.solution.get_current_positions()
    assert np.all(p[:, 2] > 0)
    np.testing.assert_allclose(p[:, :2], 0.0)
    # exact Euler update: 11 steps of F / (stokes * r) * dt
    v = 1e-12 / (s.stokes_factor * 50e-6)
    np.testing.assert_allclose(p[:, 2], 11 * v * 1e-3, rtol=1e-12)


def test_solver_stops_when_flag_set():
    s = _solver()
    s.post_processor.stop_simulation = True
    s.solve(dt=1e-3, total_time=1.0)
    assert s.solution.current_step == 0


def test_positions_wrap_in_periodic_mode():
    s = _solver(F=1e-9)
    s.force_calculator.boundary_mode = "periodic"
    s.force_calculator.L = 0.001
    s.solve(dt=1e-4, total_time=1e-3)
    p = s.solution.get_current_positions()
    assert np.all((p[:, 2] >= 0) & (p[:, 2] < 0.001))


def test_collision_restarts_step_without_advancing_time():
    s = _solver(n=3, det=StubDetector())
    s.solution.initial_droplet_state.positions[2] = [1e-3, 0, 0]
    s.solve(dt=1e-3, total_time=3e-3)
    first = s.solution.get_first_solution()
    assert first._next is s.solution
    assert first.get_times()[-1][0] == 0.0  # merged at t = 0, before any step
    assert s.solution.num_particles == 2
    assert np.isclose(s.solution.radii[-1] ** 3, 2 * (50e-6) ** 3)


def test_merge_and_chain_persistence(tmp_path: Path):
    pos = np.array([[0.0, 0.0, 0.0], [0.1, 0.0, 0.0], [0.5, 0.5, 0.5]])
    sol = DropletSolution(DropletState(pos, np.full(3, 0.05), time=0.0), length=5)
    sol.collided_droplets = np.array([[0, 1]])
    nxt = sol.generate_next_solution()
    assert nxt.num_particles == 2 and np.isclose(nxt.radii[-1] ** 3, 2 * 0.05**3)
    assert nxt._prev is sol and sol._next is nxt
    path = tmp_path / "chain.npz"
    nxt.save_step(0.5, np.ones((2, 3)))
    sol.save_chain_to_file(str(path))
    loaded = DropletSolution.load_chain_from_file(str(path))
    np.testing.assert_allclose(loaded.radii, sol.radii)
    np.testing.assert_allclose(loaded._next.trajectories, nxt.get_trajectories())


def test_state_npz_and_xlsx_round_trip(tmp_path: Path):
    rng = np.random.default_rng(0)
    st = DropletState(rng.random((7, 3)), rng.random(7) * 1e-5, time=2.5)
    st.save(str(tmp_path / "s.npz"))
    back = DropletState(filename=str(tmp_path / "s.npz"))
    np.testing.assert_array_equal(back.positions, st.positions)
    assert back.time == 2.5
    st.export_to_xlsx(str(tmp_path / "s.xlsx"))
    x = DropletState(filename=str(tmp_path / "s.xlsx"))
    np.testing.assert_array_equal(x.positions, st.positions)
    np.testing.assert_array_equal(x.radii, st.radii)
    assert x.time == 0


def test_next_power_of_2():
    assert [_next_power_of_2(k) for k in (1, 2, 3, 5, 16, 17)] == [1, 2, 4, 8, 16, 32]


def test_flat_octree_params_without_gpu():
    tree = FlatOctree(theta=0.5, mpl=1, num_particles=16)
    tree.update_params(theta=0.3, mpl=4)
    assert tree.theta == 0.3 and tree.mpl == 4
    assert tree.theta_sq[None] == 0.3 * 0.3 and tree.mpl_field[None] == 4
    assert tree.max_nodes == max(4 * 64 // 1, 1024)
    assert tree.nodes.count.to_numpy().shape == (tree.max_nodes,)


def test_taichi_stand_in_and_script_inputs(tmp_path):
    import importlib
    import sys

    from dem import run_script

    run_script.ensure_paths()
    ti = importlib.import_module("taichi")
    ti.init(arch=ti.cpu, cpu_max_num_threads=16, default_fp=ti.f64)
    assert ti.sync() is None
    # the box the scripts compute (main.py:35-40)
    r = np.array([2.5e-6, 7.5e-6])
    assert run_script.script_box(1000, 0.02) == np.cbrt(
        (np.pi * 1000 * np.sum(r) * np.sum(np.square(r))) / (3 * 0.02))
    sys.modules.pop("taichi", None)


def test_generator_produces_non_overlapping_droplets():
    from dem.particle_generator import UniformDropletGenerator

    g = UniformDropletGenerator(coord_range=(0, 2e-4), radii_range=(2.5e-6, 7.5e-6),
                                num_particles=300, minimum_distance=1e-6, seed=1)
    st = g.generate()
    d = np.linalg.norm(st.positions[:, None] - st.positions[None], axis=2)
    np.fill_diagonal(d, np.inf)
    assert np.all(d >= st.radii[:, None] + st.radii[None] + 1e-6 - 1e-18)


def test_merge_plan_rows_matches_the_reference_merge_bit_for_bit():
    """The batch's float merge plan equals merge_collided (the reference's
    numpy code) on random systems with chains of shared droplets."""
    from dem.solution.droplet_solution import merge_collided, merge_plan_rows

    rng = np.random.default_rng(7)
    for trial in range(200):
        n = int(rng.integers(2, 60))
        pos = rng.uniform(0, 1e-3, size=(n, 3))
        rad = np.exp(rng.uniform(np.log(2e-6), np.log(2e-5), size=n))
        k = int(rng.integers(1, 3 * n))
        i = rng.integers(0, n, size=k)
        j = rng.integers(0, n, size=k)
        pairs = np.stack([np.minimum(i, j), np.maximum(i, j)], 1)
        pairs = np.unique(pairs[pairs[:, 0] < pairs[:, 1]], axis=0).astype(np.int32)
        if len(pairs) == 0:
            continue
        ref_p, ref_r, _ = merge_collided(pos, rad, pairs)
        vals = np.concatenate([pos[pairs[:, 0]], rad[pairs[:, 0], None], pos[pairs[:, 1]],
                               rad[pairs[:, 1], None]], 1)
        removed, merged = merge_plan_rows(pairs, vals)
        keep = np.delete(np.arange(n), removed)
        m = np.asarray(merged, dtype=np.float64).reshape(-1, 4)
        got_p = np.vstack([pos[keep], m[:, :3]])
        got_r = np.concatenate([rad[keep], m[:, 3]])
        assert np.array_equal(got_p, ref_p) and np.array_equal(got_r, ref_r), trial


def test_get_state_rows_interpolation_and_chain():
    """get_state (reference droplet_solution.py:423-484) on compacted
    segments: a recorded time returns its row, a time between two rows
    interpolates linearly, and a time in a later segment of the chain is
    found there."""
    rng = np.random.default_rng(4)
    n = 5
    p0, p1, p2 = (rng.random((n, 3)) for _ in range(3))
    r = np.full(n, 1e-6)
    sol = DropletSolution(DropletState(p0.copy(), r.copy(), 0.0), length=2)
    sol.save_step(1.0, p1)
    sol.save_step(2.0, p2)  # doubles the buffers
    sol.compact()
    assert np.array_equal(sol.get_state(1.0).positions, p1)
    assert np.array_equal(sol.get_state(2.0).positions, p2)
    np.testing.assert_allclose(sol.get_state(1.25).positions, p1 + 0.25 * (p2 - p1), rtol=0,
                               atol=1e-15)
    nxt = sol.next_solution(p2[1:], r[1:], np.array([[0, 1]]))
    nxt.save_step(3.0, p0[1:])
    nxt.compact()
    got = sol.get_state(2.5).positions
    np.testing.assert_allclose(got, p2[1:] + 0.5 * (p0[1:] - p2[1:]), rtol=0, atol=1e-15)
