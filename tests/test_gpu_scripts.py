"""API replay of the reference's entry scripts on the GPU operators.

The scripts themselves (main.py, main_tree.py, main_multi.py) are not part
of this repository; these tests make the same constructor calls, method calls
and solver runs in the same order (cited line by line), with a short t_stop.
``python -m dem.run_script <script> --inputs`` runs the real files.
"""

from __future__ import annotations

import numpy as np
import pytest

from dem.collision_detector import SpatialHashCollisionDetector
from dem.force_calculator import DirectDropletForceCalculator
from dem.octree import compute_tree_stats, print_tree_stats, visualize_tree
from dem.octree.force_tree import TreeDropletForceCalculator
from dem.particle_state import DropletState
from dem.periodic_correction import COMSOLLatticeCorrection
from dem.post_processor import DropletPostProcessor
from dem.run_script import script_box, write_inputs
from dem.solution import DropletSolution
from dem.solver import EulerDropletSolver

pytestmark = pytest.mark.gpu

PHYS = dict(eps_oil=2.85, eta_oil=0.065, eta_water=0.001, E=3e5)


@pytest.fixture(scope="module")
def inputs(tmp_path_factory):
    d = tmp_path_factory.mktemp("scripts")
    write_inputs(d)
    return d


@pytest.fixture(scope="module")
def lattice():
    return COMSOLLatticeCorrection.load_default()


def test_main_direct_hash_comsol(gpu, inputs, lattice, tmp_path):
    """main.py:21-134 with t_stop = 0.2 (5 steps of dt = 0.04)."""
    box = script_box(1000, 0.02)
    state = DropletState(filename=str(inputs / "results" / "droplets_N1000_vol0.02_0.xlsx"))
    n = len(state.radii)
    fc = DirectDropletForceCalculator(num_particles=n, rho_water=1000, rho_oil=900, L=box,
                                      boundary_mode="periodic",
                                      correction_grid_resolution=lattice.grid_resolution, **PHYS)
    fc.load_periodic_correction(lattice, L_sim=box)
    det = SpatialHashCollisionDetector(num_particles=n, L=box, boundary_mode="periodic")
    sol = DropletSolution(initial_droplet_state=state, real_time_visualization=False)
    pp = DropletPostProcessor(sol, box_size=box)
    solver = EulerDropletSolver(force_calculator=fc, solution=sol, post_processor=pp,
                                collision_detector=det)
    solver.solve(0.04, 0.2)
    out = tmp_path / "results_N1000.npz"
    solver.solution.save_chain_to_file(str(out), precision="float32")
    chain = DropletSolution.load_chain_from_file(str(out))
    last = chain
    while last._next is not None:
        last = last._next
    assert np.isclose(last.get_times()[-1][0], 0.2, atol=1e-6)
    pos = solver.solution.get_current_positions()
    assert np.all(np.isfinite(pos)) and np.all((pos >= 0) & (pos < box))
    # volume is conserved through any merges
    np.testing.assert_allclose(np.sum(solver.solution.radii**3), np.sum(state.radii**3), rtol=1e-12)


def test_main_tree_diagnostics_and_run(gpu, inputs, lattice):
    """main_tree.py:98-138: tree operator, one diagnostic build + stats, solve."""
    box = script_box(1000, 0.02)
    state = DropletState(filename=str(inputs / "results" / "droplets_N1000_vol0.02_0.xlsx"))
    n = len(state.radii)
    fc = TreeDropletForceCalculator(num_particles=n, theta=0.35, mpl=1, L=box, periodic=True,
                                    correction_grid_resolution=lattice.grid_resolution, **PHYS)
    fc.load_periodic_correction(lattice, L_sim=box)
    det = SpatialHashCollisionDetector(num_particles=n, L=box, boundary_mode="periodic")
    sol = DropletSolution(initial_droplet_state=state, real_time_visualization=False)
    pp = DropletPostProcessor(sol, box_size=box)
    solver = EulerDropletSolver(force_calculator=fc, solution=sol, post_processor=pp,
                                collision_detector=det)
    p0, r0 = state.positions.copy(), state.radii.copy()
    fc.octree.build(p0, r0, box, True)
    _ = fc.octree.compute_forces(fc.m_const)
    stats = compute_tree_stats(fc.octree, build_time_ms=1.0, compute_time_ms=1.0)
    print_tree_stats(stats)
    assert stats.total_particles == n
    with pytest.warns(RuntimeWarning):
        visualize_tree(fc.octree, p0, r0)
    solver.solve(0.04, 0.12)
    pos = solver.solution.get_current_positions()
    assert np.all(np.isfinite(pos))


def test_main_multi_sequential_runs(gpu, inputs):
    """main_multi.py:43-71: default L = 1.0 'periodic' operator, no detector."""
    for dt, k in ((0.08, 1), (0.04, 2)):
        state = DropletState(filename=str(inputs / "results" / f"droplets_N10000_vol0.01_{k}.xlsx"))
        n = len(state.radii)
        fc = DirectDropletForceCalculator(num_particles=n, rho_water=1000, rho_oil=900, **PHYS)
        sol = DropletSolution(initial_droplet_state=state, real_time_visualization=False)
        pp = DropletPostProcessor(sol, box_size=script_box(10000, 0.01))
        solver = EulerDropletSolver(force_calculator=fc, solution=sol, post_processor=pp)
        solver.solve(dt, 2 * dt)
        assert solver.solution.current_step == 3
        assert np.all(np.isfinite(solver.solution.get_current_positions()))
