"""The device-resident Euler loop reproduces the host loop bit for bit
(same operators, same merge, bit-exact Euler kernel), merges included."""

from __future__ import annotations

import numpy as np
import pytest

from dem.collision_detector import SpatialHashCollisionDetector
from dem.force_calculator import DirectDropletForceCalculator
from dem.octree import TreeDropletForceCalculator
from dem.particle_state import DropletState
from dem.post_processor import DropletPostProcessor
from dem.solution import DropletSolution
from dem.solver import EulerDropletSolver
from dem.solver.device_solver import DeviceEulerDropletSolver
from dem.workloads import make_system

pytestmark = pytest.mark.gpu


def _run(solver_cls, s, tree=False, steps=12, dt=None):
    if tree:
        fc = TreeDropletForceCalculator(num_particles=s.n, theta=0.5, mpl=4, eps_oil=s.eps_oil,
                                        eta_oil=s.eta_oil, eta_water=s.eta_water, E=s.E, L=s.L,
                                        periodic=True)
    else:
        fc = DirectDropletForceCalculator(num_particles=s.n, eps_oil=s.eps_oil, eta_oil=s.eta_oil,
                                          eta_water=s.eta_water, E=s.E, L=s.L,
                                          boundary_mode="periodic")
    det = SpatialHashCollisionDetector(num_particles=s.n, L=s.L, boundary_mode="periodic")
    sol = DropletSolution(DropletState(s.positions.copy(), s.radii.copy(), 0.0), length=8)
    pp = DropletPostProcessor(sol, box_size=s.L)
    kw = {"verbose": False} if solver_cls is DeviceEulerDropletSolver else {}
    solver = solver_cls(force_calculator=fc, solution=sol, post_processor=pp,
                        collision_detector=det, **kw)
    solver.solve(dt, dt * steps)
    return solver


def _chain(sol):
    out = []
    cur = sol.get_first_solution()
    while cur is not None:
        out.append((cur.radii.copy(), cur.get_trajectories().copy(), cur.get_times().copy()))
        cur = cur._next
    return out


@pytest.mark.parametrize("tree", [False, True])
def test_device_loop_is_bit_identical_to_host_loop(gpu, tree):
    s = make_system(1500, seed=4, phi=0.04)
    # a velocity scale from the first evaluation: dt moves droplets ~0.5 r_min
    fc = DirectDropletForceCalculator(num_particles=s.n, eps_oil=s.eps_oil, eta_oil=s.eta_oil,
                                      eta_water=s.eta_water, E=s.E, L=s.L)
    F, V = fc.calculate_forces_and_convection(s.positions, s.radii)
    vmax = np.max(np.linalg.norm(F / (s.stokes_factor * s.radii[:, None]) + V, axis=1))
    dt = 0.5 * s.radii.min() / vmax
    a = _run(EulerDropletSolver, s, tree, dt=dt)
    b = _run(DeviceEulerDropletSolver, s, tree, dt=dt)
    ca, cb = _chain(a.solution), _chain(b.solution)
    assert len(ca) == len(cb) and len(ca) >= 2, "the run should contain merges"
    for (ra, ta, ha), (rb, tb, hb) in zip(ca, cb):
        assert np.array_equal(ra, rb)
        assert np.array_equal(ta, tb)
        assert np.array_equal(ha, hb)
