"""The device-resident Euler loop reproduces the host loop bit for bit
(same operators, same merge, bit-exact Euler kernel), merges included: with
save_interval 1 the merges are compacted on the device
(``dem_merge_compact``), with a longer interval the reference's merge from
stale saved positions runs on the host."""

from __future__ import annotations

import numpy as np
import pytest

import oracle

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


def _run(solver_cls, s, tree=False, steps=12, dt=None, save_interval=1):
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
                        collision_detector=det, save_interval=save_interval, **kw)
    solver.solve(dt, dt * steps)
    return solver


def _chain(sol):
    out = []
    cur = sol.get_first_solution()
    while cur is not None:
        out.append((cur.radii.copy(), cur.get_trajectories().copy(), cur.get_times().copy(),
                    np.array(cur.initial_droplet_state.positions, copy=True)))
        cur = cur._next
    return out


def _dt(s):
    # a velocity scale from the first evaluation: dt moves droplets ~0.5 r_min
    fc = DirectDropletForceCalculator(num_particles=s.n, eps_oil=s.eps_oil, eta_oil=s.eta_oil,
                                      eta_water=s.eta_water, E=s.E, L=s.L)
    F, V = fc.calculate_forces_and_convection(s.positions, s.radii)
    vmax = np.max(np.linalg.norm(F / (s.stokes_factor * s.radii[:, None]) + V, axis=1))
    return 0.5 * s.radii.min() / vmax


@pytest.mark.parametrize("tree,save_interval", [(False, 1), (True, 1), (False, 3)])
def test_device_loop_is_bit_identical_to_host_loop(gpu, tree, save_interval):
    s = make_system(1500, seed=4, phi=0.04)
    dt = _dt(s)
    a = _run(EulerDropletSolver, s, tree, dt=dt, save_interval=save_interval)
    b = _run(DeviceEulerDropletSolver, s, tree, dt=dt, save_interval=save_interval)
    ca, cb = _chain(a.solution), _chain(b.solution)
    assert len(ca) == len(cb) and len(ca) >= 2, "the run should contain merges"
    for (ra, ta, ha, pa), (rb, tb, hb, pb) in zip(ca, cb):
        assert np.array_equal(ra, rb)
        assert np.array_equal(ta, tb)
        assert np.array_equal(ha, hb)
        # the reference advances each segment's initial positions in place
        assert np.array_equal(pa, pb)


def test_merge_compact_matches_merge_collided(gpu):
    import torch

    from dem import _device, _native
    from dem.solution.droplet_solution import merge_collided, merge_plan

    rng = np.random.default_rng(7)
    for n, k in ((2, 1), (50, 6), (4000, 300)):
        pos = rng.random((n, 3))
        r = rng.uniform(1e-3, 2e-3, n)
        # random pairs with repeats and droplets in several pairs (greedy skips)
        a = rng.integers(0, n, k)
        b = (a + 1 + rng.integers(0, max(n - 1, 1), k)) % n
        pairs = np.unique(np.sort(np.stack([a, b], 1), axis=1), axis=0).astype(np.int32)
        pairs = pairs[pairs[:, 0] != pairs[:, 1]]
        ref_p, ref_r, ref_c = merge_collided(pos, r, pairs)
        removed, new_r, new_p, coll = merge_plan(pairs, lambda i: r[i], lambda i: pos[i])
        np.testing.assert_array_equal(coll, ref_c)
        rem = np.asarray(sorted(removed), dtype=np.int32)
        m = len(new_r)
        merged = np.column_stack([np.asarray(new_p).reshape(m, 3), new_r])
        n_out = n - len(rem) + m
        out_p = torch.full((n_out, 3), np.nan, dtype=torch.float64, device="cuda")
        out_r = torch.full((n_out,), np.nan, dtype=torch.float64, device="cuda")
        # every device buffer stays referenced until the results are read: a
        # tensor freed right after its pointer is taken can be handed to the
        # next allocation before the kernel runs
        pd, rd = _device.to_device(pos), _device.to_device(r)
        rem_d, merged_d = _device.to_device(rem, np.int32), _device.to_device(merged)
        _native.call("dem_merge_compact", _device.ptr(pd), _device.ptr(rd), n,
                     _device.ptr(rem_d), len(rem), _device.ptr(merged_d), m, _device.ptr(out_p),
                     _device.ptr(out_r), _device.stream())
        assert np.array_equal(out_p.cpu().numpy(), ref_p)
        assert np.array_equal(out_r.cpu().numpy(), ref_r)
        # the oracle's merge, visiting the pairs in the same (np.argsort) order
        op, orr, _ = oracle.merge(pos, r, pairs[np.argsort(pairs[:, 0])])
        assert np.array_equal(out_p.cpu().numpy(), op)
        assert np.array_equal(out_r.cpu().numpy(), orr)


def test_config0_trajectory_against_fp64_oracle(gpu):
    """The plain direct workload on the CPU-path size: N = 2000, seed 0,
    phi = 0.02, E = 3e5 V/m, no coalescence, dt = 0.1 r_min / max|v(0)| from
    the FP64 oracle, 100 steps of the device-resident loop against the FP64
    oracle's rollout (C restatement of the force, Oseen and Euler kernels) on
    the same initial state.

    Tolerance, at every saved step: the displacement from the initial state,
    d(t) = x(t) - x(0) (minimum image), has field error
    |d - d_ref| / |d_ref| <= 1e-5 and a per-droplet error (floor 1e-3 rms)
    of at most 1e-3; positions agree to within 1e-3 r_min."""
    from ._errors import summary

    s = make_system(2000, seed=0, phi=0.02)
    F0 = oracle.direct_forces(s.positions, s.radii, s.m_const, s.L, True)
    V0 = oracle.direct_convection(s.positions, F0, s.eta_const, s.L, True)
    v0 = F0 / (s.stokes_factor * s.radii[:, None]) + V0
    dt = 0.1 * s.radii.min() / np.max(np.linalg.norm(v0, axis=1))
    fc = DirectDropletForceCalculator(num_particles=s.n, eps_oil=s.eps_oil, eta_oil=s.eta_oil,
                                      eta_water=s.eta_water, E=s.E, L=s.L,
                                      boundary_mode="periodic")
    sol = DropletSolution(DropletState(s.positions.copy(), s.radii.copy(), 0.0), length=128)
    solver = DeviceEulerDropletSolver(force_calculator=fc, solution=sol,
                                      post_processor=DropletPostProcessor(sol, box_size=s.L),
                                      collision_detector=None, verbose=False)
    solver.solve(dt, 100 * dt)
    traj = solver.solution.get_trajectories()
    assert solver.solution.get_first_solution() is solver.solution  # no merges
    steps = traj.shape[0] - 1
    assert steps >= 100
    x = s.positions.copy()
    worst = {"norm": 0.0, "max": 0.0, "abs": 0.0}
    for k in range(1, steps + 1):
        F = oracle.direct_forces(x, s.radii, s.m_const, s.L, True)
        V = oracle.direct_convection(x, F, s.eta_const, s.L, True)
        x = oracle.euler_step(x, F, V, s.radii, s.stokes_factor, dt, s.L, True)
        d = traj[k] - s.positions
        d_ref = x - s.positions
        d -= s.L * np.round(d / s.L)
        d_ref -= s.L * np.round(d_ref / s.L)
        e = summary(d, d_ref)
        gap = traj[k] - x
        gap -= s.L * np.round(gap / s.L)
        worst["norm"] = max(worst["norm"], e["norm"])
        worst["max"] = max(worst["max"], e["max"])
        worst["abs"] = max(worst["abs"], float(np.abs(gap).max() / s.radii.min()))
    print(f"\n[config 0 trajectory, {steps} steps] worst displacement field error "
          f"{worst['norm']:.3e}, per-droplet {worst['max']:.3e}, "
          f"position gap {worst['abs']:.3e} r_min")
    assert worst["norm"] <= 1e-5 and worst["max"] <= 1e-3 and worst["abs"] <= 1e-3
