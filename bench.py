"""Benchmark of the direct O(N^2) dipole force + Oseen convection step.

Workload: the plain direct DEM system at scale — N = 65536 droplets, seed 0,
volume fraction 0.02 in a periodic cube, log-normal radii (median 5 um,
sigma 0.3, clipped to [2, 20] um), E = 3e5 V/m, eps_oil = 2.85,
eta_oil = 0.05 Pa s.  One step = one full ``dem_direct_forces_and_convection``
call (quantise/pack, phase 1 dipole forces, phase 2 Oseen velocities, both
reductions) with positions already resident in HBM.

Metric: ordered pair-interactions per second, N(N-1) per step (each ordered
pair passes through both the force and the Oseen phase).

Also reported: ``roofline`` for the dominant pair kernel (its own HIP-event
duration, recorded by the library on the launch stream), ``cpu_baseline``
(the FP64 C/OpenMP oracle timed on a bounded sample of targets of the same
system on this host), and ``parity`` (FP32 result vs the library's FP64 checker kernel).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for _p in (str(ROOT), str(ROOT / "particle-simulator_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

PEAK_FP32 = 157.3e12  # MI355X dense FP32 vector (packed) FLOP/s
PEAK_FP32_UNPACKED = 78.6e12
FLOP_FORCE = 33  # algorithmic flops per ordered pair, force phase
FLOP_CONV = 31   # Oseen phase
# flops the pair-once kernels execute per ordered pair: one evaluation per
# unordered pair plus the reversed side's products and accumulations
# (force 33 + 6, Oseen 31 + 18, halved)
EXEC_FORCE_SYM = 19.5
EXEC_CONV_SYM = 24.5
# HBM bytes per launch at N = 65536 from two rocprofv3 --pmc runs (FETCH_SIZE
# scaled by 2.0, WRITE_SIZE by 1.0: the factors a plain 1 GiB copy reads in
# the same program); profiles/direct_force_oseen_n65536.md
TRAFFIC_N65536 = {"k_force_tile": (15384.9 * 2.0 + 12288.0) * 1024,
                  "k_conv_tile": (23247.7 * 2.0 + 12288.0) * 1024,
                  "k_force_sym": (17396.9 * 2.0 + 121481.7) * 1024,
                  "k_conv_sym": (52468.8 * 2.0 + 107136.0) * 1024}


def _errors(x, ref, floor=1e-3) -> dict:
    """Per-particle e_i = |x_i - ref_i| / max(|ref_i|, floor * rms|ref|), the
    share of particles with e_i > 1e-5 and the field error ||x - ref|| / ||ref||."""
    import numpy as np

    nr = np.linalg.norm(ref, axis=1)
    rms = float(np.sqrt(np.mean(nr**2)))
    e = np.linalg.norm(x - ref, axis=1) / np.maximum(nr, floor * rms)
    return {"max": float(e.max()), "p999": float(np.quantile(e, 0.999)),
            "frac_above_1e-5": float(np.mean(e > 1e-5)),
            "field": float(np.linalg.norm(x - ref) / np.linalg.norm(ref))}


def cpu_baseline(system, fc, F_any, budget_s: float) -> dict:
    """FP64 C/OpenMP oracle on the first T targets (all N sources), both phases."""
    import oracle

    n = system.n
    calib = min(n, 2048)
    t0 = time.perf_counter()
    oracle.direct_forces(system.positions, system.radii, fc.m_const, system.L, True,
                         targets=(0, calib))
    oracle.direct_convection(system.positions, F_any, fc.eta_const, system.L, True,
                             targets=(0, calib))
    per_target = (time.perf_counter() - t0) / calib
    T = int(max(64, min(n, budget_s / max(per_target, 1e-9))))
    t0 = time.perf_counter()
    oracle.direct_forces(system.positions, system.radii, fc.m_const, system.L, True, targets=(0, T))
    t1 = time.perf_counter()
    oracle.direct_convection(system.positions, F_any, fc.eta_const, system.L, True, targets=(0, T))
    t2 = time.perf_counter()
    rate = T * (n - 1) / (t2 - t0)
    return {
        "value": rate,
        "unit": "pair-interactions/s",
        "cores": oracle.num_threads(),
        "kind": "port",
        "sample": f"{T} of {n} targets x {n} sources, force + Oseen phases, float64 "
                  f"({t2 - t0:.1f} s)",
    }


def config0_cpu(steps: int = 100) -> dict:
    """The FP64 C/OpenMP oracle stepping the N = 2000 system (forces, Oseen
    velocities, Euler step and wrap) for `steps` steps: the CPU side of the
    plain direct workload at the reference's CPU-runnable size."""
    import numpy as np

    import oracle
    from dem.workloads import make_system

    s = make_system(2000, seed=0, phi=0.02)
    F = oracle.direct_forces(s.positions, s.radii, s.m_const, s.L, True)
    V = oracle.direct_convection(s.positions, F, s.eta_const, s.L, True)
    v0 = F / (s.stokes_factor * s.radii[:, None]) + V
    dt = 0.1 * s.radii.min() / np.max(np.linalg.norm(v0, axis=1))
    x = s.positions.copy()
    t0 = time.perf_counter()
    for _ in range(steps):
        F = oracle.direct_forces(x, s.radii, s.m_const, s.L, True)
        V = oracle.direct_convection(x, F, s.eta_const, s.L, True)
        x = oracle.euler_step(x, F, V, s.radii, s.stokes_factor, dt, s.L, True)
    ms = (time.perf_counter() - t0) / steps * 1e3
    return {"ms_per_step": ms, "pair_interactions_per_s": s.n * (s.n - 1) / (ms * 1e-3),
            "steps": steps, "cores": oracle.num_threads(), "kind": "port"}


def config0_leg(steps: int = 100) -> dict:
    """The plain direct workload at N = 2000 (seed 0, phi = 0.02, periodic,
    E = 3e5 V/m): 100 steps of the device-resident solver without contact
    detection (no coalescence), dt = 0.1 r_min / max|v(0)| from the FP32
    path; positions copied to the host trajectory every step."""
    import numpy as np
    import torch

    from dem.force_calculator import DirectDropletForceCalculator
    from dem.particle_state import DropletState
    from dem.post_processor import DropletPostProcessor
    from dem.solution import DropletSolution
    from dem.solver.device_solver import DeviceEulerDropletSolver
    from dem.workloads import make_system

    s = make_system(2000, seed=0, phi=0.02)
    fc = DirectDropletForceCalculator(num_particles=s.n, eps_oil=s.eps_oil, eta_oil=s.eta_oil,
                                      eta_water=s.eta_water, E=s.E, L=s.L,
                                      boundary_mode="periodic")
    F, V = fc.calculate_forces_and_convection(s.positions, s.radii)
    vmax = np.max(np.linalg.norm(F / (s.stokes_factor * s.radii[:, None]) + V, axis=1))
    dt = 0.1 * s.radii.min() / vmax

    def run(nsteps):
        sol = DropletSolution(DropletState(s.positions.copy(), s.radii.copy(), 0.0),
                              length=nsteps + 8)
        solver = DeviceEulerDropletSolver(force_calculator=fc, solution=sol,
                                          post_processor=DropletPostProcessor(sol, box_size=s.L),
                                          collision_detector=None, verbose=False)
        t0 = time.perf_counter()
        solver.solve(dt, (nsteps - 0.5) * dt)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, solver

    run(5)
    el, solver = run(steps)
    k = solver.steps_done
    ms = el / max(k, 1) * 1e3
    return {"workload": "N=2000, phi=0.02, periodic, seed 0, direct operator, no coalescence",
            "steps": k, "ms_per_step": ms, "pair_interactions_per_s": s.n * (s.n - 1) / (ms * 1e-3),
            "note": "device-resident loop, positions copied to the host trajectory every "
                    "step; at this size a step is a handful of dependent launches"}


def hash_pass(s, reps: int = 20, radius_scale: float = 1.6) -> dict:
    """Spatial-hash contact pass on the same positions with radii x 1.6 (a
    system with contacts, so the emit path is timed too); device-resident
    inputs, the pair count read back by the host as the drop-in API does."""
    import torch

    from dem import _device
    from dem.collision_detector import SpatialHashCollisionDetector

    det = SpatialHashCollisionDetector(num_particles=s.n, L=s.L, boundary_mode="periodic")
    pd, rd = _device.to_device(s.positions), _device.to_device(s.radii * radius_scale)
    det.detect_device(pd, rd, s.L, True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        k, _pairs = det.detect_device(pd, rd, s.L, True)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / reps * 1e3
    n, T = s.n, det._hash_table_size
    # compulsory traffic: assign N(16+4), zero/scan 3T*4, scatter N(4+4),
    # candidate scan N(16 + 27*8) + 8K (pairs)
    nbytes = n * 20 + 3 * T * 4 + n * 8 + n * (16 + 27 * 8) + 8 * k
    return {"ms_per_pass": ms, "pairs": int(k), "radius_scale": radius_scale,
            "algorithmic_bytes": nbytes,
            "achieved_GBps": nbytes / (ms * 1e-3) / 1e9, "peak_GBps": 8000.0,
            "note": "includes the host read of the pair count (drop-in semantics): the emit "
                    "kernel stores it in pinned host memory as soon as the pair counts are "
                    "scanned and the host polls it; the pairs are complete in stream order"}


def step_leg(steps: int = 20) -> dict:
    """End-to-end DEM step at N = 65536 (phi = 0.02, periodic): the reference
    loop (contact detection, merges, forces + Oseen velocities, Euler step and
    wrap) on the device-resident solver, dt = 0.1 r_min / max|v(0)|."""
    import numpy as np
    import torch

    from dem.collision_detector import SpatialHashCollisionDetector
    from dem.force_calculator import DirectDropletForceCalculator
    from dem.particle_state import DropletState
    from dem.post_processor import DropletPostProcessor
    from dem.solution import DropletSolution
    from dem.solver.device_solver import DeviceEulerDropletSolver
    from dem.workloads import make_system

    s = make_system(65536, seed=0, phi=0.02)
    fc = DirectDropletForceCalculator(num_particles=s.n, eps_oil=s.eps_oil, eta_oil=s.eta_oil,
                                      eta_water=s.eta_water, E=s.E, L=s.L,
                                      boundary_mode="periodic")
    F, V = fc.calculate_forces_and_convection(s.positions, s.radii)
    vmax = np.max(np.linalg.norm(F / (s.stokes_factor * s.radii[:, None]) + V, axis=1))
    dt = 0.1 * s.radii.min() / vmax
    det = SpatialHashCollisionDetector(num_particles=s.n, L=s.L, boundary_mode="periodic")

    def run(nsteps):
        # a fresh solution per run: like the reference's, a solution compacted at the
        # end of solve() does not take further steps
        sol = DropletSolution(DropletState(s.positions.copy(), s.radii.copy(), 0.0), length=8)
        pp = DropletPostProcessor(sol, box_size=s.L)
        # save_interval 1 (the reference's default): every step's positions reach the
        # host trajectory, and merges start from the current positions
        solver = DeviceEulerDropletSolver(force_calculator=fc, solution=sol, post_processor=pp,
                                          collision_detector=det, save_interval=1,
                                          verbose=False)
        t0 = time.perf_counter()
        solver.solve(dt, (nsteps - 0.5) * dt)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, solver

    run(2)
    el, solver = run(steps)
    k = solver.steps_done
    ms = el / max(k, 1) * 1e3
    n = s.n
    return {"workload": "N=65536, phi=0.02, periodic, seed 0, direct operator, hash every step",
            "dt_s": dt, "steps": k, "merges": solver.merges, "ms_per_step": ms,
            "pair_interactions_per_s": n * (n - 1) / (ms * 1e-3),
            "note": "device-resident loop (DeviceEulerDropletSolver), save_interval 1: contact "
                    "pass with the pair count read back every step, positions copied to the "
                    "host trajectory every step (pinned staging); merge plans on the host "
                    "in float64 (the reference's greedy arithmetic), compaction on the "
                    "device (dem_merge_compact)"}


def pbm_counts(reps: int = 5) -> dict:
    """Population-balance collision counts at N = 65536 (phi = 0.02, periodic,
    30 volume classes): the tree-pruned walk (leaf 16, after a build) and the
    all-pairs float64 kernel."""
    import numpy as np
    import torch

    from dem import _device, _native
    from dem.octree.flat_tree import FlatOctree
    from dem.workloads import make_system
    from pbm import VolumeGrid
    from pbm.collision_frequency import DirectCollisionFrequency
    from pbm.volume_grid import sphere_volumes

    s = make_system(65536, seed=0, phi=0.02)
    g = VolumeGrid.from_radii_range(float(s.radii.min()), float(s.radii.max()), 30)
    bins = _device.to_device(g.bin_indices(sphere_volumes(s.radii)).astype(np.int32), np.int32)
    pd, rd = _device.to_device(s.positions), _device.to_device(s.radii)
    tree = FlatOctree(theta=0.5, mpl=16, num_particles=s.n)
    tree.build_device(pd, rd, s.L, True)
    dc = DirectCollisionFrequency(g)
    for _ in range(2):
        tree.collision_counts_device(bins, 30)
        dc.compute_device(pd, rd, bins, s.L, True)
    torch.cuda.synchronize()
    lib = _native.load()
    lib.dem_profile_start()
    for _ in range(reps):
        M = tree.collision_counts_device(bins, 30)
        D = dc.compute_device(pd, rd, bins, s.L, True)
    torch.cuda.synchronize()
    lib.dem_profile_stop()
    k = {name: _native.profile_read(name)[0] / reps
         for name in ("k_walk_collisions", "k_pair_collisions")}
    return {"workload": "N=65536, phi=0.02, periodic, 30 volume classes, tree leaf 16",
            "tree_walk_ms": k["k_walk_collisions"], "all_pairs_ms": k["k_pair_collisions"],
            "all_pairs_per_s": s.n * (s.n - 1) / 2 / (k["k_pair_collisions"] * 1e-3),
            "touching_pairs_tree": int(M.sum().item()) // 2,
            "touching_pairs_all": int(D.sum().item()) // 2}


def tree_step(reps: int = 5) -> dict:
    """Barnes-Hut step at N = 2^20, theta = 0.5, leaf size 16, periodic cube:
    build + force walk + total-velocity walk; error vs FP64 direct on 8192
    seeded targets (reference metric: relative error with a 1 %-of-mean floor)."""
    import numpy as np
    import torch

    from dem import _device, _native
    from dem.force_calculator import DirectDropletForceCalculator
    from dem.octree import TreeDropletForceCalculator
    from dem.workloads import make_system

    n = 1 << 20
    s = make_system(n, seed=0, phi=0.02)
    tc = TreeDropletForceCalculator(num_particles=n, theta=0.5, mpl=16, eps_oil=s.eps_oil,
                                    eta_oil=s.eta_oil, eta_water=s.eta_water, E=s.E, L=s.L,
                                    periodic=True)
    pd, rd = _device.to_device(s.positions), _device.to_device(s.radii)
    for _ in range(2):
        F, V = tc.forces_and_total_velocity_device(pd, rd, s.stokes_factor)
    torch.cuda.synchronize()
    lib = _native.load()
    lib.dem_profile_start()
    t0 = time.perf_counter()
    for _ in range(reps):
        F, V = tc.forces_and_total_velocity_device(pd, rd, s.stokes_factor)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / reps * 1e3
    lib.dem_profile_stop()
    walks = {k: _native.profile_read(k)[0] / reps for k in ("k_walk_forces", "k_walk_total")}
    t0 = time.perf_counter()
    for _ in range(reps):
        tc.octree.build_device(pd, rd, s.L, True)
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t0) / reps * 1e3
    dc = DirectDropletForceCalculator(num_particles=n, eps_oil=s.eps_oil, eta_oil=s.eta_oil,
                                      eta_water=s.eta_water, E=s.E, L=s.L)
    Fall = dc.forces_device(pd, rd)
    tg = np.sort(np.random.default_rng(0).choice(n, 8192, replace=False)).astype(np.int32)
    Ft, Vc = dc.f64_targets_device(pd, rd, _device.to_device(tg, np.int32), Fall)
    Ft, Vc = Ft.cpu().numpy(), Vc.cpu().numpy()
    vt_ref = Vc + Ft / (s.stokes_factor * s.radii[tg, None])

    def err(ref, x):
        nr = np.linalg.norm(ref, axis=1)
        e = np.linalg.norm(x - ref, axis=1) / (nr + 0.01 * nr.mean())
        return {"mean_pct": float(e.mean() * 100), "max_pct": float(e.max() * 100)}

    info = tc.octree._info
    # algorithmic bytes (HBM roofline, 8 TB/s): build N (24 key/index + 24 p radix
    # passes, p = ceil(3 D / 8) + 32 permute + 8 D level scans + 16 leaf
    # aggregates) + nodes x 96 B; walks compulsory N (16 in + 24 out) + nodes x
    # 64 B (force) or + nodes x 128 B and 24 B more per target (total velocity)
    D = int(info.max_depth)
    M = int(info.node_count)
    p = -(-3 * D // 8)
    b_build = n * (24 + 24 * p + 32 + 8 * D + 16) + M * 96
    b_force = n * (16 + 24) + M * 64
    b_total = n * (16 + 24 + 24) + M * 128
    hbm = {"build_bytes": b_build, "build_GBps": b_build / (build_ms * 1e-3) / 1e9,
           "force_walk_bytes": b_force,
           "force_walk_GBps": b_force / (walks["k_walk_forces"] * 1e-3) / 1e9,
           "total_walk_bytes": b_total,
           "total_walk_GBps": b_total / (walks["k_walk_total"] * 1e-3) / 1e9,
           "peak_GBps": 8000.0,
           "note": "compulsory bytes; node arrays and leaf records are L2/MALL-resident, so "
                   "the walks are bound by their dependent node loads and FP32 interactions, "
                   "the build by its ~100 dependent launches"}
    return {
        "workload": "N=1048576, phi=0.02, periodic, theta=0.5, leaf 16, seed 0",
        "ms_per_step": ms, "build_ms": build_ms,
        "walk_ms": walks, "hbm": hbm, "nodes": M, "max_nodes": int(tc.octree.max_nodes),
        "depth": D,
        "force_err_vs_fp64_direct": err(Ft, F.cpu().numpy()[tg]),
        "total_velocity_err_vs_fp64_direct": err(vt_ref, V.cpu().numpy()[tg]),
        "note": "errors are the reference algorithm's Barnes-Hut errors (the GPU walk matches "
                "the CPU restatement of the reference walk to FP32 accuracy)",
    }


def ensembles_leg(systems, steps: int, plain_steps: int, tune: bool = True) -> dict:
    """256 uneven periodic systems: batched graph step vs the single-system
    path run system after system.  Both include contact checks, host merges
    and Euler steps; rates count Euler steps actually taken (a system that
    merges does not advance in that global step)."""
    import numpy as np
    import torch

    from dem import _native
    from dem.batch import BatchedEnsembles, SequentialEnsembles, Tuning, autotune
    from dem.force_calculator import DirectDropletForceCalculator

    # dt: the fastest droplet moves 0.1 r_min in the first step (FP32 path)
    dts = []
    for sy in systems:
        fc = DirectDropletForceCalculator(num_particles=sy.n, eps_oil=sy.eps_oil,
                                          eta_oil=sy.eta_oil, eta_water=sy.eta_water, E=sy.E,
                                          L=sy.L, boundary_mode="periodic")
        F, V = fc.calculate_forces_and_convection(sy.positions, sy.radii)
        v = np.linalg.norm(F / (sy.stokes_factor * sy.radii[:, None]) + V, axis=1)
        dts.append(0.1 * sy.radii.min() / v.max())
    ns = np.array([sy.n for sy in systems], dtype=np.float64)
    big = 1 << 30
    times = None
    best = Tuning(256, 4096, pair_once=True)  # the autotuner's choice on this workload
    if tune:
        best, times = autotune(systems, dts)
    eng = BatchedEnsembles(systems, dts, [big] * len(systems), best, capture=True)
    eng.run(2)
    torch.cuda.synchronize()
    d0 = eng.steps_done.cpu().numpy().astype(np.int64)
    lib = _native.load()
    t0 = time.perf_counter()
    eng.run(steps)
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    adv_b = int((eng.steps_done.cpu().numpy().astype(np.int64) - d0).sum())
    merges_b = eng.merges
    rounds_b = eng.merge_rounds
    # per-phase wall time on a few more steps, each phase synchronised
    eng.phase_ms = {}
    prof_steps = max(1, min(5, steps))
    eng.run(prof_steps)
    phases = {k: v / prof_steps for k, v in eng.phase_ms.items()}
    eng.close()
    del eng

    seq = SequentialEnsembles(systems, dts, [big] * len(systems))
    seq.step_all()
    torch.cuda.synchronize()
    a0 = int(seq.steps_done.sum())
    t0 = time.perf_counter()
    for _ in range(plain_steps):
        seq.step_all()
    torch.cuda.synchronize()
    tp = time.perf_counter() - t0
    adv_p = int(seq.steps_done.sum()) - a0
    rate_b = adv_b / tb
    rate_p = adv_p / tp
    pairs = float((ns * (ns - 1)).sum())
    return {
        "metric": "ensemble-steps/s batched vs plain",
        "systems": len(systems),
        "sum_n": int(ns.sum()),
        "pairs_per_global_step": pairs,
        "batched": {"ensemble_steps_per_s": rate_b, "global_steps": steps,
                    "ms_per_global_step": tb / steps * 1e3, "ensemble_steps": adv_b,
                    "merges": merges_b, "merge_rounds": rounds_b,
                    "phase_ms_per_global_step": phases,
                    "tuning": {"block": best.block, "jlen": best.jlen,
                               "pair_once": bool(best.pair_once)},
                    "pair_interactions_per_s": pairs * adv_b / len(systems) / tb,
                    "autotune_ms_per_step": ({(f"{b}x{j}" if b != "pair_once" else f"pair_once_{j}"): v
                                              for (b, j), v in times.items()}
                                             if times else None)},
        "plain": {"ensemble_steps_per_s": rate_p, "global_steps": plain_steps,
                  "ms_per_global_step": tp / plain_steps * 1e3, "ensemble_steps": adv_p,
                  "merges": seq.merges,
                  "path": "single-system device path per system (hash + count sync, "
                          "dem_direct_forces_and_convection, dem_euler_step)"},
        "speedup": rate_b / rate_p if rate_p > 0 else None,
        "note": "dt_e = 0.1 r_min / max|v(0)| per system; the batched rate is timed over "
                "the whole run of global_steps steps, the plain rate over its own steps"
    }


def adjoint_leg(steps: int = 200) -> dict:
    """Differentiable direct DEM: N = 4096, seed 0, phi = 0.02, T Euler steps
    without merging, loss = sum_{i<j} softplus(1.05 (r_i + r_j) - |r_ij|) in
    micrometres; gradients w.r.t. E, mu_oil and x0 by the hand-written adjoint
    (checkpoints every 20 steps).  Timed here; the agreement of the same
    gradients with autograd through the plain FP64 path and with FP64 central
    differences on this workload is checked by
    tests/test_gpu_adjoint.py::test_full_workload_gradients_match_autograd_and_fd."""
    import numpy as np
    import torch

    from dem.adjoint import DirectTrajectory, contact_loss, rollout
    from dem.force_calculator import DirectDropletForceCalculator
    from dem.workloads import differentiable, differentiable_dt

    s = differentiable()
    E0, mu0, mu_w, eps = 3e5, 0.05, 0.001, 2.85
    dev = torch.device("cuda")
    dt = differentiable_dt(s, E0, mu0, eps, mu_w)
    traj = DirectTrajectory(s.radii, s.L, dt, steps, eps_oil=eps, eta_water=mu_w,
                            checkpoint_every=20)
    x0 = torch.tensor(s.positions, device=dev, requires_grad=True)
    E = torch.tensor(E0, dtype=torch.float64, requires_grad=True)
    mu = torch.tensor(mu0, dtype=torch.float64, requires_grad=True)
    # warm-up forward + backward (builds workspaces), then timed forward + backward
    xw = torch.tensor(s.positions, device=dev, requires_grad=True)
    contact_loss(rollout(xw, E0, mu0, traj), traj).backward()
    torch.cuda.synchronize()
    rec0 = traj.recomputed_steps
    t0 = time.perf_counter()
    loss = contact_loss(rollout(x0, E, mu, traj), traj)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    loss.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    return {
        "workload": f"differentiable direct DEM, N={s.n}, T={steps}, phi=0.02, periodic, seed 0",
        "loss": loss.item(), "dt_s": dt,
        "forward_ms": (t1 - t0) * 1e3, "backward_ms": (t2 - t1) * 1e3,
        "ms_per_step_forward": (t1 - t0) * 1e3 / steps,
        "ms_per_step_backward": (t2 - t1) * 1e3 / steps,
        "recomputed_steps": traj.recomputed_steps - rec0,
        "grad": {"E": E.grad.item(), "mu_oil": mu.grad.item(),
                 "x0_norm": float(np.linalg.norm(x0.grad.cpu().numpy()))},
        "checked_by": "tests/test_gpu_adjoint.py::test_full_workload_gradients_match_autograd_and_fd "
                      "(FP64 autograd and FP64 central differences, tolerance 1e-4)",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=int(os.environ.get("WORLD_SIZE", "1")))
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=65536)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the hash-pass, Barnes-Hut tree and ensemble measurements")
    ap.add_argument("--ens-steps", type=int, default=500,
                    help="timed batched global steps of the 256-system workload")
    ap.add_argument("--ens-plain-steps", type=int, default=500,
                    help="timed steps of the system-after-system comparison")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 1 GiB device copy (scale for PMC byte counters)")
    ap.add_argument("--one-sided", action="store_true",
                    help="time the one-sided direct kernels instead of the pair-once ones")
    ap.add_argument("--legs", default="config0,hash,step,tree,pbm,ensembles,adjoint",
                    help="comma list of extra measurements: config0, hash, step, tree, pbm, ensembles, "
                         "adjoint")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported "
              "(the multi-GPU sharding is design-only)", file=sys.stderr)
        return 2

    ens_systems = None
    legs = set() if a.no_extras else {x.strip() for x in a.legs.split(",") if x.strip()}
    if "ensembles" in legs and a.ens_steps > 0:
        from dem.workloads import ensembles

        # placement runs in forked workers: before the GPU is initialised
        ens_systems = ensembles(256, base_seed=0)

    import numpy as np
    import torch

    from dem import _device, _native
    from dem.force_calculator import DirectDropletForceCalculator
    from dem.workloads import make_system

    torch.cuda.set_device(0)
    if a.one_sided:
        _native.call("dem_direct_set_symmetric", 0)
    s = make_system(a.n, seed=0, phi=0.02)
    fc = DirectDropletForceCalculator(num_particles=s.n, eps_oil=s.eps_oil, eta_oil=s.eta_oil,
                                      eta_water=s.eta_water, E=s.E, L=s.L,
                                      boundary_mode="periodic")
    pd, rd = _device.to_device(s.positions), _device.to_device(s.radii)
    # the FP64 checker of the parity block runs first (~60 ms of GPU work), so
    # the timed steps start on a GPU that has been busy, as in a long run
    F64, V64 = fc.forces_and_convection_f64_device(pd, rd)
    for _ in range(a.warmup):
        F, V = fc.forces_and_convection_device(pd, rd)
    torch.cuda.synchronize()
    lib = _native.load()
    lib.dem_profile_start()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        F, V = fc.forces_and_convection_device(pd, rd)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    lib.dem_profile_stop()
    ms_step = (t1 - t0) / a.steps * 1e3
    if a.copy_calibration:
        # a copy of known size: 1 GiB read + 1 GiB written per launch
        src = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    npairs = s.n * (s.n - 1)
    value = npairs / (ms_step * 1e-3)

    kern = {}
    for name, flop, ex in (("k_force_tile", FLOP_FORCE, FLOP_FORCE),
                           ("k_conv_tile", FLOP_CONV, FLOP_CONV),
                           ("k_force_sym", FLOP_FORCE, EXEC_FORCE_SYM),
                           ("k_conv_sym", FLOP_CONV, EXEC_CONV_SYM)):
        tot, cnt = _native.profile_read(name)
        if cnt == 0:
            continue
        avg = tot / cnt
        kern[name] = {"avg_ms": avg, "launches": cnt, "flops": flop * npairs,
                      "tflops": flop * npairs / (avg * 1e-3) / 1e12 if avg > 0 else 0.0,
                      "executed_tflops": ex * npairs / (avg * 1e-3) / 1e12 if avg > 0 else 0.0}
    dom = max(kern, key=lambda k: kern[k]["avg_ms"])
    achieved = kern[dom]["flops"] / (kern[dom]["avg_ms"] * 1e-3)
    roof = {"bound": "valu", "compute": "FP32 VALU (pair sums are not a contraction: no MFMA "
                                       "instructions; peak = MI355X FP32 vector, packed)",
            "kernel": dom, "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12,
            "unit": "TFLOP/s", "frac": achieved / PEAK_FP32,
            "frac_unpacked_peak": achieved / PEAK_FP32_UNPACKED,
            "executed": kern[dom]["executed_tflops"],
            "executed_frac_unpacked_peak": kern[dom]["executed_tflops"] * 1e12 / PEAK_FP32_UNPACKED,
            "note": "achieved counts the algorithmic flops of every ordered pair; the "
                    "pair-once kernels evaluate each unordered pair once (executed)",
            "traffic": TRAFFIC_N65536.get(dom) if s.n == 65536 else None,
            "traffic_unit": "bytes per launch",
            "traffic_source": "recorded constant, not measured in this run: FETCH_SIZE and "
                              "WRITE_SIZE of rocprofv3 --pmc runs of this kernel at N = 65536, "
                              "scaled by a plain copy's counters "
                              "(profiles/direct_force_oseen_n65536.md)",
            "kernels": kern}

    # the drop-in API: numpy float64 in and out (host<->device copies included)
    fc.calculate_forces_and_convection(s.positions, s.radii)
    th0 = time.perf_counter()
    for _ in range(3):
        fc.calculate_forces_and_convection(s.positions, s.radii)
    host_ms = (time.perf_counter() - th0) / 3 * 1e3

    torch.cuda.synchronize()
    Fn, Vn = F.cpu().numpy(), V.cpu().numpy()
    parity = {"force": _errors(Fn, F64.cpu().numpy()),
              "velocity": _errors(Vn, V64.cpu().numpy()),
              "reference": "FP64 checker kernel on the same float64 inputs; per-particle "
                           "error floored at 1e-3 rms"}

    base = None
    if not a.no_cpu_baseline:
        base = cpu_baseline(s, fc, Fn, a.cpu_seconds)
    extras = {}
    if "config0" in legs:
        extras["config0"] = config0_leg()
        if base is not None:
            extras["config0"]["cpu_baseline"] = config0_cpu()
    if "hash" in legs:
        extras["hash"] = hash_pass(s)
    if "tree" in legs:
        extras["tree"] = tree_step()
    if "pbm" in legs:
        extras["pbm"] = pbm_counts()
    if "step" in legs:
        extras["step"] = step_leg()
    if ens_systems is not None:
        extras["ensembles"] = ensembles_leg(ens_systems, a.ens_steps, a.ens_plain_steps,
                                            tune=not a.no_autotune)
    if "adjoint" in legs:
        extras["adjoint"] = adjoint_leg()

    out = {
        "metric": "pair-interactions/s direct N=65536; ensemble-steps/s batched vs plain",
        "value": value,
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded log-normal droplets, periodic cube)",
        "config": {"workload": f"direct DEM force+Oseen, N={s.n}, phi=0.02, periodic, seed 0",
                   "n": s.n, "L_m": s.L, "parallelism": "single GPU"},
        "drop_in_host_arrays": {"ms_per_step": host_ms, "pair_interactions_per_s":
                                npairs / (host_ms * 1e-3),
                                "note": "calculate_forces_and_convection on numpy arrays: "
                                        "PCIe copies in and out included (not the metric)"},
        "roofline": roof,
        "cpu_baseline": base,
        "parity": parity,
        **extras,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
