"""Euler time loop whose state stays on the GPU.

Same control flow and results as ``EulerDropletSolver`` (the reference's
loop, dem/solver/euler_droplet_solver.py:43-159) — contact detection first,
on contact a volume-conserving merge and a restart of the step without
advancing t, otherwise forces/velocities, the Euler update and the periodic
wrap — but positions and radii live in device memory between steps:

* contact detection: ``dem_hash_pairs`` on the device arrays; only the pair
  count crosses to the host each step;
* merge (rare): the pairs and the state are copied to the host and merged by
  ``DropletSolution.generate_next_solution`` — the reference's float64 numpy
  merge, bit-exact — then copied back;
* forces + velocities: the operator's ``*_device`` entry point;
* Euler update and wrap: ``dem_euler_step`` / ``dem_euler_step_total``,
  bit-exact with the reference's numpy update;
* ``solution.save_step`` receives a host copy only on saved steps.

Given identical operator outputs, trajectories are bit-identical to
``EulerDropletSolver``'s.
"""

from __future__ import annotations

import time

import numpy as np

from .. import _device, _native


class DeviceEulerDropletSolver:
    def __init__(self, force_calculator, solution, post_processor, collision_detector=None,
                 save_interval: int = 1, pbm_coupling=None, verbose: bool = True):
        self.force_calculator = force_calculator
        self.solution = solution
        self.post_processor = post_processor
        self.collision_detector = collision_detector
        self.pbm_coupling = pbm_coupling
        self.simulation_time = 0
        self.save_interval = max(save_interval, 1)
        self.verbose = verbose
        eo, ew = force_calculator.eta_oil, force_calculator.eta_water
        self.stokes_factor = 2 * np.pi * eo * (2 * eo + 3 * ew) / (eo + ew)
        self.steps_done = 0
        self.merges = 0

    def _upload(self):
        st = self.solution.initial_droplet_state
        return (_device.to_device(np.asarray(st.positions, dtype=np.float64).reshape(-1, 3)),
                _device.to_device(np.asarray(st.radii, dtype=np.float64)))

    def _forces(self, pos_d, r_d):
        fc = self.force_calculator
        if hasattr(fc, "calculate_forces_and_total_velocity"):
            return "total", fc.forces_and_total_velocity_device(pos_d, r_d, self.stokes_factor)
        return "direct", fc.forces_and_convection_device(pos_d, r_d)

    def solve(self, dt, total_time):
        fc = self.force_calculator
        t = float(self.solution.get_current_time())
        t_start, t_end = t, t + float(total_time)
        time_start = time.time()
        last_percent = -1
        is_aborted = False
        periodic = fc.boundary_mode == "periodic"
        L = float(fc.L)
        while t < t_end and not is_aborted:
            if self.post_processor.stop_simulation:
                break
            pos_d, r_d = self._upload()
            radii_host = self.solution.initial_droplet_state.radii
            n = int(pos_d.shape[0])
            remaining = int((t_end - t) / dt) + 1
            for step in range(remaining):
                percent = int(100 * (t - t_start) / max(t_end - t_start, 1e-12))
                if self.verbose and percent != last_percent:
                    last_percent = percent
                    print(f"{percent}%: time {t:.3f} of {t_end:.3f}, droplets: {n}, "
                          f"elapsed: {time.time() - time_start:.1f} s", flush=True)
                if self.post_processor.stop_simulation:
                    is_aborted = True
                    break
                det = self.collision_detector
                if det is not None:
                    k, pairs_d = det.detect_device(pos_d, r_d, det._L, det._boundary_mode == "periodic")
                    if k > 0:
                        # as in the reference, the merge starts from the last *saved*
                        # positions (trajectories[current_step]); with save_interval = 1
                        # those are the current ones
                        pairs = pairs_d.cpu().numpy()
                        if self.pbm_coupling is not None:
                            self.pbm_coupling.on_collision(pairs, radii_host)
                        sol = self.solution
                        sol.is_collision = True
                        sol.collided_droplets = pairs
                        sol.compact()
                        self.solution = sol.generate_next_solution()
                        self.post_processor.update_solution(self.solution)
                        self.merges += 1
                        break
                kind, (F, V) = self._forces(pos_d, r_d)
                if kind == "total":
                    _native.call("dem_euler_step_total", _device.ptr(pos_d), _device.ptr(V), n,
                                 float(dt), L, 1 if periodic else 0, _device.stream())
                else:
                    _native.call("dem_euler_step", _device.ptr(pos_d), _device.ptr(F),
                                 _device.ptr(V), _device.ptr(r_d), n, float(self.stokes_factor),
                                 float(dt), L, 1 if periodic else 0, _device.stream())
                if self.pbm_coupling is not None:
                    self.pbm_coupling.step(t + dt, radii_host, dt)
                t += dt
                self.steps_done += 1
                if (self.save_interval <= 1 or step % self.save_interval == 0
                        or step == remaining - 1):
                    self.solution.save_step(t, pos_d.cpu().numpy())
                if self.solution.should_update_visualization():
                    self.post_processor.live_plot()
        self.solution.compact_all()
        self.simulation_time = time.time() - time_start
        if self.verbose:
            print(f"Total compute time: {self.simulation_time:.2f} s")
        if self.solution.real_time_visualization:
            self.post_processor.finalize_live_plot()
