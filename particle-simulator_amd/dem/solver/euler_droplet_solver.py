"""Explicit Euler time loop (host orchestration).

Drop-in for the reference's ``EulerDropletSolver``
(dem/solver/euler_droplet_solver.py:8-159) with the same control flow, which
callers depend on:
* stokes_factor = 2 pi eta_o (2 eta_o + 3 eta_w) / (eta_o + eta_w) (:37-41);
* per step: contact detection; on contact -> optional PBM hook, compact,
  merge into a new solution segment and restart the step loop *without*
  advancing t (:96-110); otherwise forces/velocities via
  ``calculate_forces_and_total_velocity`` when the operator has it (tree)
  or ``calculate_forces_and_convection`` (direct), Euler update, periodic
  wrap ``positions %= L`` (:112-131), PBM step, save (:134-144);
* positions alias the current segment's initial state (:71).

The operators it drives (direct / tree / hash) run on the GPU; this loop
keeps the reference's numpy state so that any caller of the reference's
solver sees identical objects.  For a loop whose state stays on the GPU use
``dem.solver.device_solver.DeviceEulerDropletSolver``.
"""

from __future__ import annotations

import time

import numpy as np

from .solver_base import Solver


class EulerDropletSolver(Solver):
    def __init__(
        self,
        force_calculator,
        solution,
        post_processor,
        collision_detector=None,
        save_interval: int = 1,
        pbm_coupling=None,
    ):
        self.force_calculator = force_calculator
        self.solution = solution
        self.post_processor = post_processor
        self.collision_detector = collision_detector
        self.pbm_coupling = pbm_coupling
        self.simulation_time = 0
        self.save_interval = max(save_interval, 1)

        eta_oil = self.force_calculator.eta_oil
        eta_water = self.force_calculator.eta_water
        self.stokes_factor = (
            2 * np.pi * eta_oil * (2 * eta_oil + 3 * eta_water) / (eta_oil + eta_water)
        )

    def solve(self, dt, total_time):
        t = float(self.solution.get_current_time())
        t_start = t
        t_end = t_start + float(total_time)
        is_aborted = False
        time_start = time.time()
        last_percent = -1

        while (t < t_end) and not is_aborted:
            if self.post_processor.stop_simulation:
                print("Simulation stopped by the user.")
                break

            positions = self.solution.initial_droplet_state.positions
            radii = self.solution.initial_droplet_state.radii

            remaining_steps = int((t_end - t) / dt) + 1

            for step in range(remaining_steps):
                progress_denom = max(t_end - t_start, 1e-12)
                percent = int(100 * (t - t_start) / progress_denom)
                if percent != last_percent:
                    last_percent = percent
                    elapsed = time.time() - time_start
                    print(
                        f"{percent}%: time {t:.3f} of {t_end:.3f}, "
                        f"droplets: {self.solution.num_particles}, "
                        f"elapsed: {elapsed:.1f} s",
                        flush=True,
                    )

                if self.post_processor.stop_simulation:
                    print("Simulation stopped by the user.")
                    is_aborted = True
                    break

                if self.collision_detector is not None:
                    is_collision, collided_pairs = self.collision_detector.detect(positions, radii)
                    if is_collision:
                        if self.pbm_coupling is not None:
                            self.pbm_coupling.on_collision(collided_pairs, radii)
                        self.solution.is_collision = True
                        self.solution.collided_droplets = collided_pairs
                        self.solution.compact()
                        new_solution = self.solution.generate_next_solution()
                        self.solution = new_solution
                        self.post_processor.update_solution(self.solution)
                        if self.solution.real_time_visualization:
                            self.post_processor.update_live_plot()
                        break

                if hasattr(self.force_calculator, "calculate_forces_and_total_velocity"):
                    forces, total_velocities = (
                        self.force_calculator.calculate_forces_and_total_velocity(
                            positions, radii, stokes_factor=self.stokes_factor
                        )
                    )
                    positions += total_velocities * dt
                else:
                    forces, convection_velocities = (
                        self.force_calculator.calculate_forces_and_convection(positions, radii)
                    )
                    migration_velocities = forces / (self.stokes_factor * radii[:, np.newaxis])
                    positions += (migration_velocities + convection_velocities) * dt

                L = self.force_calculator.L
                boundary_mode = self.force_calculator.boundary_mode
                if boundary_mode == "periodic":
                    positions %= L

                if self.pbm_coupling is not None:
                    self.pbm_coupling.step(t + dt, radii, dt)

                t += dt
                if self.save_interval <= 1 or step % self.save_interval == 0:
                    self.solution.save_step(t, positions)
                elif step == remaining_steps - 1:
                    self.solution.save_step(t, positions)

                if self.solution.should_update_visualization():
                    self.post_processor.live_plot()

        self.solution.compact_all()
        self.simulation_time = time.time() - time_start
        print(f"Total compute time: {self.simulation_time:.2f} s")
        if self.solution.real_time_visualization:
            self.post_processor.finalize_live_plot()
