"""Explicit Euler time loop over the drop-in operators.

``EulerDropletSolver`` is the reference's solver
(dem/solver/euler_droplet_solver.py:8-159) for callers that hand it the
operators and a ``DropletSolution``: the time loop below owns the control
flow, and a *state backend* — a handful of methods — owns where the droplets
live.  This class's backend keeps them in the segment's own numpy arrays, as
the reference does; ``DeviceEulerDropletSolver`` (dem/solver/device_solver.py)
runs the same loop with the state in device memory.

What the loop guarantees (the reference's observable semantics):
* t starts at the solution's current time and runs for ``total_time``;
  each pass over a segment takes int((t_end - t) / dt) + 1 steps unless a
  contact or a stop request ends it;
* a step first asks the collision detector (when there is one); on contact
  the pairs are recorded on the segment, the PBM hook (if any) is told, the
  droplets merge into a new segment at the same t and the pass restarts
  there *without advancing t*;
* otherwise the operator gives forces and velocities — ``forces_and_total_
  velocity`` when the operator has it (tree: x += v dt), else ``forces_and_
  convection`` (direct: x += (F / (sf r) + V) dt) — and a periodic box wraps
  positions with ``%= L``;
* rows are saved every ``save_interval``-th step of a pass and on its last
  step; the backend's positions are those of the segment's
  ``initial_droplet_state``, advanced in place.
stokes_factor = 2 pi eta_o (2 eta_o + 3 eta_w) / (eta_o + eta_w) (:37-41).
"""

from __future__ import annotations

import time

import numpy as np

from .solver_base import Solver


def stokes_factor(eta_oil: float, eta_water: float) -> float:
    """Hadamard-Rybczynski drag factor: migration velocity = F / (factor r)."""
    return 2 * np.pi * eta_oil * (2 * eta_oil + 3 * eta_water) / (eta_oil + eta_water)


class _Progress:
    """Prints one line whenever the integer percentage of simulated time moves."""

    def __init__(self, t0: float, t_end: float, enabled: bool):
        self.t0, self.t_end, self.enabled = t0, t_end, enabled
        self.span = max(t_end - t0, 1e-12)
        self.shown = None
        self.wall0 = time.time()

    def update(self, t: float, droplets: int) -> None:
        pct = int(100 * (t - self.t0) / self.span)
        if self.enabled and pct != self.shown:
            self.shown = pct
            print(f"{pct}%: time {t:.3f} of {self.t_end:.3f}, droplets: {droplets}, "
                  f"elapsed: {time.time() - self.wall0:.1f} s", flush=True)

    def elapsed(self) -> float:
        return time.time() - self.wall0


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
        self.save_interval = max(save_interval, 1)
        self.simulation_time = 0
        self.stokes_factor = stokes_factor(force_calculator.eta_oil, force_calculator.eta_water)
        self.verbose = True
        self.steps_done = 0
        self.merges = 0

    # ------------------------------------------------ host state backend

    def _open_segment(self) -> None:
        """Bind the state of the current segment (its initial state's arrays)."""
        st = self.solution.initial_droplet_state
        self._x, self._r = st.positions, st.radii

    def _droplet_radii(self):
        """Host radii of the current segment (for the PBM hooks)."""
        return self._r

    def _contact_pairs(self):
        """The detector's pairs, or None without contact."""
        hit, pairs = self.collision_detector.detect(self._x, self._r)
        return pairs if hit else None

    def _merge(self, pairs) -> None:
        """Record the contact on the segment and open the merged one."""
        sol = self.solution
        sol.is_collision = True
        sol.collided_droplets = pairs
        sol.compact()
        self.solution = sol.generate_next_solution()

    def _advance(self, dt: float) -> None:
        fc = self.force_calculator
        if hasattr(fc, "calculate_forces_and_total_velocity"):
            _, v = fc.calculate_forces_and_total_velocity(self._x, self._r,
                                                          stokes_factor=self.stokes_factor)
            self._x += v * dt
        else:
            F, V = fc.calculate_forces_and_convection(self._x, self._r)
            self._x += (F / (self.stokes_factor * self._r[:, np.newaxis]) + V) * dt
        if fc.boundary_mode == "periodic":
            self._x %= fc.L

    def _positions_to_save(self):
        return self._x

    def _save(self, t: float) -> None:
        """A saved step: the positions' row in the solution."""
        self.solution.save_step(t, self._positions_to_save())

    def _close(self) -> None:
        """End of solve(): nothing is held outside the solution."""

    # ------------------------------------------------ the time loop

    def _stop_requested(self) -> bool:
        if self.post_processor.stop_simulation:
            print("Simulation stopped by the user.")
            return True
        return False

    def _pass(self, t: float, t_end: float, dt: float, progress: _Progress):
        """Steps over the current segment.  Returns (t, why) with why one of
        'done' (the pass's steps are taken), 'merged', 'stopped'."""
        n_steps = int((t_end - t) / dt) + 1
        for step in range(n_steps):
            progress.update(t, self.solution.num_particles)
            if self._stop_requested():
                return t, "stopped"
            if self.collision_detector is not None:
                pairs = self._contact_pairs()
                if pairs is not None:
                    if self.pbm_coupling is not None:
                        self.pbm_coupling.on_collision(pairs, self._droplet_radii())
                    self._merge(pairs)
                    self.merges += 1
                    self.post_processor.update_solution(self.solution)
                    if self.solution.real_time_visualization:
                        self.post_processor.update_live_plot()
                    return t, "merged"
            self._advance(dt)
            if self.pbm_coupling is not None:
                self.pbm_coupling.step(t + dt, self._droplet_radii(), dt)
            t += dt
            self.steps_done += 1
            if self.save_interval <= 1 or step % self.save_interval == 0 or step == n_steps - 1:
                self._save(t)
            if self.solution.should_update_visualization():
                self.post_processor.live_plot()
        return t, "done"

    def solve(self, dt, total_time):
        t = float(self.solution.get_current_time())
        t_end = t + float(total_time)
        progress = _Progress(t, t_end, self.verbose)
        while t < t_end:
            if self._stop_requested():
                break
            self._open_segment()
            t, why = self._pass(t, t_end, dt, progress)
            if why == "stopped":
                break
        self._close()
        self.solution.compact_all()
        self.simulation_time = progress.elapsed()
        if self.verbose:
            print(f"Total compute time: {self.simulation_time:.2f} s")
        if self.solution.real_time_visualization:
            self.post_processor.finalize_live_plot()
