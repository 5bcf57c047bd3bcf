"""Euler time loop whose state stays on the GPU.

Same control flow and results as ``EulerDropletSolver`` (the reference's
loop, dem/solver/euler_droplet_solver.py:43-159) — contact detection first,
on contact a volume-conserving merge and a restart of the step without
advancing t, otherwise forces/velocities, the Euler update and the periodic
wrap — but positions and radii live in device memory between steps:

* contact detection: ``dem_hash_pairs`` on the device arrays; only the pair
  count crosses to the host each step;
* merge: the reference merges from the last *saved* positions
  (``trajectories[current_step]``).  When those equal the device state (always
  with save_interval 1), the greedy plan runs on the host over the pairs only
  (``merge_plan``, the reference's float64 arithmetic) and
  ``dem_merge_compact`` compacts the state on the device (survivors in order,
  merged droplets appended), which stays there; otherwise the state is merged
  on the host by ``DropletSolution.generate_next_solution`` and uploaded.
  Either way the result is ``merge_collided``'s, bit for bit, and the host
  copy of the merged state starts the next solution segment;
* forces + velocities: the operator's ``*_device`` entry point;
* Euler update and wrap: ``dem_euler_step`` / ``dem_euler_step_total``,
  bit-exact with the reference's numpy update;
* ``solution.save_step`` receives a host copy only on saved steps; the
  solution's ``initial_droplet_state.positions``, which the reference's loop
  advances in place, receives the state at the end of each segment.

Given identical operator outputs, trajectories are bit-identical to
``EulerDropletSolver``'s.
"""

from __future__ import annotations

import time

import numpy as np

from .. import _device, _native
from ..solution.droplet_solution import merge_plan


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
        self._pin = None  # pinned staging for device -> host copies

    def _upload(self):
        st = self.solution.initial_droplet_state
        return (_device.to_device(np.asarray(st.positions, dtype=np.float64).reshape(-1, 3)),
                _device.to_device(np.asarray(st.radii, dtype=np.float64)))

    def _merge_on_device(self, sol, pairs, saved, pos_d, r_d):
        """``sol.generate_next_solution()`` with the compaction on the device;
        ``saved`` (the saved positions the reference merges from) equals the
        device state.  Returns the next solution and its device state."""
        radii = sol.radii
        removed, new_r, new_p, collided = merge_plan(pairs, lambda i: radii[i],
                                                     lambda i: saved[i])
        rem = np.asarray(sorted(removed), dtype=np.int32)
        m = len(new_r)
        merged = np.empty((m, 4), dtype=np.float64)
        if m:
            merged[:, :3] = np.asarray(new_p, dtype=np.float64).reshape(m, 3)
            merged[:, 3] = np.asarray(new_r, dtype=np.float64)
        n = int(pos_d.shape[0])
        n_out = n - len(rem) + m
        rem_d = _device.to_device(rem, np.int32) if len(rem) else None
        merged_d = _device.to_device(merged) if m else None
        out_p = _device.empty((n_out, 3))
        out_r = _device.empty((n_out,))
        _native.call("dem_merge_compact", _device.ptr(pos_d), _device.ptr(r_d), n,
                     _device.ptr(rem_d), len(rem), _device.ptr(merged_d), m, _device.ptr(out_p),
                     _device.ptr(out_r), _device.stream())
        new_p_h = np.array(self._staged(out_p))
        new_r_h = np.array(self._staged(out_r))
        return sol.next_solution(new_p_h, new_r_h, collided), (out_p, out_r)

    def _staged(self, d):
        """Device float64 tensor -> numpy view of a pinned staging buffer (one
        copy and one stream sync); valid until the next call."""
        t = _device.torch()
        k = d.numel()
        if self._pin is None or self._pin.numel() < k:
            self._pin = t.empty(max(k, 1 << 16), dtype=t.float64, pin_memory=True)
        v = self._pin[:k].view(d.shape)
        v.copy_(d, non_blocking=True)
        t.cuda.current_stream().synchronize()
        return v.numpy()

    def _write_back(self, cur):
        """The reference advances ``initial_droplet_state.positions`` in place."""
        st = self.solution.initial_droplet_state
        if np.shape(st.positions) == cur.shape:
            st.positions[...] = cur

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
        carried = None  # device state of the new segment after a device merge
        pos_d = None
        while t < t_end and not is_aborted:
            if self.post_processor.stop_simulation:
                break
            if carried is not None:
                pos_d, r_d = carried
                carried = None
                saved_is_current = True  # the new segment saved its initial state
            else:
                pos_d, r_d = self._upload()
                sol = self.solution
                saved_is_current = np.array_equal(sol.trajectories[sol.current_step],
                                                  sol.initial_droplet_state.positions)
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
                        # positions (trajectories[current_step])
                        pairs = pairs_d.cpu().numpy()
                        if self.pbm_coupling is not None:
                            self.pbm_coupling.on_collision(pairs, radii_host)
                        sol = self.solution
                        sol.is_collision = True
                        sol.collided_droplets = pairs
                        sol.compact()
                        saved = sol.trajectories[sol.current_step]
                        cur = saved if saved_is_current else self._staged(pos_d)
                        self._write_back(cur)
                        if saved_is_current or (saved.shape == cur.shape
                                                and np.array_equal(saved, cur)):
                            self.solution, carried = self._merge_on_device(sol, pairs, saved,
                                                                           pos_d, r_d)
                        else:
                            self.solution = sol.generate_next_solution()
                        self.post_processor.update_solution(self.solution)
                        if self.solution.real_time_visualization:
                            self.post_processor.update_live_plot()
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
                saved_is_current = (self.save_interval <= 1 or step % self.save_interval == 0
                                    or step == remaining - 1)
                if saved_is_current:
                    self.solution.save_step(t, self._staged(pos_d))
                if self.solution.should_update_visualization():
                    self.post_processor.live_plot()
        if pos_d is not None and carried is None:
            self._write_back(self._staged(pos_d))
        self.solution.compact_all()
        self.simulation_time = time.time() - time_start
        if self.verbose:
            print(f"Total compute time: {self.simulation_time:.2f} s")
        if self.solution.real_time_visualization:
            self.post_processor.finalize_live_plot()
