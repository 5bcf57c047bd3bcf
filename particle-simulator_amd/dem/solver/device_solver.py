"""Euler time loop whose state stays on the GPU.

``DeviceEulerDropletSolver`` runs ``EulerDropletSolver``'s time loop (the
reference's control flow, dem/solver/euler_droplet_solver.py:43-159) with a
device state backend: positions and radii live in device memory between
steps.

* contact detection: ``dem_hash_pairs`` on the device arrays; only the pair
  count crosses to the host each step;
* merge: the reference merges from the last *saved* positions
  (``trajectories[current_step]``).  When those equal the device state
  (always with save_interval 1), the greedy plan runs on the host over the
  pairs only (``merge_plan``, the reference's float64 arithmetic) and
  ``dem_merge_compact`` compacts the state on the device (survivors in order,
  merged droplets appended), which stays there; otherwise the state is merged
  on the host by ``DropletSolution.generate_next_solution`` and uploaded.
  Either way the result is ``merge_collided``'s, bit for bit, and the host
  copy of the merged state starts the next solution segment;
* forces + velocities: the operator's ``*_device`` entry point;
* Euler update and wrap: ``dem_euler_step`` / ``dem_euler_step_total``,
  bit-exact with the reference's numpy update;
* ``solution.save_step`` receives a host copy only on saved steps.  The
  copy is queued at the save and the row is written after the next step's
  kernels are launched (or before anything reads the solution: a merge, a
  new segment, the end of ``solve``), so the host copy runs while the GPU
  computes; with live plots on, rows are written at once.  The segment's
  ``initial_droplet_state.positions``, which the host backend advances in
  place, receives the state when the segment ends.

Given identical operator outputs, trajectories are bit-identical to
``EulerDropletSolver``'s.
"""

from __future__ import annotations

import numpy as np

from .. import _device, _native
from ..solution.droplet_solution import merge_plan
from .euler_droplet_solver import EulerDropletSolver


class DeviceEulerDropletSolver(EulerDropletSolver):
    def __init__(self, force_calculator, solution, post_processor, collision_detector=None,
                 save_interval: int = 1, pbm_coupling=None, verbose: bool = True):
        super().__init__(force_calculator, solution, post_processor, collision_detector,
                         save_interval, pbm_coupling)
        self.verbose = verbose
        self._pin = None      # pinned staging for device -> host copies
        self._ring = None     # pinned slot of a queued save
        self._pending = None  # (t, slot view, event, solution) of a row not yet written
        self._carried = None  # device state of a segment opened by a device merge
        self._pos = self._rad = None
        self._saved_is_state = False  # the last saved row equals the device state

    # ------------------------------------------------ device state backend

    def _open_segment(self) -> None:
        self._finish_save()
        sol = self.solution
        if self._carried is not None:
            self._pos, self._rad = self._carried
            self._carried = None
            self._saved_is_state = True  # the new segment saved its initial state
        else:
            st = sol.initial_droplet_state
            self._pos = _device.to_device(np.asarray(st.positions, dtype=np.float64).reshape(-1, 3))
            self._rad = _device.to_device(np.asarray(st.radii, dtype=np.float64))
            self._saved_is_state = np.array_equal(sol.trajectories[sol.current_step],
                                                  st.positions)

    def _droplet_radii(self):
        return self.solution.initial_droplet_state.radii

    def _contact_pairs(self):
        det = self.collision_detector
        k, pairs_d = det.detect_device(self._pos, self._rad, det._L,
                                       det._boundary_mode == "periodic")
        return pairs_d.cpu().numpy() if k > 0 else None

    def _merge(self, pairs) -> None:
        self._finish_save()
        sol = self.solution
        sol.is_collision = True
        sol.collided_droplets = pairs
        sol.compact()
        saved = sol.trajectories[sol.current_step]  # what the reference merges from
        state = saved if self._saved_is_state else self._staged(self._pos)
        self._hand_back(state)
        if self._saved_is_state or (saved.shape == state.shape and np.array_equal(saved, state)):
            self.solution, self._carried = self._merge_on_device(sol, pairs, saved)
        else:
            self.solution = sol.generate_next_solution()

    def _advance(self, dt: float) -> None:
        fc = self.force_calculator
        n = int(self._pos.shape[0])
        periodic = 1 if fc.boundary_mode == "periodic" else 0
        if hasattr(fc, "calculate_forces_and_total_velocity"):
            _, V = fc.forces_and_total_velocity_device(self._pos, self._rad, self.stokes_factor)
            _native.call("dem_euler_step_total", _device.ptr(self._pos), _device.ptr(V), n,
                         float(dt), float(fc.L), periodic, _device.stream())
        else:
            F, V = fc.forces_and_convection_device(self._pos, self._rad)
            _native.call("dem_euler_step", _device.ptr(self._pos), _device.ptr(F), _device.ptr(V),
                         _device.ptr(self._rad), n, float(self.stokes_factor), float(dt),
                         float(fc.L), periodic, _device.stream())
        self._saved_is_state = False
        self._finish_save()  # the last saved row, written while the step computes

    def _positions_to_save(self):
        self._saved_is_state = True
        return self._staged(self._pos)

    def _save(self, t: float) -> None:
        self._finish_save()
        if self.solution.real_time_visualization:  # live plots read the row at once
            self.solution.save_step(t, self._positions_to_save())
            return
        torch = _device.torch()
        k = self._pos.numel()
        if self._ring is None or self._ring.numel() < k:
            self._ring = torch.empty(max(k, 1 << 16), dtype=torch.float64, pin_memory=True)
        v = self._ring[:k].view(self._pos.shape)
        v.copy_(self._pos, non_blocking=True)  # ordered before the next step's kernels
        ev = torch.cuda.Event()
        ev.record()
        self._pending = (t, v, ev, self.solution)
        self._saved_is_state = True

    def _finish_save(self) -> None:
        """Writes a queued row into the solution it was saved from."""
        if self._pending is None:
            return
        t, v, ev, sol = self._pending
        self._pending = None
        ev.synchronize()
        sol.save_step(t, v.numpy())

    def _close(self) -> None:
        self._finish_save()
        if self._pos is not None and self._carried is None:
            self._hand_back(self._staged(self._pos))

    # ------------------------------------------------ helpers

    def _merge_on_device(self, sol, pairs, saved):
        """``sol.generate_next_solution()`` with the compaction on the device;
        ``saved`` (the saved positions the reference merges from) equals the
        device state.  Returns the next solution and its device state."""
        radii = sol.radii
        removed, new_r, new_p, visit = merge_plan(pairs, lambda i: radii[i], lambda i: saved[i])
        rem = np.asarray(sorted(removed), dtype=np.int32)
        m = len(new_r)
        merged = np.empty((m, 4), dtype=np.float64)
        if m:
            merged[:, :3] = np.asarray(new_p, dtype=np.float64).reshape(m, 3)
            merged[:, 3] = np.asarray(new_r, dtype=np.float64)
        n = int(self._pos.shape[0])
        n_out = n - len(rem) + m
        rem_d = _device.to_device(rem, np.int32) if len(rem) else None
        merged_d = _device.to_device(merged) if m else None
        out_p = _device.empty((n_out, 3))
        out_r = _device.empty((n_out,))
        _native.call("dem_merge_compact", _device.ptr(self._pos), _device.ptr(self._rad), n,
                     _device.ptr(rem_d), len(rem), _device.ptr(merged_d), m, _device.ptr(out_p),
                     _device.ptr(out_r), _device.stream())
        # one host copy of the merged state out of the pinned staging buffer,
        # adopted by the new segment's initial state
        new_r_h = np.array(self._staged(out_r))
        new_p_h = np.array(self._staged(out_p))
        return sol.next_solution(new_p_h, new_r_h, visit, adopt=True), (out_p, out_r)

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

    def _hand_back(self, state):
        """The host loop advances ``initial_droplet_state.positions`` in place;
        the device loop writes the state there when the segment ends."""
        st = self.solution.initial_droplet_state
        if np.shape(st.positions) == state.shape:
            st.positions[...] = state
