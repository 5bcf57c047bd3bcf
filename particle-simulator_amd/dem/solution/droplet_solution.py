"""Trajectory chain and the volume-conserving merge (host, float64 numpy).

Drop-in for the reference's ``DropletSolution``
(dem/solution/droplet_solution.py:10-499).  The merge
(``generate_next_solution``, :208-287) is the reference's algorithm kept
bit-exact on purpose — it runs on the host in float64 with the same numpy
calls, in the same order:
* pairs ordered by their first index with numpy's default ``argsort``;
* greedy: a pair is skipped when either droplet already merged;
* r_new = (r1^3 + r2^3) ** (1/3) (float power, not cbrt);
* R^3-weighted centre from raw coordinates (no minimum image, no wrap);
* ``np.delete`` of the merged indices keeps survivors in order and the merged
  droplets are appended last in processing order.
Storage (growing buffers, chain links, npz chain I/O as a list of dicts,
time interpolation) follows the same layout so files interoperate.
"""

from __future__ import annotations

import os.path
import time

import numpy as np

from ..particle_state import DropletState
from .solution_base import Solution


def merge_plan(collided, radius_of, position_of):
    """The greedy merge decisions of the reference (droplet_solution.py:
    208-262), reading droplet values through ``radius_of(i)`` (float64
    scalar) and ``position_of(i)`` ((3,) float64).

    Pairs are visited in ``np.argsort(collided[:, 0])`` order; a pair is
    skipped when either droplet already merged; r = (r1^3 + r2^3) ** (1/3)
    (float power); the centre is the r^3-weighted mean of the raw coordinates
    (no minimum image).  Returns (indices to remove, descending; merged radii;
    merged positions; collided sorted by first index).
    """
    sorted_indices = np.argsort(collided[:, 0])
    collided = collided[sorted_indices]
    new_droplets_radii = []
    new_droplets_positions = []
    collided_indices = set()
    for idx1, idx2 in collided:
        if idx1 in collided_indices or idx2 in collided_indices:
            continue
        radius1, radius2 = radius_of(idx1), radius_of(idx2)
        position1, position2 = position_of(idx1), position_of(idx2)
        radius1_cube = radius1**3
        radius2_cube = radius2**3
        new_radius = (radius1_cube + radius2_cube) ** (1 / 3)
        new_position = (position1 * radius1_cube + position2 * radius2_cube) / (
            radius1_cube + radius2_cube
        )
        new_droplets_radii.append(new_radius)
        new_droplets_positions.append(new_position)
        collided_indices.add(idx1)
        collided_indices.add(idx2)
    return sorted(collided_indices, reverse=True), new_droplets_radii, new_droplets_positions, collided


def merge_plan_rows(pairs, values):
    """``merge_plan`` for one system whose pair list carries the droplets'
    values: pairs (k, 2) local indices, values (k, 8) = (x, y, z, r) of the
    first droplet then of the second.

    Same decisions and the same IEEE operations in the same order as
    ``merge_plan`` — argsort by first index, greedy skip, r^3 = r ** 3,
    r = (r1^3 + r2^3) ** (1/3), centre (x1 r1^3 + x2 r2^3) / (r1^3 + r2^3) per
    axis — on Python floats (C ``pow`` and unfused multiply/add/divide, as
    numpy's float64 scalars and element-wise array operations do), which is
    what makes it cheap for thousands of systems.  Returns (indices to remove,
    ascending; merged (x, y, z, r) tuples in processing order).
    """
    order = np.argsort(pairs[:, 0])
    pl = pairs[order].tolist()
    vl = values[order].tolist()
    value = {}
    for (i, j), v in zip(pl, vl):
        value[i] = v[0:4]
        value[j] = v[4:8]
    done = set()
    merged = []
    for i, j in pl:
        if i in done or j in done:
            continue
        x1, y1, z1, r1 = value[i]
        x2, y2, z2, r2 = value[j]
        c1 = r1**3
        c2 = r2**3
        s = c1 + c2
        merged.append(((x1 * c1 + x2 * c2) / s, (y1 * c1 + y2 * c2) / s,
                       (z1 * c1 + z2 * c2) / s, s ** (1 / 3)))
        done.add(i)
        done.add(j)
    return sorted(done), merged


def merge_collided(positions, radii, collided):
    """Volume-conserving greedy merge (reference droplet_solution.py:208-271).

    Returns (new_positions, new_radii, collided sorted by first index):
    survivors keep their order (``np.delete``) and merged droplets are
    appended in processing order (see ``merge_plan``).
    """
    new_radii = np.copy(radii)
    new_positions = np.copy(positions)
    indices_to_remove, new_droplets_radii, new_droplets_positions, collided = merge_plan(
        collided, lambda i: new_radii[i], lambda i: new_positions[i])
    new_radii = np.delete(new_radii, indices_to_remove)
    new_positions = np.delete(new_positions, indices_to_remove, axis=0)
    new_radii = np.append(new_radii, new_droplets_radii)
    new_positions = np.vstack([new_positions, new_droplets_positions])
    return new_positions, new_radii, collided


class DropletSolution(Solution):
    def __init__(
        self,
        initial_droplet_state,
        real_time_visualization=False,
        update_interval=0.05,
        length=10,
        previous=None,
        filename=None,
    ):
        self.initial_droplet_state = initial_droplet_state
        self.num_particles = len(initial_droplet_state.radii)
        self.radii = initial_droplet_state.radii
        self.length = length
        self.real_time_visualization = real_time_visualization
        self.update_interval = update_interval
        self.times = np.zeros((self.length, 1)) + initial_droplet_state.time
        self.trajectories = np.zeros((self.length, self.num_particles, 3))
        self.last_update_time = time.time()
        self._prev = previous
        self._next = None
        self.is_collision = False
        self.collided_droplets = []
        self.resulting_droplet = []
        self.current_step = -1
        self.save_step(initial_droplet_state.time, initial_droplet_state.positions)
        self.filename = filename

    def save_step(self, t, positions):
        self.current_step += 1
        if self.current_step >= self.length:
            new_length = self.length * 2
            new_times = np.empty((new_length, 1), dtype=self.times.dtype)
            new_times[: self.length] = self.times
            self.times = new_times
            new_traj = np.empty((new_length, self.num_particles, 3), dtype=self.trajectories.dtype)
            new_traj[: self.length] = self.trajectories
            self.trajectories = new_traj
            self.length = new_length
        self.times[self.current_step] = t
        self.trajectories[self.current_step] = positions

    def get_times(self):
        return self.times[: self.current_step + 1]

    def get_trajectories(self):
        return self.trajectories[: self.current_step + 1]

    def get_current_time(self):
        return self.times[self.current_step][0]

    def get_last_time(self):
        valid_steps = np.where(np.any(self.times != 0, axis=0))[0]
        return self.times[valid_steps[-1]]

    def get_chain_start_time(self):
        cur = self
        while cur._prev is not None:
            cur = cur._prev
        return cur.times[0][0]

    def get_chain_end_time(self):
        cur = self
        while cur._next is not None:
            cur = cur._next
        return cur.times[-1][0]

    def get_first_solution(self):
        cur = self
        while cur._prev is not None:
            cur = cur._prev
        return cur

    def get_current_positions(self):
        return self.trajectories[self.current_step]

    def compact(self):
        self.times = self.get_times()
        self.trajectories = self.get_trajectories()

    def compact_all(self):
        self.compact()
        nxt = self._next
        while nxt is not None:
            nxt.compact()
            nxt = nxt._next
        prev = self._prev
        while prev is not None:
            prev.compact()
            prev = prev._prev

    def should_update_visualization(self):
        if self.real_time_visualization:
            now = time.time()
            if now - self.last_update_time >= self.update_interval:
                self.last_update_time = now
                return True
        return False

    def generate_next_solution(self):
        """Merge the collided pairs and start the next trajectory segment."""
        new_positions, new_radii, collided = merge_collided(
            self.trajectories[self.current_step], self.radii, self.collided_droplets)
        return self.next_solution(new_positions, new_radii, collided)

    def next_solution(self, new_positions, new_radii, collided):
        """The second half of ``generate_next_solution``: start the next
        trajectory segment from already merged arrays (``merge_collided``'s
        outputs, or the same values compacted on the device)."""
        self.collided_droplets = collided
        new_initial_state = DropletState(new_positions, new_radii, self.times[self.current_step])
        next_solution = DropletSolution(
            initial_droplet_state=new_initial_state,
            real_time_visualization=self.real_time_visualization,
            update_interval=self.update_interval,
            length=self.length,
            previous=self,
        )
        self._next = next_solution
        next_solution.is_collision = False
        return next_solution

    def save_to_file(self, filename):
        np.savez(
            filename,
            num_particles=self.num_particles,
            radii=self.radii,
            times=self.get_times(),
            trajectories=self.get_trajectories(),
        )
        print(f"Solution saved to {filename}.")

    def load_from_file(self, filename):
        data = np.load(filename)
        self.num_particles = data["num_particles"]
        self.radii = data["radii"]
        self.times = data["times"]
        self.trajectories = data["trajectories"]
        self.length = len(self.times)
        self.filename = os.path.splitext(os.path.basename(filename))[0]
        print(f"Solution loaded from {filename}.")

    def save_chain_to_file(self, filename, precision="float64"):
        if precision == "float32":
            dtype = np.float32
        elif precision == "float16":
            dtype = np.float16
        else:
            dtype = np.float64
        first = self.get_first_solution()
        solution_list = []
        cur = first
        while cur is not None:
            cur.compact()
            solution_list.append(
                {
                    "num_particles": cur.num_particles,
                    "radii": cur.radii.astype(dtype),
                    "length": cur.length,
                    "real_time_visualization": cur.real_time_visualization,
                    "update_interval": cur.update_interval,
                    "times": cur.get_times().astype(dtype),
                    "trajectories": cur.get_trajectories().astype(dtype),
                    "is_collision": cur.is_collision,
                    "collided_droplets": cur.collided_droplets,
                    "resulting_droplet": cur.resulting_droplet,
                }
            )
            cur = cur._next
        np.savez(filename, solutions=solution_list)
        print(f"Solution chain saved to {filename}.")

    @staticmethod
    def load_chain_from_file(filename):
        """Load a chain written by ``save_chain_to_file`` (a pickled list of
        dicts inside an npz, the reference's format).  Only load files this
        package (or the reference) wrote: the format unpickles."""
        fname = os.path.splitext(os.path.basename(filename))[0]
        data = np.load(filename, allow_pickle=True)
        solution_list = data["solutions"]
        first = None
        prev = None
        for sd in solution_list:
            state = DropletState(sd["trajectories"][0], sd["radii"], sd["times"][0])
            sol = DropletSolution(
                initial_droplet_state=state,
                real_time_visualization=sd["real_time_visualization"],
                update_interval=sd["update_interval"],
                length=sd["length"],
                previous=prev,
                filename=fname,
            )
            sol.times = sd["times"]
            sol.trajectories = sd["trajectories"]
            sol.is_collision = sd["is_collision"]
            sol.collided_droplets = sd["collided_droplets"]
            sol.resulting_droplet = sd["resulting_droplet"]
            if prev is not None:
                prev._next = sol
            else:
                first = sol
            prev = sol
        print(f"Solution chain loaded from {filename}.")
        return first

    def get_state(self, target_time):
        cur = self
        while True:
            min_time = cur.times[0][0]
            max_time = cur.times[-1][0]
            if target_time < min_time:
                if cur._prev is not None:
                    cur = cur._prev
                else:
                    raise ValueError(f"time {target_time} precedes the chain start")
            elif target_time > max_time:
                if cur._next is not None:
                    cur = cur._next
                else:
                    raise ValueError(f"time {target_time} exceeds the chain end")
            else:
                break
        times = cur.times.flatten()
        ib = np.searchsorted(times, target_time) - 1
        ia = ib + 1
        if ib < 0:
            ib = 0
        if ia >= len(times):
            ia = len(times) - 1
        tb, ta = times[ib], times[ia]
        pb, pa = cur.trajectories[ib], cur.trajectories[ia]
        if target_time == ta:
            pos = pa
        elif target_time == tb:
            pos = pb
        else:
            f = (target_time - tb) / (ta - tb)
            pos = pb + f * (pa - pb)
        return DropletState(positions=pos, radii=cur.radii, time=target_time)

    def get_name(self):
        first = self.get_first_solution()
        return first.filename if first.filename is not None else ""
