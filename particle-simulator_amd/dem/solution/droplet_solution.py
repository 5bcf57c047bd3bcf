"""Trajectory segments and the volume-conserving merge (host, float64 numpy).

A run is a doubly linked chain of segments.  A segment records one fixed set
of droplets (radii and count never change inside it) at successive times;
a merge closes the segment and opens the next one with the merged set.  The
public surface — class and method names, attributes read by the solver and
post-processors (``times`` (rows, 1), ``trajectories`` (rows, N, 3),
``radii``, ``num_particles``, ``length``, ``current_step``,
``initial_droplet_state``, ``is_collision``, ``collided_droplets``,
``resulting_droplet``, ``_prev`` / ``_next``) and the npz files — is that of
the reference's ``DropletSolution`` (dem/solution/droplet_solution.py:10-499),
so its callers and files work unchanged.

Recording: rows [0, current_step] of ``times`` / ``trajectories`` are
recorded; the buffers double when full (``length`` doubles with them, and is
what a saved chain stores), and ``compact`` trims them to the recorded rows.
Unrecorded rows of ``times`` hold the segment's start time; unrecorded rows
of ``trajectories`` are not initialised: the reference zero-fills a new
segment's buffer (milliseconds per segment at 65536 droplets) but leaves the
rows a doubling adds uninitialised, and every accessor reads recorded rows
only.

The merge (``generate_next_solution``, reference :208-287) is kept bit-exact
with the reference: pairs visited in numpy's default ``argsort`` order of the
first droplet, a pair skipped when either droplet already merged, the merged
radius (r1^3 + r2^3) ** (1/3) as a float power, the centre the r^3-weighted
mean of the raw (unwrapped) coordinates, survivors kept in order and merged
droplets appended in visiting order.
"""

from __future__ import annotations

import os.path
import time

import numpy as np

from ..particle_state import DropletState
from .solution_base import Solution

# ---------------------------------------------------------------- the merge


def merge_plan(collided, radius_of, position_of):
    """Greedy merge decisions over the contact pairs ``collided`` (k, 2),
    reading droplet values through ``radius_of(i)`` (float64 scalar) and
    ``position_of(i)`` ((3,) float64).

    Returns (indices to remove, descending; merged radii; merged centres;
    the pairs in visiting order).  Reference: droplet_solution.py:214-262.
    """
    visit = collided[np.argsort(collided[:, 0])]  # numpy's default kind, as the reference
    merged_away = set()
    radii_out, centres_out = [], []
    for a, b in visit:
        if a in merged_away or b in merged_away:
            continue
        va = radius_of(a) ** 3
        vb = radius_of(b) ** 3
        radii_out.append((va + vb) ** (1 / 3))
        centres_out.append((position_of(a) * va + position_of(b) * vb) / (va + vb))
        merged_away.update((a, b))
    return sorted(merged_away, reverse=True), radii_out, centres_out, visit


def merge_plan_rows(pairs, values):
    """``merge_plan`` for one system whose pair list carries the droplets'
    values: pairs (k, 2) local indices, values (k, 8) = (x, y, z, r) of the
    first droplet then of the second.

    Same decisions and the same IEEE operations in the same order as
    ``merge_plan`` (r^3 = r ** 3, r = (r1^3 + r2^3) ** (1/3), centre
    (x1 r1^3 + x2 r2^3) / (r1^3 + r2^3) per axis), on Python floats — C
    ``pow`` and unfused multiply/add/divide, as numpy's float64 scalars and
    element-wise operations do — which is what makes it cheap for thousands
    of systems.  Returns (indices to remove, ascending; merged (x, y, z, r)
    tuples in visiting order).
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

    Returns (new_positions, new_radii, collided in visiting order): survivors
    keep their order and merged droplets follow in visiting order.
    """
    gone, add_r, add_x, visit = merge_plan(collided, lambda i: radii[i], lambda i: positions[i])
    keep = np.ones(len(radii), dtype=bool)
    keep[gone] = False
    new_radii = np.concatenate([np.asarray(radii, dtype=np.float64)[keep],
                                np.asarray(add_r, dtype=np.float64)])
    survivors = np.asarray(positions, dtype=np.float64)[keep]
    new_positions = (np.concatenate([survivors, np.asarray(add_x, dtype=np.float64).reshape(-1, 3)])
                     if add_x else survivors)
    return new_positions, new_radii, visit


# ---------------------------------------------------------------- storage helpers


def _enlarged(a: np.ndarray, rows: int) -> np.ndarray:
    """``a`` with its leading axis enlarged to ``rows``; the old rows copied,
    the new ones left uninitialised (they are written before being read)."""
    out = np.empty((rows,) + a.shape[1:], dtype=a.dtype)
    out[: a.shape[0]] = a
    return out


_PRECISIONS = {"float32": np.float32, "float16": np.float16}


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
        state = initial_droplet_state
        self.initial_droplet_state = state
        self.radii = state.radii
        self.num_particles = len(state.radii)
        self.real_time_visualization = real_time_visualization
        self.update_interval = update_interval
        self.filename = filename
        self._prev = previous
        self._next = None
        # collision record of the segment (set when it ends in a merge)
        self.is_collision = False
        self.collided_droplets = []
        self.resulting_droplet = []
        self.last_update_time = time.time()
        # unrecorded rows: the start time in ``times``, uninitialised trajectories
        self.length = length
        self.times = np.zeros((length, 1)) + state.time
        self.trajectories = np.empty((length, self.num_particles, 3))
        self.current_step = -1
        self.save_step(state.time, state.positions)

    # -- recording

    def save_step(self, t, positions):
        row = self.current_step + 1
        if row >= self.length:
            rows = 2 * self.length
            self.times = _enlarged(self.times, rows)
            self.trajectories = _enlarged(self.trajectories, rows)
            self.length = rows
        self.times[row] = t
        self.trajectories[row] = positions
        self.current_step = row

    def _recorded(self) -> slice:
        return slice(0, self.current_step + 1)

    def get_times(self):
        return self.times[self._recorded()]

    def get_trajectories(self):
        return self.trajectories[self._recorded()]

    def get_current_time(self):
        return self.times[self.current_step, 0]

    def get_current_positions(self):
        return self.trajectories[self.current_step]

    def get_last_time(self):
        """The reference evaluates ``times[where(any(times != 0, axis=0))[0][-1]]``
        (:126-127).  The ``any`` reduces over the step axis of the (rows, 1)
        array, so the result is the segment's *first* time row whenever some
        recorded time is non-zero (IndexError when none is).  The
        post-processor's collision windows read this value, so it is kept."""
        if not np.any(self.times != 0):
            raise IndexError("get_last_time: every time of the segment is zero")
        return self.times[0]

    def compact(self):
        self.times = self.get_times()
        self.trajectories = self.get_trajectories()

    # -- the chain

    def _segments(self):
        """Every segment of the chain, first to last."""
        seg = self.get_first_solution()
        while seg is not None:
            yield seg
            seg = seg._next

    def get_first_solution(self):
        seg = self
        while seg._prev is not None:
            seg = seg._prev
        return seg

    def get_chain_start_time(self):
        return self.get_first_solution().times[0][0]

    def get_chain_end_time(self):
        seg = self
        while seg._next is not None:
            seg = seg._next
        return seg.times[-1][0]

    def compact_all(self):
        for seg in self._segments():
            seg.compact()

    def get_name(self):
        name = self.get_first_solution().filename
        return "" if name is None else name

    def should_update_visualization(self):
        if not self.real_time_visualization:
            return False
        now = time.time()
        if now - self.last_update_time < self.update_interval:
            return False
        self.last_update_time = now
        return True

    # -- merging

    def generate_next_solution(self):
        """Merge the collided pairs of the last recorded positions and open the
        next segment (reference :208-287)."""
        new_positions, new_radii, visit = merge_collided(
            self.trajectories[self.current_step], self.radii, self.collided_droplets)
        return self.next_solution(new_positions, new_radii, visit)

    def next_solution(self, new_positions, new_radii, collided, adopt=False):
        """Open the next segment from already merged arrays (``merge_collided``'s
        outputs, or the same values compacted on the device); ``collided`` is
        the pair list in visiting order, kept as this segment's record.  With
        ``adopt`` the new initial state holds the arrays themselves instead of
        copies (arrays that nothing else refers to)."""
        self.collided_droplets = collided
        t0 = self.times[self.current_step]
        start = (DropletState.adopt(new_positions, new_radii, t0) if adopt
                 else DropletState(new_positions, new_radii, t0))
        nxt = DropletSolution(
            initial_droplet_state=start,
            real_time_visualization=self.real_time_visualization,
            update_interval=self.update_interval,
            length=self.length,
            previous=self,
        )
        self._next = nxt
        return nxt

    # -- time lookup

    def _segment_holding(self, t):
        seg = self
        while True:
            if t < seg.times[0][0]:
                if seg._prev is None:
                    raise ValueError(f"time {t} precedes the chain start")
                seg = seg._prev
            elif t > seg.times[-1][0]:
                if seg._next is None:
                    raise ValueError(f"time {t} exceeds the chain end")
                seg = seg._next
            else:
                return seg

    def get_state(self, target_time):
        """Positions at ``target_time``, linearly interpolated between the two
        recorded rows around it (a recorded time returns its row as is)."""
        seg = self._segment_holding(target_time)
        ts = seg.times.ravel()
        k = int(np.searchsorted(ts, target_time))
        lo, hi = max(k - 1, 0), min(k, len(ts) - 1)
        if target_time == ts[hi]:
            pos = seg.trajectories[hi]
        elif target_time == ts[lo]:
            pos = seg.trajectories[lo]
        else:
            w = (target_time - ts[lo]) / (ts[hi] - ts[lo])
            pos = seg.trajectories[lo] + w * (seg.trajectories[hi] - seg.trajectories[lo])
        return DropletState(positions=pos, radii=seg.radii, time=target_time)

    # -- files (npz layouts of the reference, :289-421)

    def save_to_file(self, filename):
        """One segment: keys num_particles, radii, times, trajectories."""
        np.savez(filename, num_particles=self.num_particles, radii=self.radii,
                 times=self.get_times(), trajectories=self.get_trajectories())
        print(f"Solution saved to {filename}.")

    def load_from_file(self, filename):
        data = np.load(filename)
        for key in ("num_particles", "radii", "times", "trajectories"):
            setattr(self, key, data[key])
        self.length = len(self.times)
        self.filename = os.path.splitext(os.path.basename(filename))[0]
        print(f"Solution loaded from {filename}.")

    def _record(self, dtype) -> dict:
        """This segment as one entry of a saved chain (the reference's dict keys)."""
        return {
            "num_particles": self.num_particles,
            "radii": self.radii.astype(dtype),
            "length": self.length,
            "real_time_visualization": self.real_time_visualization,
            "update_interval": self.update_interval,
            "times": self.get_times().astype(dtype),
            "trajectories": self.get_trajectories().astype(dtype),
            "is_collision": self.is_collision,
            "collided_droplets": self.collided_droplets,
            "resulting_droplet": self.resulting_droplet,
        }

    def save_chain_to_file(self, filename, precision="float64"):
        """The whole chain as ``solutions``: an object array of per-segment dicts
        (float64 unless precision is 'float32' or 'float16')."""
        dtype = _PRECISIONS.get(precision, np.float64)
        self.compact_all()
        np.savez(filename, solutions=[seg._record(dtype) for seg in self._segments()])
        print(f"Solution chain saved to {filename}.")

    @staticmethod
    def load_chain_from_file(filename):
        """Load a chain written by ``save_chain_to_file`` (or the reference).
        The format unpickles: load only files this package or the reference
        wrote."""
        name = os.path.splitext(os.path.basename(filename))[0]
        records = np.load(filename, allow_pickle=True)["solutions"]
        first = prev = None
        for rec in records:
            seg = DropletSolution(
                initial_droplet_state=DropletState(rec["trajectories"][0], rec["radii"],
                                                   rec["times"][0]),
                real_time_visualization=rec["real_time_visualization"],
                update_interval=rec["update_interval"],
                length=rec["length"],
                previous=prev,
                filename=name,
            )
            for key in ("times", "trajectories", "is_collision", "collided_droplets",
                        "resulting_droplet"):
                setattr(seg, key, rec[key])
            if prev is None:
                first = seg
            else:
                prev._next = seg
            prev = seg
        print(f"Solution chain loaded from {filename}.")
        return first
