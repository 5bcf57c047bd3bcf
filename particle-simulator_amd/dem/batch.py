"""Many independent periodic droplet systems stepped together on one GPU.

The reference runs independent systems one after another: ``main_multi.py``
(:43-71) builds a ``DirectDropletForceCalculator`` and an
``EulerDropletSolver`` per input file and calls ``solve``; each step of
``solve`` (dem/solver/euler_droplet_solver.py:43-159) checks contacts with
the spatial hash, merges on contact (droplet_solution.py:208-287) and
retries, otherwise computes forces + Oseen velocities and takes an Euler step
with the periodic wrap.

``BatchedEnsembles`` keeps exactly that per-system sequence of states but runs
one global step for all systems in five launches captured as a HIP graph
(``dem_batch_*`` in include/dem_hip.h).  A system that finds a contact does not
advance in that global step; the host merges it with the same float64 merge
as ``DropletSolution`` and the next global step re-checks it.  Systems stop
after their own number of Euler steps.

``SequentialEnsembles`` is the drop-in path run system after system (the
single-system device kernels, a host synchronisation for the contact count on
every step): the comparison point for the batched throughput, and — run with
the batched source split — a bit-for-bit check of it.

``autotune`` times the batched step for each work-item shape (targets per
item x sources per item) on a copy of the state and returns the fastest.
"""

from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass

import numpy as np

from . import _device, _native
from .collision_detector import SpatialHashCollisionDetector
from .force_calculator import DirectDropletForceCalculator
from .solution.droplet_solution import merge_plan, merge_plan_rows

BLOCKS = (64, 128, 256)
JLENS = (1024, 2048, 4096, 8192)
PAIR_ONCE_JLENS = (1024, 2048, 4096, 8192)  # 4 ... 32 partner blocks per work item


@dataclass
class Tuning:
    block: int = 256   # targets per work item (workgroup size)
    jlen: int = 4096   # sources per work item
    force_tile: int = 0
    conv_tile: int = 0
    contacts: bool = True  # False: no contact pass (solver without a detector)
    pair_once: bool = False  # True: pair-once sums, jlen / 256 partner blocks per work item

    def native(self) -> _native.DemBatchTuning:
        return _native.DemBatchTuning(self.block, self.jlen, self.force_tile, self.conv_tile,
                                      1 if self.contacts else 0, 1 if self.pair_once else 0)


def split_for(n: int, jlen: int, pair_once: bool = False) -> int:
    """Source split (one-sided) or partner-list chunks (pair-once) the batch
    uses for a system of n droplets."""
    lib = _native.load()
    if pair_once:
        return int(lib.dem_batch_pair_once_split(int(n), int(jlen)))
    return int(lib.dem_batch_split(int(n), int(jlen)))


def _consts(s):
    eo, ew = s.eta_oil, s.eta_water
    sf = 2 * np.pi * eo * (2 * eo + 3 * ew) / (eo + ew)
    return float(s.m_const), float(s.eta_const), float(sf)


class BatchedEnsembles:
    """Independent periodic systems stepped together.

    ``systems``: objects with ``positions`` (N,3), ``radii`` (N,), ``L``,
    ``m_const``, ``eta_const``, ``eta_oil`` and ``eta_water`` (e.g.
    ``dem.workloads.System``); ``dts``/``steps``: per-system time step and
    Euler step count.

    One global step (``step``): the contact pass of every active system; while
    any system has contacts, merge all of them on the host (float64, the
    reference's greedy merge) and check again; then force + Oseen + Euler for
    every active system.  Each system therefore goes through exactly the
    states of the reference loop, and every active system advances once per
    global step.
    """

    def __init__(self, systems, dts, steps, tuning: Tuning | None = None, capture: bool = True):
        t = _device.torch()
        dev = _device.device()
        self.n_ens = len(systems)
        if self.n_ens == 0:
            raise ValueError("need at least one system")
        self.tuning = tuning or Tuning()
        self.capacity = np.array([s.radii.shape[0] for s in systems], dtype=np.int64)
        self.offset = np.concatenate([[0], np.cumsum(self.capacity)[:-1]]).astype(np.int64)
        self.n = self.capacity.copy()
        self.L = np.array([float(s.L) for s in systems])
        self.dt = np.asarray(dts, dtype=np.float64).reshape(self.n_ens)
        self.steps = np.asarray(steps, dtype=np.int64).reshape(self.n_ens)
        total = int(self.capacity.sum())
        pos = np.concatenate([np.asarray(s.positions, dtype=np.float64).reshape(-1, 3)
                              for s in systems])
        rad = np.concatenate([np.asarray(s.radii, dtype=np.float64) for s in systems])
        self.pos = t.from_numpy(np.ascontiguousarray(pos)).to(dev)
        self.radii = t.from_numpy(np.ascontiguousarray(rad)).to(dev)
        self.F = t.zeros((total, 3), dtype=t.float64, device=dev)
        self.V = t.zeros((total, 3), dtype=t.float64, device=dev)
        self._desc = (_native.DemBatchEnsemble * self.n_ens)()
        for e, s in enumerate(systems):
            m, eta, sf = _consts(s)
            self._desc[e] = _native.DemBatchEnsemble(int(self.offset[e]), int(self.n[e]),
                                                     int(self.capacity[e]), int(self.steps[e]),
                                                     float(self.L[e]), float(self.dt[e]), sf, m,
                                                     eta)
        self._tune = self.tuning.native()
        lib = _native.load()
        nbytes = lib.dem_batch_workspace_bytes(self._desc, self.n_ens, ctypes.byref(self._tune))
        if nbytes == 0:
            raise _native.NativeError("dem_batch_workspace_bytes", -1, _native.last_error())
        self.ws = t.empty(int(nbytes), dtype=t.uint8, device=dev)
        self._h = ctypes.c_void_p()
        _native.call("dem_batch_create", self._desc, self.n_ens, ctypes.byref(self._tune),
                     _device.ptr(self.pos), _device.ptr(self.radii), _device.ptr(self.F),
                     _device.ptr(self.V), _device.ptr(self.ws), self.ws.numel(),
                     _device.stream(), ctypes.byref(self._h))
        self.active = self._field("active", t.int32)
        self.stepped = self._field("stepped", t.int32)
        self.pair_count = self._field("pair_count", t.int32)
        self.pair_base = self._field("pair_base", t.int32)
        self.steps_done = self._field("steps_done", t.int32)
        self.incremental_rechecks = self._field("incremental", t.int32)
        self.time = self._field("time", t.float64)
        self.pairs = self._field("pairs", t.int32).view(-1, 2)
        self.pair_values = self._field("pair_values", t.float64).view(-1, 8)
        self.pair_capacity = int(self.pairs.shape[0])
        # pinned staging of the check results, copied before the advance
        k = min(self.pair_capacity, 1 << 16)
        self._h_counts = t.empty(self.n_ens, dtype=t.int32, pin_memory=True)
        self._h_pairs = t.empty((k, 2), dtype=t.int32, pin_memory=True)
        self._h_vals = t.empty((k, 8), dtype=t.float64, pin_memory=True)
        self._staged = t.cuda.Event()
        # pairs staged per round; grows to twice the largest total seen
        self._stage_k = min(k, 2048)
        # every merge below is the greedy plan of a check's complete pair list
        self.incremental = True
        _native.call("dem_batch_set_incremental", self._h, 1)
        self.captured = False
        if capture:
            _native.call("dem_batch_capture", self._h)
            self.captured = True
        self._detectors: dict[int, SpatialHashCollisionDetector] = {}
        self.merges = 0
        self.merge_rounds = 0
        self.global_steps = 0
        self.phase_ms: dict | None = None
        # merge rounds on a second stream beside the advance of the
        # contact-free systems (default priority, like every stream here)
        self.overlap = True      # False: check, advance and merge rounds one after another
        self.merge_first = False  # merge rounds first, then one advance of every system
        self._side = t.cuda.Stream()

    def _field(self, name, dtype):
        p = ctypes.c_void_p()
        nb = ctypes.c_size_t()
        _native.call("dem_batch_field", self._h, name.encode(), ctypes.byref(p), ctypes.byref(nb))
        off = int(p.value) - self.ws.data_ptr()
        return self.ws[off: off + nb.value].view(dtype)

    # ------------------------------------------------------------------ stepping

    def check(self, new_step: bool = True):
        _native.call("dem_batch_check", self._h, 1 if new_step else 0, _device.stream())

    def advance(self):
        _native.call("dem_batch_advance", self._h, _device.stream())

    def _stage(self):
        """Queue the copies of the last check's results to pinned memory: the
        counts and the first ``self._stage_k`` pairs with their values (a
        larger total is read from the device by ``_merge_staged``)."""
        k = self._stage_k
        self._h_counts.copy_(self.pair_count, non_blocking=True)
        self._h_pairs[:k].copy_(self.pairs[:k], non_blocking=True)
        self._h_vals[:k].copy_(self.pair_values[:k], non_blocking=True)
        self._staged.record()

    def _tick(self, name, t0):
        if self.phase_ms is not None:
            _device.torch().cuda.synchronize()
            now = time.perf_counter()
            self.phase_ms[name] = self.phase_ms.get(name, 0.0) + (now - t0) * 1e3
            return now
        return t0

    def step(self, merge: bool = True) -> int:
        """One global step; returns the number of system merges in it.

        With ``self.merge_first``: check -> merge the systems in contact,
        check them again, ... -> one advance of every system.  The systems
        are independent, so each goes through the same check / merge /
        advance sequence as in the other schedules and ends in the same
        state; the advance is one launch over all systems.  Otherwise (the
        default): check -> (advance the systems without contacts | merge the others,
        check them again, ... on a second stream) -> advance the rest.  The
        merge rounds touch only the systems in contact, whose descriptors,
        cells and records ``dem_batch_merge_async`` rewrites, and start once
        the advance has fixed which systems it moves (``dem_batch_wait_go``);
        the work tables are rebuilt before the second advance.  With
        ``self.phase_ms`` set to a dict (or ``self.overlap`` false) the
        phases run one after another, synchronised, and their wall times (ms)
        are accumulated under "check", "advance", "merge"."""
        if self.phase_ms is not None or not self.overlap or not merge:
            return self._step_serial(merge)
        if self.merge_first:
            return self._step_merge_first()
        t = _device.torch()
        main = t.cuda.current_stream()
        self.check(new_step=True)
        self._stage()
        self.advance()
        _native.call("dem_batch_wait_go", self._h, ctypes.c_void_p(self._side.cuda_stream))
        merged = 0
        with t.cuda.stream(self._side):
            while True:
                self._staged.synchronize()
                k = self._merge_staged(defer_tables=True)
                if k == 0:
                    break
                merged += k
                self.check(new_step=False)
                self._stage()
        main.wait_stream(self._side)
        if merged:
            self.advance()
        self.global_steps += 1
        return merged

    def _step_merge_first(self) -> int:
        self.check(new_step=True)
        self._stage()
        merged = 0
        while True:
            self._staged.synchronize()
            k = self._merge_staged(defer_tables=True)
            if k == 0:
                break
            merged += k
            self.check(new_step=False)
            self._stage()
        self.advance()  # the work tables are rebuilt here when merges changed them
        self.global_steps += 1
        return merged

    def _step_serial(self, merge: bool = True) -> int:
        t0 = time.perf_counter()
        self.check(new_step=True)
        if merge:
            self._stage()
        t0 = self._tick("check", t0)
        self.advance()
        t0 = self._tick("advance", t0)
        merged = 0
        while merge:
            self._staged.synchronize()
            k = self._merge_staged()
            t0 = self._tick("merge", t0)
            if k == 0:
                break
            merged += k
            self.check(new_step=False)
            self._stage()
            t0 = self._tick("check", t0)
        if merged:
            self.advance()
            self._tick("advance", t0)
        self.global_steps += 1
        return merged

    def run(self, global_steps: int, merge: bool = True) -> int:
        return sum(self.step(merge) for _ in range(int(global_steps)))

    def run_until_done(self, max_global_steps: int = 1_000_000) -> int:
        k = 0
        while not self.done() and k < max_global_steps:
            self.step()
            k += 1
        return k

    def done(self) -> bool:
        return bool(np.all(self.steps_done.cpu().numpy() >= self.steps))

    def _system_pairs(self, e, a, c, pairs_h, vals_h):
        """(pairs, values) of system e from the packed list, or — when the
        packed buffer overflowed — from the single-system pass."""
        if a + c <= self.pair_capacity:
            return pairs_h[a: a + c], vals_h[a: a + c]
        det = self._detectors.get(e)
        if det is None:
            det = SpatialHashCollisionDetector(num_particles=int(self.capacity[e]),
                                               L=float(self.L[e]), boundary_mode="periodic")
            self._detectors[e] = det
        t = _device.torch()
        off, n = int(self.offset[e]), int(self.n[e])
        k, pd = det.detect_device(self.pos[off: off + n], self.radii[off: off + n],
                                  float(self.L[e]), True)
        pairs = pd.cpu().numpy()
        g = t.from_numpy(off + pairs.reshape(-1).astype(np.int64)).to(self.pos.device)
        v = t.cat([self.pos.index_select(0, g), self.radii.index_select(0, g)[:, None]], 1)
        return pairs, v.cpu().numpy().reshape(-1, 8)

    def _merge_staged(self, defer_tables: bool = False) -> int:
        """Merge every system the last check found in contact; returns how
        many systems were merged.

        The merge decisions and merged droplets come from the reference's
        greedy merge with its float64 arithmetic, on the droplets named in
        the pairs, whose values the check packed next to them: for all
        systems at once by ``dem_merge_plan`` (pairs visited in np.argsort
        order of their first droplet, as ``merge_plan_rows`` does), or, when
        the packed pair buffer overflowed, system by system with
        ``merge_plan_rows``.  ``dem_batch_merge`` then compacts each system on
        the device (survivors in order, merged droplets appended) — exactly
        ``merge_collided``'s result."""
        counts = self._h_counts.numpy().astype(np.int64)
        hit = np.nonzero(counts)[0]
        if len(hit) == 0:
            return 0
        base = np.concatenate([[0], np.cumsum(counts)])
        k = self._stage_k
        total = int(min(base[-1], self.pair_capacity))
        pairs_h = self._h_pairs.numpy()
        vals_h = self._h_vals.numpy()
        if total > k:  # more pairs than this round staged
            pairs_h = self.pairs[:total].cpu().numpy()
            vals_h = self.pair_values[:total].cpu().numpy()
            self._stage_k = min(self._h_pairs.shape[0], 1 << int(2 * total - 1).bit_length())
        vp = ctypes.c_void_p
        if base[-1] <= self.pair_capacity:
            ro, rm, mo, mg = self._plan_native(hit, counts, base, pairs_h, vals_h)
        else:
            ro, rm, mo, mg = self._plan_python(hit, counts, base, pairs_h, vals_h)
        es = np.ascontiguousarray(hit, dtype=np.int32)
        _native.call("dem_batch_merge_async" if defer_tables else "dem_batch_merge", self._h,
                     len(es), es.ctypes.data_as(vp),
                     ro.ctypes.data_as(vp), rm.ctypes.data_as(vp), mo.ctypes.data_as(vp),
                     mg.ctypes.data_as(vp), _device.stream())
        self.n[hit] -= ro[1:] - ro[:-1]
        self.n[hit] += mo[1:] - mo[:-1]
        self.merges += len(hit)
        self.merge_rounds += 1
        return len(hit)

    @staticmethod
    def _plan_native(hit, counts, base, pairs_h, vals_h):
        """Merge lists of every system in contact from one ``dem_merge_plan``
        call.  The systems' pair segments are consecutive in the packed list;
        a system whose first column has no repeated droplet is visited in
        list order (the only order a sort can give), the others in np.argsort
        order, which is what the reference's merge uses."""
        c = counts[hit]
        total = int(base[-1])
        P = np.ascontiguousarray(pairs_h[:total], dtype=np.int32)
        V = np.ascontiguousarray(vals_h[:total], dtype=np.float64)
        seg = np.append(base[hit], total).astype(np.int32)
        sys_of = np.repeat(np.arange(len(hit)), c)
        order = (np.arange(total) - np.repeat(seg[:-1], c)).astype(np.int32)
        first = P[:, 0]
        tie = (sys_of[1:] == sys_of[:-1]) & (first[1:] == first[:-1])
        if tie.any():
            for q in np.unique(sys_of[1:][tie]).tolist():
                a, b = int(seg[q]), int(seg[q + 1])
                order[a:b] = np.argsort(first[a:b])
        ro = np.zeros(len(hit) + 1, np.int32)
        mo = np.zeros(len(hit) + 1, np.int32)
        rm = np.zeros(max(2 * total, 1), np.int32)
        mg = np.zeros(max(4 * total, 1), np.float64)
        vp = ctypes.c_void_p
        _native.call("dem_merge_plan", len(hit), seg.ctypes.data_as(vp), P.ctypes.data_as(vp),
                     V.ctypes.data_as(vp), order.ctypes.data_as(vp), ro.ctypes.data_as(vp),
                     rm.ctypes.data_as(vp), mo.ctypes.data_as(vp), mg.ctypes.data_as(vp))
        return ro, rm, mo, mg

    def _plan_python(self, hit, counts, base, pairs_h, vals_h):
        """System by system with ``merge_plan_rows`` (overflowed pair lists
        come from the single-system pass)."""
        roff, moff, removed, merged = [0], [0], [], []
        cap = self.pair_capacity
        for e in hit.tolist():
            a, c = int(base[e]), int(counts[e])
            if c == 1 and a + 1 <= cap:
                # one pair: merge_plan_rows' arithmetic without its bookkeeping
                (i, j), = pairs_h[a: a + 1].tolist()
                x1, y1, z1, r1, x2, y2, z2, r2 = vals_h[a].tolist()
                c1 = r1**3
                c2 = r2**3
                sm = c1 + c2
                removed.extend((i, j) if i < j else (j, i))
                merged.append(((x1 * c1 + x2 * c2) / sm, (y1 * c1 + y2 * c2) / sm,
                               (z1 * c1 + z2 * c2) / sm, sm ** (1 / 3)))
            else:
                pe, ve = self._system_pairs(e, a, c, pairs_h, vals_h)
                rem, mg = merge_plan_rows(pe, ve)
                removed.extend(rem)
                merged.extend(mg)
            roff.append(len(removed))
            moff.append(len(merged))
        ro = np.asarray(roff, dtype=np.int32)
        mo = np.asarray(moff, dtype=np.int32)
        rm = np.asarray(removed, dtype=np.int32)
        mg = np.ascontiguousarray(np.asarray(merged, dtype=np.float64).reshape(-1))
        return ro, rm, mo, mg

    # ------------------------------------------------------------------ state

    def positions(self, e: int) -> np.ndarray:
        off = int(self.offset[e])
        return self.pos[off: off + int(self.n[e])].cpu().numpy()

    def radii_of(self, e: int) -> np.ndarray:
        off = int(self.offset[e])
        return self.radii[off: off + int(self.n[e])].cpu().numpy()

    def forces(self, e: int) -> np.ndarray:
        off = int(self.offset[e])
        return self.F[off: off + int(self.n[e])].cpu().numpy()

    def velocities(self, e: int) -> np.ndarray:
        off = int(self.offset[e])
        return self.V[off: off + int(self.n[e])].cpu().numpy()

    def contact_pairs(self, e: int) -> np.ndarray:
        """Pairs of system e found by the last check (lexicographic order)."""
        b = int(self.pair_base[e].item())
        c = int(self.pair_count[e].item())
        return self.pairs[b: min(b + c, self.pair_capacity)].cpu().numpy()

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _native.load().dem_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SequentialEnsembles:
    """The single-system device path, system after system.

    Per step of system e: contact pairs (``dem_hash_pairs``; the pair count
    crosses to the host), merge and re-check while there are contacts, then
    ``dem_direct_forces_and_convection`` and ``dem_euler_step``.  A merge is
    the device-resident loop's: the greedy plan on the host from the pair
    list and the pair droplets' values (``merge_plan``), the compaction on
    the device (``dem_merge_compact``).  With
    ``jlen`` set, each force evaluation uses the batched split
    (``split_for(n, jlen, pair_once)``; pair-once mode 2 when ``pair_once``),
    which makes its sums bit-identical to ``BatchedEnsembles`` with the same
    ``jlen`` and ``pair_once``.
    """

    def __init__(self, systems, dts, steps, jlen: int | None = None, contacts: bool = True,
                 pair_once: bool = False):
        self.n_ens = len(systems)
        self.jlen = jlen
        self.pair_once = pair_once
        self.contacts = contacts
        self.dt = np.asarray(dts, dtype=np.float64).reshape(self.n_ens)
        self.steps = np.asarray(steps, dtype=np.int64).reshape(self.n_ens)
        self.steps_done = np.zeros(self.n_ens, dtype=np.int64)
        self.L = [float(s.L) for s in systems]
        self.pos = [_device.to_device(np.asarray(s.positions).reshape(-1, 3)) for s in systems]
        self.radii = [_device.to_device(np.asarray(s.radii)) for s in systems]
        self.sf = []
        self.fc = []
        self.det = []
        for s in systems:
            n = int(s.radii.shape[0])
            self.fc.append(DirectDropletForceCalculator(
                num_particles=n, eps_oil=s.eps_oil, eta_oil=s.eta_oil, eta_water=s.eta_water,
                E=s.E, L=s.L, boundary_mode="periodic"))
            self.det.append(SpatialHashCollisionDetector(num_particles=n, L=s.L,
                                                         boundary_mode="periodic"))
            self.sf.append(_consts(s)[2])
        self.F = [None] * self.n_ens
        self.V = [None] * self.n_ens
        self.merges = 0

    def step(self, e: int):
        if self.steps_done[e] >= self.steps[e]:
            return
        L = self.L[e]
        while self.contacts:
            k, pairs = self.det[e].detect_device(self.pos[e], self.radii[e], L, True)
            if k == 0:
                break
            self.pos[e], self.radii[e] = self._merge(self.pos[e], self.radii[e],
                                                     pairs.cpu().numpy())
            self.merges += 1
        n = int(self.radii[e].shape[0])
        prev = prev_sym = None
        if self.jlen is not None:
            if self.pair_once:
                prev_sym = _native.call("dem_direct_set_symmetric", 2)
            prev = _native.call("dem_direct_set_tuning", split_for(n, self.jlen, self.pair_once), 0)
        try:
            F, V = self.fc[e].forces_and_convection_device(self.pos[e], self.radii[e])
        finally:
            if prev is not None:
                _native.call("dem_direct_set_tuning", prev >> 16, prev & 0xFFFF)
            if prev_sym is not None:
                _native.call("dem_direct_set_symmetric", prev_sym)
        _native.call("dem_euler_step", _device.ptr(self.pos[e]), _device.ptr(F), _device.ptr(V),
                     _device.ptr(self.radii[e]), n, float(self.sf[e]), float(self.dt[e]), L, 1,
                     _device.stream())
        self.F[e], self.V[e] = F, V
        self.steps_done[e] += 1

    @staticmethod
    def _merge(pos_d, r_d, pairs):
        """``merge_collided`` of the device state: plan on the host over the
        droplets named in the pairs (the reference's float64 arithmetic),
        compaction on the device."""
        t = _device.torch()
        ids = np.unique(pairs).astype(np.int32)
        n = int(pos_d.shape[0])
        ids_d = _device.to_device(ids, np.int32)
        vals_d = _device.empty((len(ids), 4))
        _native.call("dem_gather_droplets", _device.ptr(pos_d), _device.ptr(r_d), n,
                     _device.ptr(ids_d), len(ids), _device.ptr(vals_d), _device.stream())
        vals = vals_d.cpu().numpy()
        p_h, r_h = vals[:, :3], vals[:, 3]
        at = {int(i): q for q, i in enumerate(ids.tolist())}
        removed, new_r, new_p, _ = merge_plan(pairs, lambda i: r_h[at[int(i)]],
                                              lambda i: p_h[at[int(i)]])
        rem = np.asarray(sorted(removed), dtype=np.int32)
        m = len(new_r)
        merged = np.empty((m, 4), dtype=np.float64)
        merged[:, :3] = np.asarray(new_p, dtype=np.float64).reshape(m, 3)
        merged[:, 3] = np.asarray(new_r, dtype=np.float64)
        n_out = n - len(rem) + m
        rem_d = _device.to_device(rem, np.int32)
        merged_d = _device.to_device(merged)
        out_p = _device.empty((n_out, 3))
        out_r = _device.empty((n_out,))
        _native.call("dem_merge_compact", _device.ptr(pos_d), _device.ptr(r_d), n,
                     _device.ptr(rem_d), len(rem), _device.ptr(merged_d), m,
                     _device.ptr(out_p), _device.ptr(out_r), _device.stream())
        # the inputs of the launch stay referenced until it has run
        t.cuda.current_stream().synchronize()
        return out_p, out_r

    def step_all(self):
        for e in range(self.n_ens):
            self.step(e)

    def run_until_done(self):
        while np.any(self.steps_done < self.steps):
            self.step_all()

    def positions(self, e: int) -> np.ndarray:
        return self.pos[e].cpu().numpy()

    def radii_of(self, e: int) -> np.ndarray:
        return self.radii[e].cpu().numpy()


def autotune(systems, dts, blocks=BLOCKS, jlens=JLENS, pair_once_jlens=PAIR_ONCE_JLENS,
             warmup: int = 1, reps: int = 3, verbose: bool = False):
    """Fastest work shape for these systems: one-sided (block, jlen) and
    pair-once (jlen) candidates each run ``warmup`` + ``reps`` captured
    force/Oseen/Euler steps of every system (contact pass off, so every system
    advances) on a throw-away copy of the state.  Returns (Tuning with
    contacts on, {candidate: ms per step}); candidates are (block, jlen) and
    ("pair_once", jlen)."""
    t = _device.torch()
    times = {}
    big = 1 << 30
    cands = [((b, j), Tuning(b, j, contacts=False)) for b in blocks for j in jlens]
    cands += [(("pair_once", j), Tuning(256, j, contacts=False, pair_once=True))
              for j in pair_once_jlens]
    for key, tune in cands:
        eng = BatchedEnsembles(systems, dts, [big] * len(systems), tune, capture=True)
        st = _device.stream()
        _native.call("dem_batch_run", eng._h, warmup, st)
        t.cuda.synchronize()
        t0 = time.perf_counter()
        _native.call("dem_batch_run", eng._h, reps, st)
        t.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        times[key] = ms
        if verbose:
            print(f"autotune {key}: {ms:.3f} ms/step", flush=True)
        eng.close()
        del eng
    best = min(times, key=times.get)
    if best[0] == "pair_once":
        return Tuning(256, best[1], pair_once=True), times
    return Tuning(best[0], best[1]), times


__all__ = ["BatchedEnsembles", "SequentialEnsembles", "Tuning", "autotune", "split_for"]
