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
from .solution.droplet_solution import merge_collided

BLOCKS = (64, 128, 256)
JLENS = (1024, 2048, 4096, 8192)


@dataclass
class Tuning:
    block: int = 256   # targets per work item (workgroup size)
    jlen: int = 4096   # sources per work item
    force_tile: int = 0
    conv_tile: int = 0

    def native(self) -> _native.DemBatchTuning:
        return _native.DemBatchTuning(self.block, self.jlen, self.force_tile, self.conv_tile)


def split_for(n: int, jlen: int) -> int:
    """Source split the batch uses for a system of n droplets."""
    return int(_native.load().dem_batch_split(int(n), int(jlen)))


def _consts(s):
    eo, ew = s.eta_oil, s.eta_water
    sf = 2 * np.pi * eo * (2 * eo + 3 * ew) / (eo + ew)
    return float(s.m_const), float(s.eta_const), float(sf)


class BatchedEnsembles:
    """Independent periodic systems stepped by one captured graph per global step.

    ``systems``: objects with ``positions`` (N,3), ``radii`` (N,), ``L``,
    ``m_const``, ``eta_const``, ``eta_oil`` and ``eta_water`` (e.g.
    ``dem.workloads.System``); ``dts``/``steps``: per-system time step and
    Euler step count.
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
        self.contact = self._field("contact", t.int32)
        self.advanced = self._field("advanced", t.int32)
        self.steps_done = self._field("steps_done", t.int32)
        self.time = self._field("time", t.float64)
        self.captured = False
        if capture:
            _native.call("dem_batch_capture", self._h)
            self.captured = True
        self._detectors: dict[int, SpatialHashCollisionDetector] = {}
        self.merges = 0
        self.global_steps = 0

    def _field(self, name, dtype):
        p = ctypes.c_void_p()
        nb = ctypes.c_size_t()
        _native.call("dem_batch_field", self._h, name.encode(), ctypes.byref(p), ctypes.byref(nb))
        off = int(p.value) - self.ws.data_ptr()
        return self.ws[off: off + nb.value].view(dtype)

    # ------------------------------------------------------------------ stepping

    def run(self, global_steps: int, merge: bool = True) -> int:
        """``global_steps`` global steps, merging after each one; returns the
        number of systems merged."""
        merged = 0
        st = _device.stream()
        for _ in range(int(global_steps)):
            _native.call("dem_batch_run", self._h, 1, st)
            self.global_steps += 1
            if merge:
                hits = np.nonzero(self.contact.cpu().numpy())[0]
                for e in hits:
                    self._merge(int(e))
                    merged += 1
        return merged

    def run_until_done(self, max_global_steps: int = 1_000_000) -> int:
        k = 0
        while not self.done() and k < max_global_steps:
            self.run(1)
            k += 1
        return k

    def done(self) -> bool:
        return bool(np.all(self.steps_done.cpu().numpy() >= self.steps))

    def _merge(self, e: int):
        """Merge the contacts of system e on the host (float64, reference order)."""
        off, n = int(self.offset[e]), int(self.n[e])
        det = self._detectors.get(e)
        if det is None:
            det = SpatialHashCollisionDetector(num_particles=int(self.capacity[e]),
                                               L=float(self.L[e]), boundary_mode="periodic")
            self._detectors[e] = det
        pos_e = self.pos[off: off + n]
        r_e = self.radii[off: off + n]
        k, pairs = det.detect_device(pos_e, r_e, float(self.L[e]), True)
        if k == 0:
            raise RuntimeError(f"system {e}: batched contact flag without a contact pair")
        new_p, new_r, _ = merge_collided(pos_e.cpu().numpy(), r_e.cpu().numpy(),
                                         pairs.cpu().numpy())
        m = int(new_r.shape[0])
        t = _device.torch()
        self.pos[off: off + m] = t.from_numpy(np.ascontiguousarray(new_p)).to(self.pos.device)
        self.radii[off: off + m] = t.from_numpy(np.ascontiguousarray(new_r)).to(self.pos.device)
        self.n[e] = m
        _native.call("dem_batch_update", self._h, e, m, _device.stream())
        self.merges += 1

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
    ``dem_direct_forces_and_convection`` and ``dem_euler_step``.  With
    ``jlen`` set, each force evaluation uses the batched source split
    (``split_for(n, jlen)``), which makes its sums bit-identical to
    ``BatchedEnsembles`` with the same ``jlen``.
    """

    def __init__(self, systems, dts, steps, jlen: int | None = None):
        self.n_ens = len(systems)
        self.jlen = jlen
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
        while True:
            k, pairs = self.det[e].detect_device(self.pos[e], self.radii[e], L, True)
            if k == 0:
                break
            new_p, new_r, _ = merge_collided(self.pos[e].cpu().numpy(),
                                             self.radii[e].cpu().numpy(), pairs.cpu().numpy())
            self.pos[e] = _device.to_device(new_p)
            self.radii[e] = _device.to_device(new_r)
            self.merges += 1
        n = int(self.radii[e].shape[0])
        prev = None
        if self.jlen is not None:
            prev = _native.call("dem_direct_set_tuning", split_for(n, self.jlen), 0)
        try:
            F, V = self.fc[e].forces_and_convection_device(self.pos[e], self.radii[e])
        finally:
            if prev is not None:
                _native.call("dem_direct_set_tuning", prev >> 16, prev & 0xFFFF)
        _native.call("dem_euler_step", _device.ptr(self.pos[e]), _device.ptr(F), _device.ptr(V),
                     _device.ptr(self.radii[e]), n, float(self.sf[e]), float(self.dt[e]), L, 1,
                     _device.stream())
        self.F[e], self.V[e] = F, V
        self.steps_done[e] += 1

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


def autotune(systems, dts, blocks=BLOCKS, jlens=JLENS, warmup: int = 1, reps: int = 3,
             verbose: bool = False):
    """Fastest (block, jlen) for these systems: each candidate runs ``warmup``
    + ``reps`` captured global steps on a throw-away copy of the state (no
    merges).  Returns (Tuning, {(block, jlen): ms per global step})."""
    t = _device.torch()
    times = {}
    big = 1 << 30
    for b in blocks:
        for j in jlens:
            eng = BatchedEnsembles(systems, dts, [big] * len(systems), Tuning(b, j), capture=True)
            st = _device.stream()
            _native.call("dem_batch_run", eng._h, warmup, st)
            t.cuda.synchronize()
            t0 = time.perf_counter()
            _native.call("dem_batch_run", eng._h, reps, st)
            t.cuda.synchronize()
            ms = (time.perf_counter() - t0) * 1e3 / reps
            times[(b, j)] = ms
            if verbose:
                print(f"autotune block={b} jlen={j}: {ms:.3f} ms/step", flush=True)
            eng.close()
            del eng
    best = min(times, key=times.get)
    return Tuning(best[0], best[1]), times


__all__ = ["BatchedEnsembles", "SequentialEnsembles", "Tuning", "autotune", "split_for"]
