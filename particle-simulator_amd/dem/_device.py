"""Device plumbing: PyTorch ROCm tensors as the buffers handed to libdemhip.

PyTorch only provides memory and the stream here; every computation on the
hot path is a libdemhip kernel.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _native

_checked = False


def torch():
    import torch as _t

    return _t


def device():
    """The current ROCm device; raises if none is visible (no CPU fallback)."""
    global _checked
    t = torch()
    if not _checked:
        _native.load()
        if not t.cuda.is_available():
            raise RuntimeError(
                "dem: no ROCm GPU is visible; the DEM operators run only on the HIP path"
            )
        _checked = True
    return t.device("cuda", t.cuda.current_device())


def stream() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch().cuda.current_stream().cuda_stream)


def ptr(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def to_device(a, dtype=np.float64):
    """numpy (or tensor) -> contiguous device tensor of `dtype`."""
    t = torch()
    if isinstance(a, t.Tensor):
        tt = {np.float64: t.float64, np.float32: t.float32, np.int32: t.int32}[dtype]
        return a.to(device=device(), dtype=tt).contiguous()
    arr = np.ascontiguousarray(a, dtype=dtype)
    return t.from_numpy(arr).to(device())


def empty(shape, dtype="float64"):
    t = torch()
    return t.empty(shape, dtype=getattr(t, dtype), device=device())


_hip = None


def _hip_runtime() -> ctypes.CDLL:
    """torch's HIP runtime (the library libdemhip links against)."""
    global _hip
    if _hip is None:
        import os

        _hip = ctypes.CDLL(os.path.join(os.path.dirname(torch().__file__), "lib", "libamdhip64.so"))
    return _hip


class HostWord:
    """One int32 in pinned host memory that kernels write through its device
    address.  arm() puts a sentinel (-1) there before the launch; read()
    polls until the kernels have published a value and returns it, with no
    device-to-host copy and no wait for the stream to drain (work the
    kernels still do is ordered before anything later on the stream); a
    word that does not arrive within a second falls back to a stream
    synchronisation."""

    def __init__(self):
        t = torch()
        device()
        self._h = t.zeros(1, dtype=t.int32, pin_memory=True)
        self._view = self._h.numpy()
        d = ctypes.c_void_p()
        rc = _hip_runtime().hipHostGetDevicePointer(
            ctypes.byref(d), ctypes.c_void_p(self._h.data_ptr()), 0)
        if rc != 0 or not d.value:
            raise RuntimeError(f"hipHostGetDevicePointer failed (status {rc})")
        self.dev = ctypes.c_void_p(d.value)

    def arm(self) -> None:
        self._view[0] = -1

    def read(self) -> int:
        import time

        v = self._view
        deadline = None
        while v[0] == -1:
            now = time.perf_counter()
            if deadline is None:
                deadline = now + 1.0
            elif now > deadline:
                torch().cuda.current_stream().synchronize()
                if v[0] == -1:
                    raise RuntimeError("dem: the kernels did not write the count word")
                break
        return int(v[0])


class Workspace:
    """A grow-only device scratch buffer (uint8)."""

    def __init__(self):
        self._buf = None

    def get(self, nbytes: int):
        t = torch()
        nbytes = max(int(nbytes), 256)
        if self._buf is None or self._buf.numel() < nbytes:
            self._buf = t.empty(nbytes, dtype=t.uint8, device=device())
        return self._buf

    def release(self):
        self._buf = None
