"""ctypes binding of libdemhip.so (the gfx950 HIP kernels behind the C ABI
declared in include/dem_hip.h).

There is no CPU fallback: if the library is missing or no GPU is visible the
operators raise.  Device buffers are PyTorch ROCm tensors; the library links
the HIP runtime that PyTorch ships, so tensor pointers and torch streams are
valid across the boundary.
"""

from __future__ import annotations

import ctypes
import os
import re
from pathlib import Path

LIB_PATH = Path(os.environ.get("DEM_HIP_LIB", Path(__file__).resolve().parent / "_lib" / "libdemhip.so"))
HEADER = Path(__file__).resolve().parents[2] / "include" / "dem_hip.h"

DEM_OK = 0
DEM_ERR_INVALID = -1
DEM_ERR_CAPACITY = -2
DEM_ERR_HIP = -3
DEM_ERR_WORKSPACE = -4


class NativeError(RuntimeError):
    """A libdemhip call returned a nonzero status."""

    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed ({code}): {msg}")
        self.code = code


class DemPhys(ctypes.Structure):
    _fields_ = [
        ("m_const", ctypes.c_double),
        ("eta_const", ctypes.c_double),
        ("L", ctypes.c_double),
        ("periodic", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class DemCorrection(ctypes.Structure):
    _fields_ = [
        ("grid", ctypes.c_void_p),
        ("grid64", ctypes.c_void_p),
        ("n", ctypes.c_int32),
        ("enabled", ctypes.c_int32),
        ("grid_min", ctypes.c_double),
        ("inv_dx", ctypes.c_double),
        ("Fz_inv", ctypes.c_double),
        ("L_ratio", ctypes.c_double),
        ("eta_ratio", ctypes.c_double),
        ("self_coeff", ctypes.c_double),
    ]


DEM_TREE_MAX_LEVELS = 64


class DemTreeInfo(ctypes.Structure):
    _fields_ = [
        ("n_capacity", ctypes.c_int32),
        ("max_nodes", ctypes.c_int32),
        ("n", ctypes.c_int32),
        ("node_count", ctypes.c_int32),
        ("parent_count", ctypes.c_int32),
        ("leaf_slot_count", ctypes.c_int32),
        ("num_levels", ctypes.c_int32),
        ("max_depth", ctypes.c_int32),
        ("level_offsets", ctypes.c_int32 * DEM_TREE_MAX_LEVELS),
        ("periodic", ctypes.c_int32),
        ("mpl", ctypes.c_int32),
        ("forces_valid", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("L", ctypes.c_double),
    ]


class DemBatchEnsemble(ctypes.Structure):
    _fields_ = [
        ("offset", ctypes.c_int32),
        ("n", ctypes.c_int32),
        ("capacity", ctypes.c_int32),
        ("steps", ctypes.c_int32),
        ("L", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("stokes_factor", ctypes.c_double),
        ("m_const", ctypes.c_double),
        ("eta_const", ctypes.c_double),
    ]


class DemBatchTuning(ctypes.Structure):
    _fields_ = [
        ("block", ctypes.c_int32),
        ("jlen", ctypes.c_int32),
        ("force_tile", ctypes.c_int32),
        ("conv_tile", ctypes.c_int32),
        ("contacts", ctypes.c_int32),
        ("pair_once", ctypes.c_int32),
    ]


_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_S = ctypes.c_size_t

# name -> (restype, argtypes)
SIGNATURES: dict[str, tuple] = {
    "dem_last_error": (ctypes.c_char_p, []),
    "dem_version": (_I, []),
    "dem_device_count": (_I, []),
    "dem_direct_workspace_bytes": (_S, [_I]),
    "dem_direct_forces": (_I, [_P, _P, _I, _P, _P, _P, _S, _P]),
    "dem_direct_convection": (_I, [_P, _P, _P, _I, _P, _P, _P, _P, _S, _P]),
    "dem_direct_forces_and_convection": (_I, [_P, _P, _I, _P, _P, _P, _P, _P, _S, _P]),
    "dem_direct_forces_and_convection_f64": (_I, [_P, _P, _I, _P, _P, _P, _P, _P]),
    "dem_direct_f64_targets": (_I, [_P, _P, _I, _P, _P, _P, _I, _P, _P, _P, _P]),
    "dem_direct_set_tuning": (_I, [_I, _I]),
    "dem_direct_set_symmetric": (_I, [_I]),
    "dem_direct_set_image32": (_I, [_I]),
    "dem_octree_set_exact_mac": (_I, [_I]),
    "dem_hash_workspace_bytes": (_S, [_I, _I]),
    "dem_hash_pairs": (_I, [_P, _P, _I, _D, _I, _I, _P, _I, _P, _P, _S, _P]),
    "dem_hash_pairs_noalloc": (_I, [_P, _P, _I, _D, _I, _I, _P, _I, _P, _P]),
    "dem_euler_step": (_I, [_P, _P, _P, _P, _I, _D, _D, _D, _I, _P]),
    "dem_euler_step_total": (_I, [_P, _P, _I, _D, _D, _I, _P]),
    "dem_merge_compact": (_I, [_P, _P, _I, _P, _I, _P, _I, _P, _P, _P]),
    "dem_gather_droplets": (_I, [_P, _P, _I, _P, _I, _P, _P]),
    "dem_octree_workspace_bytes": (_S, [_I, _I]),
    "dem_octree_build": (_I, [_P, _P, _I, _D, _I, _I, _D, _P, _P, _S, _P]),
    "dem_octree_forces": (_I, [_P, _D, _D, _P, _P, _S, _P]),
    "dem_octree_stokeslet": (_I, [_P, _D, _P, _D, _P, _P, _P, _S, _P]),
    "dem_octree_total_velocity": (_I, [_P, _D, _P, _D, _D, _P, _P, _P, _S, _P]),
    "dem_octree_field": (_I, [_P, ctypes.c_char_p, ctypes.POINTER(_S), ctypes.POINTER(_S)]),
    "dem_octree_collision_counts": (_I, [_P, _P, _I, _P, _P, _S, _P]),
    "dem_direct_collision_counts": (_I, [_P, _P, _I, _D, _I, _P, _I, _P, _P]),
    "dem_batch_pair_once_split": (_I, [_I, _I]),
    "dem_batch_split": (_I, [_I, _I]),
    "dem_batch_workspace_bytes": (_S, [_P, _I, _P]),
    "dem_batch_create": (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _S, _P, ctypes.POINTER(_P)]),
    "dem_batch_resize": (_I, [_P, _I, _P, _P, _P]),
    "dem_batch_check": (_I, [_P, _I, _P]),
    "dem_batch_advance": (_I, [_P, _P]),
    "dem_batch_merge": (_I, [_P, _I, _P, _P, _P, _P, _P, _P]),
    "dem_batch_merge_async": (_I, [_P, _I, _P, _P, _P, _P, _P, _P]),
    "dem_batch_wait_go": (_I, [_P, _P]),
    "dem_batch_set_incremental": (_I, [_P, _I]),
    "dem_merge_plan": (_I, [_I, _P, _P, _P, _P, _P, _P, _P, _P]),
    "dem_batch_capture": (_I, [_P]),
    "dem_batch_run": (_I, [_P, _I, _P]),
    "dem_batch_field": (_I, [_P, ctypes.c_char_p, ctypes.POINTER(_P), ctypes.POINTER(_S)]),
    "dem_batch_destroy": (_I, [_P]),
    "dem_adjoint_workspace_bytes": (_S, [_I]),
    "dem_direct_step_vjp": (_I, [_P, _P, _P, _P, _I, _P, _D, _P, _P, _P, _S, _P]),
    "dem_contact_loss": (_I, [_P, _P, _I, _P, _D, _D, _P, _P, _P, _S, _P]),
    "dem_profile_start": (_I, []),
    "dem_profile_stop": (_I, []),
    "dem_profile_read": (_I, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                              ctypes.POINTER(ctypes.c_int)]),
}


def profile_read(kernel: str) -> tuple[float, int]:
    """(total ms, launches) of `kernel` recorded since dem_profile_start()."""
    ms = ctypes.c_double(0.0)
    cnt = ctypes.c_int(0)
    call("dem_profile_read", kernel.encode(), ctypes.byref(ms), ctypes.byref(cnt))
    return ms.value, cnt.value

_lib: ctypes.CDLL | None = None


def load() -> ctypes.CDLL:
    """Load libdemhip.so once; raise ImportError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"libdemhip.so not found at {LIB_PATH}; build it with "
            "`python particle-simulator_amd/build.py` (no CPU fallback exists)"
        )
    lib = ctypes.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def header_functions() -> list[str]:
    """Function names declared in include/dem_hip.h."""
    text = HEADER.read_text()
    return sorted(set(re.findall(r"\b(dem_[a-z0-9_]+)\s*\(", text)))


def last_error() -> str:
    msg = load().dem_last_error()
    return msg.decode() if msg else ""


def call(name: str, *args) -> int:
    """Call `name` and raise NativeError on a negative status."""
    rc = getattr(load(), name)(*args)
    if rc < 0:
        raise NativeError(name, rc, last_error())
    return rc
