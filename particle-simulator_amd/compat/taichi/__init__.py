"""Minimal stand-in for the ``taichi`` module, for running the reference's
entry scripts on this build when Taichi is not installed.

The scripts only do ``import taichi as ti; ti.init(arch=ti.cpu,
cpu_max_num_threads=..., default_fp=ti.f64)`` at import time (main.py:11-15,
main_tree.py:12-16, main_multi.py:9-11); no computation goes through Taichi
in this build.  ``dem.run_script`` puts this package on sys.path only when
``importlib.util.find_spec("taichi")`` finds no real Taichi.
"""

cpu = "cpu"
gpu = "gpu"
cuda = "cuda"
vulkan = "vulkan"
f32 = "f32"
f64 = "f64"
i32 = "i32"
i64 = "i64"

_config: dict = {}


def init(arch=None, **kwargs):
    """Record the requested configuration; nothing to initialise."""
    _config.clear()
    _config.update({"arch": arch, **kwargs})


def sync():
    """Kernels of this build synchronise inside each operator call."""
    return None


def reset():
    _config.clear()
