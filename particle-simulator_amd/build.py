"""Build libdemhip.so (gfx950) in-tree.

Every `.hip` file under csrc/ is compiled with hipcc for gfx950 and the
objects are linked against the HIP runtime that PyTorch ships, so the
library and torch share one runtime (device pointers and streams from torch
tensors are valid in the library).  Usage: ``python build.py [--jobs N]``.
"""

from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT = HERE / "dem" / "_lib" / "libdemhip.so"
OBJ = HERE / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    # packed FP32 (v_pk_*_f32) is kept out of every kernel.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


# Per-file additions to HIP_FLAGS.  The pair-once direct kernels, single-
# system and batched, take the ILP-first instruction scheduler: under the
# default one 96 of the 32-bit Oseen loop's 258 VALU instructions wait on the
# instruction before them (batched 64-bit loop: 181 of 615), which costs
# issue slots even at four waves per SIMD.
FILE_FLAGS = {
    "direct_pairs": ["-mllvm", "-amdgpu-sched-strategy=max-ilp"],
    "batch": ["-mllvm", "-amdgpu-sched-strategy=max-ilp"],
}


def _torch_lib_dir() -> Path:
    import importlib.util

    spec = importlib.util.find_spec("torch")
    if spec is None or spec.origin is None:
        raise RuntimeError("torch is required (its HIP runtime is linked)")
    return Path(spec.origin).parent / "lib"


def _compile(src: Path) -> Path:
    obj = OBJ / (src.stem + ".o")
    cmd = [HIPCC, *HIP_FLAGS, *FILE_FLAGS.get(src.stem, []), "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs: int = 8, verbose: bool = True) -> Path:
    OBJ.mkdir(parents=True, exist_ok=True)
    OUT.parent.mkdir(parents=True, exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, srcs))
    tlib = _torch_lib_dir()
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [
        "g++", "-shared", "-o", str(tmp), *map(str, objs),
        f"-L{tlib}", "-l:libamdhip64.so", f"-Wl,-rpath,{tlib}", "-Wl,--no-undefined",
    ]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT} from {len(srcs)} sources")
    return OUT


def clean() -> None:
    shutil.rmtree(OBJ, ignore_errors=True)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    a = ap.parse_args()
    try:
        build(a.jobs)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
