"""Run one of the reference's entry scripts (main.py, main_tree.py,
main_multi.py) unchanged on this build.

    python -m dem.run_script /path/to/main.py [--inputs]

* puts this build's ``dem`` package first on sys.path;
* adds the ``taichi`` stand-in (particle-simulator_amd/compat) only when no
  real Taichi is importable;
* with ``--inputs``, writes the xlsx initial states the scripts load
  (results/droplets_N1000_vol0.02_0.xlsx, results/droplets_N10000_vol0.01_{1..10}.xlsx)
  into the script's directory when they are missing (the reference does not
  ship them);
* runs the script as ``__main__`` with the script's directory as cwd.
"""

from __future__ import annotations

import argparse
import importlib.util
import os
import runpy
import sys
from pathlib import Path

PKG_ROOT = Path(__file__).resolve().parents[1]
COMPAT = PKG_ROOT / "compat"


def ensure_paths() -> None:
    if str(PKG_ROOT) not in sys.path:
        sys.path.insert(0, str(PKG_ROOT))
    if importlib.util.find_spec("taichi") is None and str(COMPAT) not in sys.path:
        sys.path.append(str(COMPAT))


def script_box(num_particles: int, water_volume_content: float) -> float:
    """Box edge the reference's scripts compute (main.py:35-40)."""
    import numpy as np

    radii_range = np.array([2.5e-6, 7.5e-6])
    return float(np.cbrt((np.pi * num_particles * np.sum(radii_range) *
                          np.sum(np.square(radii_range))) / (3 * water_volume_content)))


def write_inputs(directory: str | os.PathLike, seed: int = 0) -> list[Path]:
    """xlsx initial states for main.py / main_tree.py / main_multi.py."""
    from dem.particle_generator import UniformDropletGenerator

    out = []
    res = Path(directory) / "results"
    res.mkdir(parents=True, exist_ok=True)
    jobs = [(1000, 0.02, 0)] + [(10000, 0.01, k) for k in range(1, 11)]
    for n, wvc, k in jobs:
        path = res / f"droplets_N{n}_vol{wvc}_{k}.xlsx"
        if path.exists():
            continue
        box = script_box(n, wvc)
        gen = UniformDropletGenerator(coord_range=(0, box), radii_range=(2.5e-6, 7.5e-6),
                                      num_particles=n, minimum_distance=1e-6, seed=seed + k)
        gen.generate().export_to_xlsx(str(path))
        out.append(path)
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("script")
    ap.add_argument("--inputs", action="store_true")
    a = ap.parse_args(argv)
    ensure_paths()
    script = Path(a.script).resolve()
    os.chdir(script.parent)
    if a.inputs:
        write_inputs(script.parent)
    sys.argv = [str(script)]
    runpy.run_path(str(script), run_name="__main__")
    return 0


if __name__ == "__main__":
    ensure_paths()
    sys.exit(main())
