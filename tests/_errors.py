"""Error metrics shared by the parity tests and bench.py.

``rel_err``: per-particle relative error of a vector field,
    e_i = |x_i - ref_i| / max(|ref_i|, floor * rms_j |ref_j|)
The floor keeps particles whose net force/velocity nearly cancels from
dividing by ~0; ``floor`` is stated next to each tolerance.
"""

from __future__ import annotations

import numpy as np


def rel_err(x, ref, floor: float = 1e-3) -> np.ndarray:
    x = np.asarray(x, dtype=np.float64).reshape(-1, 3)
    ref = np.asarray(ref, dtype=np.float64).reshape(-1, 3)
    nr = np.linalg.norm(ref, axis=1)
    rms = float(np.sqrt(np.mean(nr**2))) if len(nr) else 0.0
    den = np.maximum(nr, floor * rms)
    den = np.where(den > 0, den, 1.0)
    return np.linalg.norm(x - ref, axis=1) / den


def summary(x, ref, floor: float = 1e-3) -> dict:
    e = rel_err(x, ref, floor)
    return {
        "max": float(e.max()) if len(e) else 0.0,
        "p999": float(np.quantile(e, 0.999)) if len(e) else 0.0,
        "mean": float(e.mean()) if len(e) else 0.0,
        "norm": float(np.linalg.norm(np.asarray(x) - np.asarray(ref)) / max(np.linalg.norm(ref), 1e-300)),
    }
