"""Shared test set-up.

* puts the repository root (for ``oracle``) and ``particle-simulator_amd``
  (for the ``dem`` drop-in package) on sys.path;
* registers the ``gpu`` marker: tests that call libdemhip kernels.  They run
  with ``-m gpu`` on a machine with an MI355X; there is no CPU fallback.
"""

from __future__ import annotations

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "particle-simulator_amd"
for p in (str(ROOT), str(PKG)):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X and libdemhip.so")
    config.addinivalue_line("markers", "slow: long-running test")


@pytest.fixture(scope="session")
def gpu():
    """Fail (not skip) when the GPU path cannot run: gpu tests must exercise
    the HIP kernels."""
    import torch

    from dem import _native

    _native.load()
    assert torch.cuda.is_available(), "gpu test without a visible ROCm device"
    assert _native.load().dem_device_count() >= 1
    return torch.device("cuda", 0)
