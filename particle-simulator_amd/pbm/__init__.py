"""Population-balance side of the DEM coupling: the volume classes and the
GPU collision counts between them (the reference's ``pbm`` package,
pbm/__init__.py; its ODE solver, kernels and coupling driver are not part of
this build)."""

from .volume_grid import VolumeGrid

__all__ = ["VolumeGrid"]
