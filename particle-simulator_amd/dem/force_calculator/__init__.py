from .direct_droplet_force_calculator import DirectDropletForceCalculator

__all__ = ["DirectDropletForceCalculator"]
