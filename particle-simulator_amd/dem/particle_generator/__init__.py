from .uniform_droplet_generator import UniformDropletGenerator

__all__ = ["UniformDropletGenerator"]
