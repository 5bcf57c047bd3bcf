from .droplet_post_processor import DropletPostProcessor

__all__ = ["DropletPostProcessor"]
