"""Empty, as in the reference (dem/utils/__init__.py)."""
