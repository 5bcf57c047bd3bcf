"""Base class of the force operators (reference:
dem/force_calculator/force_calculator_base.py:1-10)."""


class ForceCalculator:
    """Abstract force operator."""

    def calculate(self):
        raise NotImplementedError("calculate must be implemented by a subclass")

    def calculate_forces_and_convection(self, positions, radii):
        raise NotImplementedError("calculate_forces_and_convection must be implemented by a subclass")
