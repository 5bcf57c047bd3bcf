"""Base class of solutions (reference: dem/solution/solution_base.py)."""


class Solution:
    def __init__(self):
        pass

    def save_step(self, positions):
        raise NotImplementedError("save_step must be implemented by a subclass")
