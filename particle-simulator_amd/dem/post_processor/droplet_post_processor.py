"""Solver-facing protocol of the reference's ``DropletPostProcessor``.

The reference's post-processor (dem/post_processor/droplet_post_processor.py,
matplotlib plots and statistics) is not on the hot path.  The solver only
uses ``stop_simulation`` (:26), ``update_solution`` (:28), ``live_plot``
(:60), ``update_live_plot`` (:195) and ``finalize_live_plot`` (:110); this
class provides exactly that protocol (live plotting is a no-op) so the
reference's scripts run unchanged.
"""

from __future__ import annotations

from .post_processor_base import PostProcessor


class DropletPostProcessor(PostProcessor):
    def __init__(self, solution, box_size):
        self.solution = solution
        self.box_size = box_size
        self.stop_simulation = False

    def update_solution(self, new_solution):
        self.solution = new_solution

    def plot(self):
        raise NotImplementedError("plotting is outside the accelerated hot path")

    def live_plot(self):
        pass

    def initialize_live_plot(self):
        pass

    def update_live_plot(self):
        pass

    def finalize_live_plot(self):
        pass

    def on_close_live_plot(self, event=None):
        self.stop_simulation = True
