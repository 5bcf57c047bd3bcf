"""Base class of post-processors (reference: dem/post_processor/post_processor_base.py)."""


class PostProcessor:
    def __init__(self, solution):
        self.solution = solution

    def plot(self):
        raise NotImplementedError

    def live_plot(self):
        raise NotImplementedError

    def initialize_live_plot(self):
        raise NotImplementedError

    def finalize_live_plot(self):
        raise NotImplementedError

    def on_close_live_plot(self):
        raise NotImplementedError
