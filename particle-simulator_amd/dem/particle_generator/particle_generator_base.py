"""Base class of particle generators (reference: dem/particle_generator/particle_generator_base.py)."""


class ParticleGenerator:
    def generate(self, count):
        raise NotImplementedError

    def save(self, filename):
        raise NotImplementedError

    def load(self, filename):
        raise NotImplementedError

    def plot(self):
        raise NotImplementedError
