"""Base class of particle states (reference: dem/particle_state/particle_state_base.py)."""


class ParticleState:
    pass
