"""Particle containers with the reference's field names (PyPIC3D/particles/particle_class.py:5-31).

``u`` holds the velocity v (not gamma*v), as in the reference.  On the GPU,
``x`` and ``u`` are (1, 1, 1, S, P, 3) float32 views over structure-of-arrays
storage (3, S, P): each component plane is contiguous.
"""

from typing import NamedTuple


class SpeciesConfig(NamedTuple):
    charge: object     # (S,)
    mass: object       # (S,)
    weight: object     # (S,)
    update_x: object   # (S, 3) bool
    update_u: object   # (S, 3) bool


class TiledParticles(NamedTuple):
    x: object          # (1, 1, 1, S, P, 3) positions
    u: object          # (1, 1, 1, S, P, 3) velocities
    active: object     # (1, 1, 1, S, P) bool
