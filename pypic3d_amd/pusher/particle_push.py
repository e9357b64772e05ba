"""Fused field gather + Boris push (PyPIC3D/pusher/particle_push.py:13-175).

One HIP kernel gathers the six Yee-staggered components with the particle
shape (CIC or TSC, reduced axes collapsed to index g) and applies the
non-relativistic (boris.py:14-56) or relativistic (boris.py:60-124) Boris
update, or the Higuera-Cary update (higuera_cary.py:57-113) when
``particle_pusher == "higuera_cary"`` (particle_push.py:109-132).
"""

from __future__ import annotations

from .. import _native as N
from .._tensors import (as_field3, check_fields, is_f64, make_desc, make_species, new_soa, particle_planes,
                        pic_particles)
from ..particles.particle_class import TiledParticles


def check_pusher(static_parameters):
    if static_parameters.particle_pusher not in ("boris", "higuera_cary"):
        raise ValueError(f"Unknown particle_pusher: {static_parameters.particle_pusher}")


def particle_push(particles, species_config, E_tiles, B_tiles, static_parameters, dynamic_parameters,
                  E_external=None, B_external=None):
    """Returns TiledParticles with the pushed velocities; x and active are passed through."""
    check_pusher(static_parameters)
    desc = make_desc(static_parameters, dynamic_parameters)
    xp, up, act = particle_planes(particles)
    S, P = xp[0].shape
    sp = make_species(species_config, P)
    E3 = [as_field3(c, "E") for c in E_tiles]
    B3 = [as_field3(c, "B") for c in B_tiles]
    E_external = None if E_external is None else [as_field3(c, "E_ext") for c in E_external]
    B_external = None if B_external is None else [as_field3(c, "B_ext") for c in B_external]
    Ee = None if E_external is None else N.p3(E_external)
    Be = None if B_external is None else N.p3(B_external)
    check_fields(desc, *E3, *B3, *(E_external or ()), *(B_external or ()), dtype=xp[0].dtype)
    store, view = new_soa(S, P, xp[0].device, xp[0].dtype)
    N.call(N.op("pic_push", is_f64(xp[0])), desc, sp, pic_particles(xp, up, act), N.p3(E3), N.p3(B3), Ee, Be,
           N.p3([store[0], store[1], store[2]]), N.stream_handle())
    return TiledParticles(x=particles.x, u=view, active=particles.active)
