"""Position update and particle boundaries on one device
(PyPIC3D/particles/particle_tile_communication.py:82-99, 41-59, 292-453).

With one tile per GPU no particle changes tile, so the refresh reduces to the
per-axis particle boundary (periodic wrap, reflection, absorption), zeroing of
x and u in inactive slots, and an overflow flag that stays False.
"""

from __future__ import annotations

import torch

from .. import _native as N
from .._tensors import is_f64, make_desc, make_species, new_soa, particle_planes, pic_particles, shape_desc
from .particle_class import TiledParticles


def update_tiled_particle_positions(tiled_particles, species_config, dt):
    """x += v*dt where active & update_x (:82-99); returns new TiledParticles."""
    xp, up, act = particle_planes(tiled_particles)
    S, P = xp[0].shape
    sp = make_species(species_config, P)
    desc = shape_desc((3, 3, 3), 1, dt=float(dt))
    store, view = new_soa(S, P, xp[0].device, xp[0].dtype)
    N.call(N.op("pic_advance_positions", is_f64(xp[0])), desc, sp, pic_particles(xp, up, act), 1.0,
           N.p3([store[0], store[1], store[2]]), N.stream_handle())
    return tiled_particles._replace(x=view)


def refresh_tiled_particle_tiles(tiled_particles, static_parameters, dynamic_parameters):
    """Returns (particles, overflow) (:440-453)."""
    desc = make_desc(static_parameters, dynamic_parameters)
    xp, up, act = particle_planes(tiled_particles)
    S, P = xp[0].shape
    xs, xv = new_soa(S, P, xp[0].device, xp[0].dtype)
    us, uv = new_soa(S, P, xp[0].device, xp[0].dtype)
    for a in range(3):
        xs[a].copy_(xp[a])
        us[a].copy_(up[a])
    act_out = act.clone()
    sp = make_species_dummy(S, P)
    N.call(N.op("pic_particle_boundaries", is_f64(xp[0])), desc, sp,
           pic_particles([xs[0], xs[1], xs[2]], [us[0], us[1], us[2]], act_out),
           N.stream_handle())
    overflow = torch.zeros((), dtype=torch.bool, device=xp[0].device)
    return TiledParticles(x=xv, u=uv, active=act_out.view(1, 1, 1, S, P)), overflow


def make_species_dummy(S, P):
    sp = N.PicSpecies()
    sp.count = S
    sp.slots = P
    return sp
