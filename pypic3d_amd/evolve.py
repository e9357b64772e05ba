"""One PIC step with the reference's signatures (PyPIC3D/evolve.py:16-161).

``fields`` is the reference's 8-tuple (E, B, J, rho, phi, external_fields,
pml_state, overflow).  The Esirkepov branch runs its particle phase -- gather,
push, deposit, position update, particle boundaries (evolve.py:33-43, 70-79)
-- as one fused HIP kernel; the direct branch and the electrostatic step call
the individual kernels in the reference's order.  float64 particles and
fields run the FP64 path (the reference's precision): the same sequence with
the FP64 kernels, the Esirkepov branch as separate push, deposit, advance and
boundary launches.  Inputs are never modified; new tensors are returned.
External fields that are all zero are skipped (their sum is the identity).
"""

from __future__ import annotations

import torch

from . import _native as N
from ._tensors import (as_field3, check_fields, is_f64, like_input, make_desc, make_species, new_soa, particle_planes,
                       pic_particles)
from .deposition.Esirkepov import deposit_esirkepov_into, finish_current
from .deposition.J_from_rhov import deposit_direct_into
from .particles.particle_class import TiledParticles
from .pusher.particle_push import check_pusher
from .solvers.electrostatic_yee import electrostatic_fields_into
from .solvers.first_order_yee import _check_pml, _pml_args, yee_b, yee_e
from .utils import is_all_zero

__all__ = ["time_loop_electrodynamic", "time_loop_electrostatic"]


def _external(desc, external_fields, dtype):
    ext_E, ext_B = external_fields
    e = [as_field3(c, "external E") for c in ext_E]
    b = [as_field3(c, "external B") for c in ext_B]
    check_fields(desc, *e, *b, dtype=dtype)
    Ee = None if all(is_all_zero(c) for c in e) else N.p3(e)
    Be = None if all(is_all_zero(c) for c in b) else N.p3(b)
    return Ee, Be


def _advance(desc, sp, xp, up, act, scale, stream):
    S, P = xp[0].shape
    store, view = new_soa(S, P, xp[0].device, xp[0].dtype)
    N.call(N.op("pic_advance_positions", is_f64(xp[0])), desc, sp, pic_particles(xp, up, act), float(scale),
           N.p3([store[0], store[1], store[2]]), stream)
    return [store[0], store[1], store[2]], view


def time_loop_electrodynamic(particles, species_config, fields, static_parameters, dynamic_parameters):
    E, B, J, rho, phi, external_fields, pml_state, overflow = fields
    _check_pml(pml_state)
    check_pusher(static_parameters)
    desc = make_desc(static_parameters, dynamic_parameters)
    stream = N.stream_handle()
    xp, up, act = particle_planes(particles)
    S, P = xp[0].shape
    sp = make_species(species_config, P)
    E3 = [as_field3(c, "E") for c in E]
    B3 = [as_field3(c, "B") for c in B]
    J_t = [as_field3(c, "J") for c in J]
    J3 = [t.new_empty(t.shape) for t in J_t]
    dtype = xp[0].dtype
    f64 = dtype == torch.float64
    check_fields(desc, *E3, *B3, *J_t, dtype=dtype)
    Ee, Be = _external(desc, external_fields, dtype)
    dev = xp[0].device

    if static_parameters.current_deposition == "esirkepov" and f64:
        if desc.n[0] == 1 and desc.n[1] == 1 and desc.n[2] == 1:
            raise ValueError("Esirkepov deposition needs at least one active axis")
        # evolve.py:33-43, 70-79 as separate FP64 launches: push, deposit from
        # (x, pushed v), x += v dt, particle boundaries
        us, uv = new_soa(S, P, dev, dtype)
        u_planes = [us[0], us[1], us[2]]
        N.call("pic_push_f64", desc, sp, pic_particles(xp, up, act), N.p3(E3), N.p3(B3), Ee, Be, N.p3(u_planes),
               stream)
        deposit_esirkepov_into(desc, sp, pic_particles(xp, u_planes, act), J3, stream)
        act_o = act.clone()
        xf, xv = _advance(desc, sp, xp, u_planes, act_o, 1.0, stream)
        N.call("pic_particle_boundaries_f64", desc, sp, pic_particles(xf, u_planes, act_o), stream)
        new_particles = TiledParticles(x=xv, u=uv, active=act_o.view(1, 1, 1, S, P))
    elif static_parameters.current_deposition == "esirkepov":
        if desc.n[0] == 1 and desc.n[1] == 1 and desc.n[2] == 1:
            raise ValueError("Esirkepov deposition needs at least one active axis")
        xs, xv = new_soa(S, P, dev)
        us, uv = new_soa(S, P, dev)
        act_out = act.new_empty(act.shape)
        for c in J3:
            N.call("pic_fill_f32", N.ptr(c), c.numel(), 0.0, stream)
        out = pic_particles([xs[0], xs[1], xs[2]], [us[0], us[1], us[2]], act_out)
        N.call("pic_fused_push_esirkepov", desc, sp, pic_particles(xp, up, act), N.p3(E3), N.p3(B3), Ee, Be,
               out, N.p3(J3), stream)
        finish_current(desc, J3, stream)
        new_particles = TiledParticles(x=xv, u=uv, active=act_out.view(1, 1, 1, S, P))
    else:
        us, uv = new_soa(S, P, dev, dtype)
        N.call(N.op("pic_push", f64), desc, sp, pic_particles(xp, up, act), N.p3(E3), N.p3(B3), Ee, Be,
               N.p3([us[0], us[1], us[2]]), stream)
        u_planes = [us[0], us[1], us[2]]
        act_h = act.clone()
        xh, _ = _advance(desc, sp, xp, u_planes, act_h, 0.5, stream)
        N.call(N.op("pic_particle_boundaries", f64), desc, sp, pic_particles(xh, u_planes, act_h), stream)
        deposit_direct_into(desc, sp, pic_particles(xh, u_planes, act_h), J3, static_parameters.current_filter,
                            stream)
        xf, xv = _advance(desc, sp, xh, u_planes, act_h, 0.5, stream)
        N.call(N.op("pic_particle_boundaries", f64), desc, sp, pic_particles(xf, u_planes, act_h), stream)
        new_particles = TiledParticles(x=xv, u=uv, active=act_h.view(1, 1, 1, S, P))

    pml_b = pml_e = None
    if pml_state is not None:      # PML memories advance in cloned arrays (first_order_yee.py:54-59, 133-138)
        b_mem, prof3, b_new = _pml_args(desc, pml_state, 1)
        e_mem, _, e_new = _pml_args(desc, pml_state, 0)
        pml_b, pml_e = (b_mem, prof3), (e_mem, prof3)
        pml_state = (e_new, b_new, pml_state[2])
    B1 = [b.new_empty(b.shape) for b in B3]
    yee_b(desc, E3, B3, B1, False, stream, pml=pml_b)
    E1 = [e.new_empty(e.shape) for e in E3]
    yee_e(desc, E3, B1, J3, E1, stream, pml=pml_e)
    B2 = [b.new_empty(b.shape) for b in B3]
    yee_b(desc, E1, B1, B2, True, stream, pml=pml_b)

    out_fields = (
        tuple(like_input(e, t) for e, t in zip(E1, E)),
        tuple(like_input(b, t) for b, t in zip(B2, B)),
        tuple(like_input(j, t) for j, t in zip(J3, J)),
        rho, phi, external_fields, pml_state, overflow,
    )
    return new_particles, out_fields


def time_loop_electrostatic(particles, species_config, fields, static_parameters, dynamic_parameters):
    E, B, J, rho, phi, external_fields, pml_state, overflow = fields
    check_pusher(static_parameters)
    desc = make_desc(static_parameters, dynamic_parameters)
    stream = N.stream_handle()
    xp, up, act = particle_planes(particles)
    if is_f64(xp[0]):
        raise ValueError("the electrostatic step runs in FP32 (float32 particles and fields; the Poisson solve "
                         "refines the potential in FP64)")
    S, P = xp[0].shape
    sp = make_species(species_config, P)
    E3 = [as_field3(c, "E") for c in E]
    B3 = [as_field3(c, "B") for c in B]
    check_fields(desc, *E3, *B3, dtype=torch.float32)
    Ee, Be = _external(desc, external_fields, torch.float32)
    dev = xp[0].device
    us, uv = new_soa(S, P, dev)
    N.call("pic_push", desc, sp, pic_particles(xp, up, act), N.p3(E3), N.p3(B3), Ee, Be,
           N.p3([us[0], us[1], us[2]]), stream)
    u_planes = [us[0], us[1], us[2]]
    act_o = act.clone()
    xf, xv = _advance(desc, sp, xp, u_planes, act_o, 1.0, stream)
    N.call("pic_particle_boundaries", desc, sp, pic_particles(xf, u_planes, act_o), stream)
    rho_t = as_field3(rho, "rho")
    rho3 = rho_t.new_empty(rho_t.shape)
    phi3 = as_field3(phi, "phi").clone()
    check_fields(desc, rho_t, phi3, dtype=torch.float32)
    E_new = [e.new_empty(e.shape) for e in E3]
    electrostatic_fields_into(desc, sp, pic_particles(xf, u_planes, act_o), rho3, phi3, E_new,
                              static_parameters.current_filter, stream)
    out_fields = (
        tuple(like_input(e, t) for e, t in zip(E_new, E)),
        B, J, like_input(rho3, rho), like_input(phi3, phi), external_fields, pml_state, overflow,
    )
    return TiledParticles(x=xv, u=uv, active=act_o.view(1, 1, 1, S, P)), out_fields
