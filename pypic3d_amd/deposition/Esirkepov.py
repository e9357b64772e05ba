"""Charge-conserving Esirkepov current (PyPIC3D/deposition/Esirkepov.py:48-362).

``particles.x`` is the old position after the push; the new position
x + v*dt (for update_x axes) is predicted inside the kernel.  The J input is
a template only: the deposit starts from zero, then the ghost fold and the
refresh use the particle boundary codes (bc_type=1, :357-359).  float64
particles and J run the FP64 path.
"""

from __future__ import annotations

from .. import _native as N
from .._tensors import as_field3, check_fields, is_f64, like_input, make_desc, make_species, particle_planes, pic_particles


def deposit_esirkepov_into(desc, sp, pp, J3, stream, fill=True):
    """Zero (optionally), deposit, fold and refresh into the 3-D tensors J3."""
    f64 = is_f64(J3[0])
    if fill:
        for c in J3:
            N.fill(c, 0.0, stream)
    N.call(N.op("pic_deposit_esirkepov", f64), desc, sp, pp, N.p3(J3), stream)
    finish_current(desc, J3, stream)


def finish_current(desc, J3, stream):
    f64 = is_f64(J3[0])
    for c in J3:
        N.call(N.op("pic_halo_fold", f64), desc, N.ptr(c), 1, stream)
        N.call(N.op("pic_halo_refresh", f64), desc, N.ptr(c), 1, stream)


def Esirkepov_current(particles, species_config, J, static_parameters, dynamic_parameters):
    desc = make_desc(static_parameters, dynamic_parameters)
    if desc.n[0] == 1 and desc.n[1] == 1 and desc.n[2] == 1:
        raise ValueError("Esirkepov deposition needs at least one active axis")
    xp, up, act = particle_planes(particles)
    sp = make_species(species_config, xp[0].shape[1])
    templ = [as_field3(c, "J") for c in J]
    check_fields(desc, *templ, dtype=xp[0].dtype)
    out = [t.new_empty(t.shape) for t in templ]
    deposit_esirkepov_into(desc, sp, pic_particles(xp, up, act), out, N.stream_handle())
    return tuple(like_input(o, t) for o, t in zip(out, J))
