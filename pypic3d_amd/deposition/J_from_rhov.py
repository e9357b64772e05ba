"""Direct current J = q v S (PyPIC3D/deposition/J_from_rhov.py:31-260).

Node weights along the other axes, face weights along the current's axis with
the face offset (x - grid0) - (anchor + 1/2) dx (:139-141) -- these
extrapolate, as in the reference.  Then fold + refresh with the particle
codes, and the optional bilinear / digital filter (+ refresh).
"""

from __future__ import annotations

from .. import _native as N
from .._tensors import as_field3, check_fields, is_f64, like_input, make_desc, make_species, particle_planes, pic_particles


def deposit_direct_into(desc, sp, pp, J3, current_filter, stream, scratch=None, tile=None, starts=None):
    """tile/starts: tile-ordered deposit for particles sorted by pic_tile_keys
    (FP32 only); float64 J runs the FP64 path."""
    f64 = is_f64(J3[0])
    for c in J3:
        N.fill(c, 0.0, stream)
    if tile and not f64:
        N.call("pic_deposit_direct_tiled", desc, sp, pp, N.p3(J3), int(tile), N.ptr(starts), stream)
    else:
        N.call(N.op("pic_deposit_direct", f64), desc, sp, pp, N.p3(J3), stream)
    finish_direct_current(desc, J3, current_filter, stream, scratch)


def finish_direct_current(desc, J3, current_filter, stream, scratch=None):
    """Fold + refresh of the raw direct deposit, then the optional filter
    (J_from_rhov.py:224-258)."""
    f64 = is_f64(J3[0])
    for c in J3:
        N.call(N.op("pic_halo_fold", f64), desc, N.ptr(c), 1, stream)
        N.call(N.op("pic_halo_refresh", f64), desc, N.ptr(c), 1, stream)
    if current_filter in ("bilinear", "digital"):
        tmp = scratch if scratch is not None else [c.clone() for c in J3]
        for c, t in zip(J3, tmp):
            if current_filter == "bilinear":
                N.call(N.op("pic_filter_bilinear", f64), desc, N.ptr(c), N.ptr(t), stream)
            else:
                N.call(N.op("pic_filter_digital", f64), desc, N.ptr(c), N.ptr(t), float(desc.alpha), stream)
            N.call(N.op("pic_halo_refresh", f64), desc, N.ptr(t), 1, stream)
            c.copy_(t)


def J_from_rhov(particles, species_config, J, static_parameters, dynamic_parameters):
    desc = make_desc(static_parameters, dynamic_parameters)
    xp, up, act = particle_planes(particles)
    sp = make_species(species_config, xp[0].shape[1])
    templ = [as_field3(c, "J") for c in J]
    check_fields(desc, *templ, dtype=xp[0].dtype)
    out = [t.new_empty(t.shape) for t in templ]
    deposit_direct_into(desc, sp, pic_particles(xp, up, act), out, static_parameters.current_filter,
                        N.stream_handle())
    return tuple(like_input(o, t) for o, t in zip(out, J))
