"""Charge density (PyPIC3D/deposition/rho.py:29-196).

rho += q w / (dx dy dz) * S_x S_y S_z at node points, fold + refresh with the
particle codes, digital filter (+ refresh) when current_filter == "digital".
"""

from __future__ import annotations

from .. import _native as N
from .._tensors import as_field3, check_fields, is_f64, like_input, make_desc, make_species, particle_planes, pic_particles


def deposit_rho_into(desc, sp, pp, rho3, current_filter, stream, scratch=None, tile=None, starts=None):
    """tile/starts: tile-ordered deposit for particles sorted by pic_tile_keys
    (FP32 only); float64 rho runs the FP64 path."""
    f64 = is_f64(rho3)
    N.fill(rho3, 0.0, stream)
    if tile and not f64:
        N.call("pic_deposit_rho_tiled", desc, sp, pp, N.ptr(rho3), int(tile), N.ptr(starts), stream)
    else:
        N.call(N.op("pic_deposit_rho", f64), desc, sp, pp, N.ptr(rho3), stream)
    N.call(N.op("pic_halo_fold", f64), desc, N.ptr(rho3), 1, stream)
    N.call(N.op("pic_halo_refresh", f64), desc, N.ptr(rho3), 1, stream)
    if current_filter == "digital":
        tmp = scratch if scratch is not None else rho3.clone()
        N.call(N.op("pic_filter_digital", f64), desc, N.ptr(rho3), N.ptr(tmp), float(desc.alpha), stream)
        N.call(N.op("pic_halo_refresh", f64), desc, N.ptr(tmp), 1, stream)
        rho3.copy_(tmp)


def compute_rho(particles, species_config, rho, static_parameters, dynamic_parameters):
    desc = make_desc(static_parameters, dynamic_parameters)
    xp, up, act = particle_planes(particles)
    sp = make_species(species_config, xp[0].shape[1])
    templ = as_field3(rho, "rho")
    check_fields(desc, templ, dtype=xp[0].dtype)
    out = templ.new_empty(templ.shape)
    deposit_rho_into(desc, sp, pic_particles(xp, up, act), out, static_parameters.current_filter,
                     N.stream_handle())
    return like_input(out, rho)
