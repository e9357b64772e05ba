"""Electrostatic Poisson CG and field (PyPIC3D/solvers/electrostatic_yee.py:71-276).

The CG keeps the potential and the residual in FP64 (recomputed from the
potential, r = rho/eps + lapl(phi)) and solves each correction with FP32 CG
vectors and FP64 dot products (mixed-precision iterative refinement).  It
stops at the reference's test ||r|| <= tol (absolute, default 1e-12) on the
FP64 residual, optionally relaxed to ``rtol * ||r0||``, when a refinement pass
no longer halves ||r|| (the FP64 rounding floor of a large grid), or after
``max_iter`` CG iterations in total.  ``last_solve_info`` holds the iteration
count and final FP64 residual norm of the most recent solve.
"""

from __future__ import annotations

import ctypes

import torch

from .. import _native as N
from .._tensors import as_field3, check_fields, like_input, make_desc, make_species, particle_planes, pic_particles
from ..deposition.rho import deposit_rho_into

DEFAULT_RTOL = 0.0
last_solve_info = {"iterations": None, "residual_norm": None}
_WORKSPACE: dict = {}


def _workspace(desc, device):
    nbytes = int(N.lib().pic_cg_workspace_bytes(desc))
    if nbytes < 0:
        raise ValueError("invalid grid for the CG workspace")
    key = (str(device), nbytes)
    ws = _WORKSPACE.get(key)
    if ws is None:
        _WORKSPACE.clear()
        ws = torch.empty(nbytes, dtype=torch.uint8, device=device)
        _WORKSPACE[key] = ws
    return ws


def poisson_cg_into(desc, rho3, phi3, tol, max_iter, stream, rtol=DEFAULT_RTOL):
    ws = _workspace(desc, phi3.device)
    iters = ctypes.c_int32(0)
    res = ctypes.c_double(0.0)
    N.call("pic_poisson_cg", desc, N.ptr(rho3), N.ptr(phi3), float(tol), float(rtol), int(max_iter),
           N.ptr(ws), ws.numel(), ctypes.byref(iters), ctypes.byref(res), stream)
    last_solve_info["iterations"] = int(iters.value)
    last_solve_info["residual_norm"] = float(res.value)
    return phi3


def solve_poisson_with_conjugate_gradient(rho, phi, static_parameters, dynamic_parameters, tol=1e-12, max_iter=5000,
                                          rtol=DEFAULT_RTOL):
    """Returns phi (electrostatic_yee.py:71-156)."""
    desc = make_desc(static_parameters, dynamic_parameters)
    rho3 = as_field3(rho, "rho")
    out = as_field3(phi, "phi").clone()
    check_fields(desc, rho3, out, dtype=torch.float32)
    poisson_cg_into(desc, rho3, out, tol, max_iter, N.stream_handle(), rtol=rtol)
    return like_input(out, phi)


def phi_boundaries_into(desc, phi3, stream):
    """_apply_tiled_phi_constant_boundaries (electrostatic_yee.py:26-43), in place."""
    applied = False
    for axis in range(3):
        if desc.bc[axis] == 1:
            N.call("pic_constant_boundary", desc, N.ptr(phi3), axis, stream)
            applied = True
    if not applied:
        N.call("pic_halo_refresh", desc, N.ptr(phi3), 0, stream)


def electrostatic_fields_into(desc, sp, pp, rho3, phi3, E3, current_filter, stream, tol=1e-12, max_iter=5000,
                              rtol=DEFAULT_RTOL, scratch=None):
    """rho -> CG -> phi BC -> digital filter -> phi BC -> E = -grad phi (:249-276)."""
    deposit_rho_into(desc, sp, pp, rho3, current_filter, stream)
    poisson_cg_into(desc, rho3, phi3, tol, max_iter, stream, rtol=rtol)
    phi_boundaries_into(desc, phi3, stream)
    if desc.alpha != 1.0:
        tmp = scratch if scratch is not None else phi3.clone()
        N.call("pic_filter_digital", desc, N.ptr(phi3), N.ptr(tmp), float(desc.alpha), stream)
        phi3.copy_(tmp)
    phi_boundaries_into(desc, phi3, stream)
    N.call("pic_efield_from_phi", desc, N.ptr(phi3), N.p3(E3), stream)


def calculate_tiled_electrostatic_fields(static_parameters, dynamic_parameters, particles, species_config, rho_tiles,
                                         phi_tiles, tol=1e-12, max_iter=5000, rtol=DEFAULT_RTOL):
    """Returns (E_tiles, phi_tiles, rho_tiles)."""
    desc = make_desc(static_parameters, dynamic_parameters)
    xp, up, act = particle_planes(particles)
    sp = make_species(species_config, xp[0].shape[1])
    rho_t = as_field3(rho_tiles, "rho")
    rho3 = rho_t.new_empty(rho_t.shape)
    phi3 = as_field3(phi_tiles, "phi").clone()
    check_fields(desc, rho_t, phi3, dtype=torch.float32)
    E3 = [rho_t.new_empty(rho_t.shape) for _ in range(3)]
    electrostatic_fields_into(desc, sp, pic_particles(xp, up, act), rho3, phi3, E3, static_parameters.current_filter,
                              N.stream_handle(), tol=tol, max_iter=max_iter, rtol=rtol)
    E = tuple(like_input(e, rho_tiles) for e in E3)
    return E, like_input(phi3, phi_tiles), like_input(rho3, rho_tiles)


def calculate_electrostatic_fields(static_parameters, dynamic_parameters, particles, rho, phi, solver, bc,
                                   tol=1e-12, max_iter=5000, rtol=DEFAULT_RTOL):
    """Single-tile electrostatic field solve from a given rho
    (electrostatic_yee.py:159-209): CG -> phi boundaries -> digital filter ->
    phi boundaries -> E = -centred gradient of the interior with periodic
    wrap (jnp.roll on the interior, :46-60) -> E ghosts refreshed with the
    field codes.  rho is not deposited here (particles, solver and bc are
    accepted and unused, as in the reference).  Returns ((Ex, Ey, Ez), phi, rho)."""
    del particles, solver, bc
    desc = make_desc(static_parameters, dynamic_parameters)
    rho3 = as_field3(rho, "rho")
    phi3 = as_field3(phi, "phi").clone()
    check_fields(desc, rho3, phi3, dtype=torch.float32)
    s = N.stream_handle()
    poisson_cg_into(desc, rho3, phi3, tol, max_iter, s, rtol=rtol)
    phi_boundaries_into(desc, phi3, s)
    if desc.alpha != 1.0:
        tmp = phi3.clone()
        N.call("pic_filter_digital", desc, N.ptr(phi3), N.ptr(tmp), float(desc.alpha), s)
        phi3.copy_(tmp)
    phi_boundaries_into(desc, phi3, s)
    # the roll wraps the interior on every axis: a periodic refresh of a copy
    # supplies exactly those neighbours
    periodic = N.PicDesc.from_buffer_copy(desc)
    for a in range(3):
        periodic.bc[a] = 0
    wrapped = phi3.clone()
    N.call("pic_halo_refresh", periodic, N.ptr(wrapped), 0, s)
    E3 = [torch.zeros_like(phi3) for _ in range(3)]
    N.call("pic_efield_from_phi", periodic, N.ptr(wrapped), N.p3(E3), s)
    for c in E3:
        N.call("pic_halo_refresh", desc, N.ptr(c), 0, s)
    E = tuple(like_input(e, phi) for e in E3)
    return E, like_input(phi3, phi), rho
