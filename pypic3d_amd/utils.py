"""Hot-path helpers from PyPIC3D/utils.py: energy diagnostic (:108-187),
momentum diagnostic (:190-202), total charge,
external-field sum (:205-216) and the Courant time step (:761-801).

``compute_energy`` accumulates in FP64 on the GPU and returns three 0-d
float64 CUDA tensors (no host synchronisation).
"""

from __future__ import annotations

import weakref

import torch

from . import _native as N
from ._tensors import as_field3, check_fields, make_desc, make_species, particle_planes, pic_particles

_ENERGY_WS: dict = {}


def compute_energy(particles, E, B, static_parameters, dynamic_parameters, species_config=None):
    """(electric, magnetic, kinetic) energy; kinetic uses mass*weight (utils.py:162-183)."""
    if species_config is None:
        raise ValueError("species_config is required for the kinetic energy")
    desc = make_desc(static_parameters, dynamic_parameters)
    xp, up, act = particle_planes(particles)
    sp = make_species(species_config, xp[0].shape[1])
    E3 = [as_field3(c, "E") for c in E]
    B3 = [as_field3(c, "B") for c in B]
    check_fields(desc, *E3, *B3, dtype=xp[0].dtype)
    f64 = xp[0].dtype == torch.float64
    nbytes = 3 * 1024 * 8 if f64 else int(N.lib().pic_energy_workspace_bytes(desc, sp))
    dev = E3[0].device
    ws = _ENERGY_WS.get((str(dev), nbytes))
    if ws is None:
        ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        _ENERGY_WS[(str(dev), nbytes)] = ws
    out = torch.empty(3, dtype=torch.float64, device=dev)
    N.call(N.op("pic_energy", f64), desc, sp, pic_particles(xp, up, act), N.p3(E3), N.p3(B3), N.ptr(out),
           N.ptr(ws), ws.numel(), N.stream_handle())
    return out[0], out[1], out[2]


def add_external_fields(E, B, external_fields):
    external_E, external_B = external_fields
    total_E = tuple(e + ee for e, ee in zip(E, external_E))
    total_B = tuple(b + bb for b, bb in zip(B, external_B))
    return total_E, total_B


def courant_condition(courant_number, dx, dy, dz, dynamic_parameters):
    """dt = cfl / (C * sum over axes with N > 1 of 1/d)."""
    C = float(dynamic_parameters.C)
    Ns = [int(dynamic_parameters.Nx), int(dynamic_parameters.Ny), int(dynamic_parameters.Nz)]
    inv = sum(1.0 / float(d) for d, n in zip((dx, dy, dz), Ns) if n > 1)
    return courant_number / (C * inv)


_ZERO_CACHE: dict = {}


def is_all_zero(t) -> bool:
    """True when a tensor holds only zeros.  Cached per live tensor object and
    its in-place version counter, so a field that is never modified is
    checked (with one host synchronisation) only once."""
    hit = _ZERO_CACHE.get(id(t))
    if hit is not None and hit[0]() is t and hit[1] == t._version:
        return hit[2]
    res = bool(torch.count_nonzero(t).item() == 0)
    if len(_ZERO_CACHE) > 256:
        _ZERO_CACHE.clear()
    _ZERO_CACHE[id(t)] = (weakref.ref(t), t._version, res)
    return res


_DIAG_WS: dict = {}
_DIAG_BYTES = 8192


def _diag_workspace(dev):
    ws = _DIAG_WS.get(str(dev))
    if ws is None:
        ws = torch.empty(_DIAG_BYTES, dtype=torch.uint8, device=dev)
        _DIAG_WS[str(dev)] = ws
    return ws


def compute_total_momentum(particles, species_config=None):
    """Scalar momentum diagnostic sum(active |v| mass weight) (utils.py:190-202),
    accumulated in FP64 on the GPU; a 0-d float64 CUDA tensor."""
    if species_config is None:
        raise ValueError("species_config is required for the momentum")
    xp, up, act = particle_planes(particles)
    if xp[0].dtype != torch.float32:
        raise ValueError("compute_total_momentum takes float32 particles")
    sp = make_species(species_config, xp[0].shape[1])
    dev = xp[0].device
    ws = _diag_workspace(dev)
    out = torch.empty(1, dtype=torch.float64, device=dev)
    N.call("pic_total_momentum", sp, pic_particles(xp, up, act), N.ptr(out), N.ptr(ws), ws.numel(),
           N.stream_handle())
    return out[0]


def compute_total_charge(rho, static_parameters, dynamic_parameters):
    """Total charge sum_interior(rho) dx dy dz, accumulated in FP64 on the GPU
    (the charge diagnostic kept beside compute_energy); a 0-d float64 CUDA
    tensor.  The reference has no function of this name: its tests sum rho
    the same way (single_particle_pipeline_test.py:429-431)."""
    desc = make_desc(static_parameters, dynamic_parameters)
    rho3 = as_field3(rho, "rho")
    check_fields(desc, rho3, dtype=torch.float32)
    ws = _diag_workspace(rho3.device)
    out = torch.empty(1, dtype=torch.float64, device=rho3.device)
    N.call("pic_total_charge", desc, N.ptr(rho3), N.ptr(out), N.ptr(ws), ws.numel(), N.stream_handle())
    return out[0]
