"""Cell sort of the particle slots.

The key is the reference's flat cell index (ix*Ny + iy)*Nz + iz with
ix = clip(floor((x + W/2)/dx), 0, Nx-1) (particle_initialization.py:89-102,
kernel_fixtures.py:554-575), offset by species so every species keeps its own
slot segment; inactive slots sort to the end of their segment.  The
permutation equals ``np.argsort(key, kind="stable")`` (the reference's packing
order, particle_initialization.py:159).  Sorting is not part of the
reference's one-device step (its slots never move there); callers sort every
k steps to keep the deposit and gather cache-local.
"""

from __future__ import annotations

import math

import torch

from . import _native as N
from ._tensors import make_desc, make_species, new_soa, particle_planes, pic_particles
from .particles.particle_class import TiledParticles

_WS: dict = {}


def cell_keys(particles, species_config, static_parameters, dynamic_parameters):
    """int32 keys of shape (S, P)."""
    desc = make_desc(static_parameters, dynamic_parameters)
    xp, up, act = particle_planes(particles)
    if xp[0].dtype != torch.float32:
        raise ValueError("cell keys and the particle sort take float32 particles")
    S, P = xp[0].shape
    sp = make_species(species_config, P)
    keys = torch.empty((S, P), dtype=torch.int32, device=xp[0].device)
    N.call("pic_cell_keys", desc, sp, pic_particles(xp, up, act), N.ptr(keys), N.stream_handle())
    return keys


def _workspace(n, device):
    nbytes = int(N.lib().pic_sort_workspace_bytes(int(n)))
    key = (str(device), n)
    ws = _WS.get(key)
    if ws is None:
        _WS.clear()
        ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=device)
        _WS[key] = ws
    return ws


def key_bits(S, n_cells):
    return max(1, math.ceil(math.log2(S * (n_cells + 1) + 1)))


def sort_permutation(keys, n_cells, S):
    """Stable argsort of int32 keys (flattened) on the GPU; returns (perm, sorted_keys)."""
    flat = keys.reshape(-1).contiguous()
    n = flat.numel()
    perm = torch.empty(n, dtype=torch.int32, device=flat.device)
    sk = torch.empty(n, dtype=torch.int32, device=flat.device)
    ws = _workspace(n, flat.device)
    N.call("pic_sort_permutation", N.ptr(flat), N.ptr(sk), N.ptr(perm), n, key_bits(S, n_cells), N.ptr(ws),
           ws.numel(), N.stream_handle())
    return perm, sk


def sort_particles_by_cell(particles, species_config, static_parameters, dynamic_parameters):
    """Returns (sorted TiledParticles, permutation over the flattened (S*P) slots)."""
    desc = make_desc(static_parameters, dynamic_parameters)
    keys = cell_keys(particles, species_config, static_parameters, dynamic_parameters)
    xp, up, act = particle_planes(particles)
    S, P = xp[0].shape
    n_cells = desc.n[0] * desc.n[1] * desc.n[2]
    perm, _ = sort_permutation(keys, n_cells, S)
    xs, xv = new_soa(S, P, xp[0].device)
    us, uv = new_soa(S, P, xp[0].device)
    act_o = torch.empty_like(act)
    N.call("pic_permute_particles", pic_particles(xp, up, act),
           pic_particles([xs[0], xs[1], xs[2]], [us[0], us[1], us[2]], act_o), N.ptr(perm), S * P,
           N.stream_handle())
    return TiledParticles(x=xv, u=uv, active=act_o.view(1, 1, 1, S, P)), perm
