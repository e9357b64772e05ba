"""Perfectly matched layer (PyPIC3D/boundary_conditions/PML.py:1-359).

Profiles and state keep the reference's structure: ``load_pml_from_toml``
returns ``(active, pml_x, pml_y, pml_z, (sigma_x, sigma_y, sigma_z))`` with
one-ghost profiles of shape ``(Nx+2, Ny+2, Nz+2)``; a PML state is
``(e_memory[6], b_memory[6], profiles[3])`` with interior-shaped memories.
The tiled state (``initialize_tiled_pml_state``) is the single-tile layout of
this library: memories ``(1, 1, 1, Nx, Ny, Nz)`` and profiles in the field
layout ``(1, 1, 1, Nx+2g, Ny+2g, Nz+2g)``.  All arrays are float32 CUDA
tensors.  The stretch itself runs inside the Yee kernels
(``first_order_yee.update_B/update_E`` with a ``pml_state``);
``apply_tiled_pml_to_e_curl`` / ``_b_curl`` call the same kernel code on
given derivative arrays.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from .. import _native as N
from .._tensors import make_desc

PML_WALLS = ["-x", "+x", "-y", "+y", "-z", "+z"]

_AXIS_INDEX = {"x": 0, "y": 1, "z": 2}


def _n(dynamic_parameters):
    return (int(dynamic_parameters.Nx), int(dynamic_parameters.Ny), int(dynamic_parameters.Nz))


def _d(dynamic_parameters):
    return (float(dynamic_parameters.dx), float(dynamic_parameters.dy), float(dynamic_parameters.dz))


def load_pml_from_toml(raw_pml, static_parameters, dynamic_parameters, device="cuda"):
    """Validate the wall list and build the profiles (PML.py:17-81).  Raises
    ValueError with the reference's messages for an unknown or duplicate wall
    and for a thickness that is not positive or exceeds the active cells."""
    raw_layers = [] if raw_pml is None else raw_pml
    raw_layers = [raw_layers] if isinstance(raw_layers, dict) else list(raw_layers)
    n, d = _n(dynamic_parameters), _d(dynamic_parameters)
    layers, seen = [], set()
    for raw in raw_layers:
        wall = raw.get("wall")
        if wall not in PML_WALLS:
            raise ValueError(f"Invalid PML wall: {wall}. Expected one of {PML_WALLS}")
        if wall in seen:
            raise ValueError(f"Duplicate PML wall: {wall}")
        seen.add(wall)
        axis = wall[1]
        thickness = int(raw.get("thickness", 0))
        active_cells = n[_AXIS_INDEX[axis]]
        if thickness <= 0:
            raise ValueError(f"PML thickness for {wall} must be positive")
        if thickness > active_cells:
            raise ValueError(f"PML thickness for {wall} exceeds active cells on {axis}: {thickness} > {active_cells}")
        order = float(raw.get("order", 3.0))
        if "sigma_max" in raw:
            sigma_max = raw["sigma_max"]
        else:
            width = thickness * d[_AXIS_INDEX[axis]]
            sigma_max = -((order + 1.0) * float(dynamic_parameters.C) *
                          math.log(float(raw.get("target_reflection", 1.0e-8)))) / (2.0 * width)
        layers.append((wall, axis, thickness, order, sigma_max))
    pml_x = any(a == "x" for _, a, _, _, _ in layers)
    pml_y = any(a == "y" for _, a, _, _, _ in layers)
    pml_z = any(a == "z" for _, a, _, _, _ in layers)
    return bool(layers), pml_x, pml_y, pml_z, build_pml(dynamic_parameters, tuple(layers), device=device)


def build_pml(dynamic_parameters, pml_layers, device="cuda"):
    """One-ghost damping profiles (PML.py:84-122): layer cell i of thickness
    T (i = 0 at the interface with the physical domain) holds
    sigma_max ((i+1)/T)^order, so the damping grows towards the outer wall."""
    shape = tuple(v + 2 for v in _n(dynamic_parameters))
    prof = [np.zeros(shape) for _ in range(3)]
    for wall, axis, thickness, order, sigma_max in pml_layers:
        a = _AXIS_INDEX[axis]
        T = int(thickness)
        for i in range(T):
            index = T - i if wall[0] == "-" else shape[a] - 1 - T + i
            sl = [slice(None)] * 3
            sl[a] = index
            prof[a][tuple(sl)] = float(sigma_max) * ((i + 1) / T) ** float(order)
    return tuple(torch.tensor(p, dtype=torch.float32, device=device) for p in prof)


def initialize_pml_state(dynamic_parameters, pml_profiles):
    """(e_memory, b_memory, profiles) with six zero (Nx, Ny, Nz) memories each
    (PML.py:125-149); memory order dEz/dy, dEy/dz, dEx/dz, dEz/dx, dEy/dx,
    dEx/dy (B for the E memory)."""
    dev = pml_profiles[0].device
    n = _n(dynamic_parameters)
    e_memory = tuple(torch.zeros(n, dtype=torch.float32, device=dev) for _ in range(6))
    b_memory = tuple(torch.zeros(n, dtype=torch.float32, device=dev) for _ in range(6))
    return e_memory, b_memory, pml_profiles


def _check_single_tile(static_parameters, n):
    tile = tuple(int(w) for w in static_parameters.tile_shape)
    if tile != tuple(n):
        raise NotImplementedError("one tile per GPU: tile_shape must equal (Nx, Ny, Nz)")


def tile_pml_profiles(static_parameters, pml_profiles, tile_shape):
    """Profiles in the single-tile field layout (PML.py:158-223): the one-ghost
    profile placed at offset g-1 of a (1, 1, 1, N+2g) array, zeros elsewhere."""
    n = tuple(int(p) - 2 for p in pml_profiles[0].shape)
    if tuple(int(w) for w in tile_shape) != n:
        raise NotImplementedError("one tile per GPU: tile_shape must equal (Nx, Ny, Nz)")
    g = int(static_parameters.guard_cells)
    out = []
    for p in pml_profiles:
        t = torch.zeros((1, 1, 1) + tuple(v + 2 * g for v in n), dtype=torch.float32, device=p.device)
        t[0, 0, 0, g - 1:g + n[0] + 1, g - 1:g + n[1] + 1, g - 1:g + n[2] + 1] = p.to(torch.float32)
        out.append(t)
    return tuple(out)


def initialize_tiled_pml_state(static_parameters, dynamic_parameters, pml_profiles, tile_shape=None):
    """(e_memory, b_memory, tiled profiles) for one tile (PML.py:226-248):
    memories (1, 1, 1, Nx, Ny, Nz), profiles in the field layout."""
    n = _n(dynamic_parameters)
    _check_single_tile(static_parameters, n)
    dev = pml_profiles[0].device
    e_memory = tuple(torch.zeros((1, 1, 1) + n, dtype=torch.float32, device=dev) for _ in range(6))
    b_memory = tuple(torch.zeros((1, 1, 1) + n, dtype=torch.float32, device=dev) for _ in range(6))
    return e_memory, b_memory, tile_pml_profiles(static_parameters, pml_profiles, n)


def stretch_spatial_derivative(derivative, memory, sigma, dt):
    """(derivative + memory', memory') with memory' = b memory + (b - 1)
    derivative, b = exp(-sigma dt) (PML.py:251-267), elementwise."""
    b = torch.exp(-sigma * dt)
    memory_new = b * memory + (b - 1.0) * derivative
    return derivative + memory_new, memory_new


def _apply(derivatives, static_parameters, dynamic_parameters, pml_state, which):
    from ..solvers.first_order_yee import _pml_args
    desc = make_desc(static_parameters, dynamic_parameters)
    mem3, prof3, clones = _pml_args(desc, pml_state, which)
    n = tuple(int(v) for v in desc.n)
    dv = [t.reshape(n).to(torch.float32).contiguous().clone() for t in derivatives]
    from ..solvers.first_order_yee import pml_struct
    ptrs = (N.c_vp * 6)(*[t.data_ptr() for t in dv])
    N.call("pic_pml_stretch", desc, ptrs, ctypes.byref(pml_struct(mem3, prof3, desc=desc)), N.stream_handle())
    shape = derivatives[0].shape
    dv = [t.reshape(shape) for t in dv]
    curls = (dv[0] - dv[1], dv[2] - dv[3], dv[4] - dv[5])
    if which == 0:
        return curls, (clones, pml_state[1], pml_state[2])
    return curls, (pml_state[0], clones, pml_state[2])


def apply_tiled_pml_to_e_curl(derivatives, static_parameters, dynamic_parameters, pml_state):
    """Stretch (dBz/dy, dBy/dz, dBx/dz, dBz/dx, dBy/dx, dBx/dy) with the E
    memory and assemble curl B (PML.py:270-313).  Returns (curl, pml_state)."""
    return _apply(derivatives, static_parameters, dynamic_parameters, pml_state, 0)


def apply_tiled_pml_to_b_curl(derivatives, static_parameters, dynamic_parameters, pml_state):
    """Stretch (dEz/dy, dEy/dz, dEx/dz, dEz/dx, dEy/dx, dEx/dy) with the B
    memory and assemble curl E (PML.py:316-359).  Returns (curl, pml_state)."""
    return _apply(derivatives, static_parameters, dynamic_parameters, pml_state, 1)
