"""Tensor layout helpers and descriptor builders shared by the op modules.

Device layout:
* a field component is a contiguous float32 CUDA tensor of shape
  ``(Nx+2g, Ny+2g, Nz+2g)``; the public API also accepts/returns the
  reference's tiled shape ``(1, 1, 1, Nx+2g, Ny+2g, Nz+2g)`` (a view).
* particle ``x``/``u`` are ``(1, 1, 1, S, P, 3)`` float32 views whose
  component planes ``x[..., a]`` are contiguous ``(S, P)`` arrays (SoA
  storage ``(3, S, P)``); ``active`` is ``(1, 1, 1, S, P)`` bool.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _native as N
from .utilities.grids import tiled_origin_spacing


def _f(v):
    if isinstance(v, torch.Tensor):
        return float(v.detach().cpu().item())
    return float(np.asarray(v))


def _i(v):
    if isinstance(v, torch.Tensor):
        return int(v.detach().cpu().item())
    return int(np.asarray(v))


_DESC_CACHE: dict = {}


def make_desc(static, dynamic) -> N.PicDesc:
    """PicDesc from (StaticParameters, DynamicParameters); cached per object pair."""
    key = (id(static), id(dynamic))
    hit = _DESC_CACHE.get(key)
    if hit is not None and hit[0] is static and hit[1] is dynamic:
        return hit[2]
    d = N.PicDesc()
    n = (_i(dynamic.Nx), _i(dynamic.Ny), _i(dynamic.Nz))
    g = int(static.guard_cells)
    if tuple(int(w) for w in static.tile_shape) != n:
        raise ValueError(
            "one tile per GPU: tile_shape must equal (Nx, Ny, Nz) "
            f"(got {tuple(static.tile_shape)} for grid {n})")
    for a in range(3):
        d.n[a] = n[a]
        d.dims[a] = n[a] + 2 * g
    d.g = g
    d.shape_factor = int(static.shape_factor)
    d.relativistic = 1 if static.relativistic else 0
    if static.particle_pusher not in N.PUSHERS:
        raise ValueError(f"Unknown particle_pusher: {static.particle_pusher}")
    d.pusher = N.PUSHERS[static.particle_pusher]
    for a in range(3):
        d.bc[a] = int(static.boundary_conditions[a])
        d.pbc[a] = int(static.particle_boundary_conditions[a])
    dd = (_f(dynamic.dx), _f(dynamic.dy), _f(dynamic.dz))
    wind = (_f(dynamic.x_wind), _f(dynamic.y_wind), _f(dynamic.z_wind))
    lines_c = dynamic.grids.tiled_center_grid if dynamic.grids is not None else ()
    lines_v = dynamic.grids.tiled_vertex_grid if dynamic.grids is not None else ()
    for a in range(3):
        d.d[a] = dd[a]
        d.wind[a] = wind[a]
        if lines_c and lines_v and len(lines_c) == 3:
            lc = np.asarray(lines_c[a], dtype=np.float64).reshape(-1, n[a] + 2 * g)[0]
            lv = np.asarray(lines_v[a], dtype=np.float64).reshape(-1, n[a] + 2 * g)[0]
            d.org_c[a], d.sp_c[a] = float(lc[0]), float(lc[1] - lc[0])
            d.org_v[a], d.sp_v[a] = float(lv[0]), float(lv[1] - lv[0])
        else:
            electro = bool(static.electrostatic)
            d.org_c[a], d.sp_c[a] = tiled_origin_spacing(wind[a], dd[a], n[a], g, "center", electro)
            d.org_v[a], d.sp_v[a] = tiled_origin_spacing(wind[a], dd[a], n[a], g, "vertex", electro)
    d.dt = _f(dynamic.dt)
    d.C = _f(dynamic.C)
    d.eps = _f(dynamic.eps)
    d.mu = _f(dynamic.mu)
    d.alpha = _f(dynamic.alpha)
    if len(_DESC_CACHE) > 64:
        _DESC_CACHE.clear()
    _DESC_CACHE[key] = (static, dynamic, d)
    return d


def shape_desc(shape3, g, dt=0.0) -> N.PicDesc:
    """Minimal descriptor for shape-only kernels (filters, position advance)."""
    d = N.PicDesc()
    for a in range(3):
        d.dims[a] = int(shape3[a])
        d.n[a] = int(shape3[a]) - 2 * g
        d.d[a] = 1.0
        d.wind[a] = float(d.n[a])
        d.sp_c[a] = d.sp_v[a] = 1.0
    if min(d.n) < 1:
        raise ValueError(f"field of shape {tuple(shape3)} has no interior with {g} guard cells")
    d.g = g
    d.shape_factor = 1
    d.dt = dt
    d.C = d.eps = d.mu = d.alpha = 1.0
    return d


_SPECIES_CACHE: dict = {}


def make_species(species_config, slots: int) -> N.PicSpecies:
    key = (id(species_config), slots)
    hit = _SPECIES_CACHE.get(key)
    if hit is not None and hit[0] is species_config:
        return hit[1]
    sp = N.PicSpecies()
    charge = np.asarray(_np(species_config.charge), dtype=np.float64).reshape(-1)
    S = charge.shape[0]
    if S > N.MAX_SPECIES:
        raise ValueError(f"species count {S} above {N.MAX_SPECIES}")
    mass = np.asarray(_np(species_config.mass), dtype=np.float64).reshape(-1)
    weight = np.asarray(_np(species_config.weight), dtype=np.float64).reshape(-1)
    ux = np.asarray(_np(species_config.update_x), dtype=bool).reshape(S, 3)
    uu = np.asarray(_np(species_config.update_u), dtype=bool).reshape(S, 3)
    sp.count = S
    sp.slots = int(slots)
    for s in range(S):
        sp.charge[s] = charge[s]
        sp.mass[s] = mass[s]
        sp.weight[s] = weight[s]
        for a in range(3):
            sp.update_x[s][a] = int(ux[s, a])
            sp.update_u[s][a] = int(uu[s, a])
    if len(_SPECIES_CACHE) > 64:
        _SPECIES_CACHE.clear()
    _SPECIES_CACHE[key] = (species_config, sp)
    return sp


def _np(v):
    if isinstance(v, torch.Tensor):
        return v.detach().cpu().numpy()
    return np.asarray(v)


# ---------------------------------------------------------------- fields

def is_f64(t) -> bool:
    return isinstance(t, torch.Tensor) and t.dtype == torch.float64


def as_field3(t, name="field"):
    """3-D CUDA view of a field component (accepts the (1,1,1,...) tile
    shape): float64 stays float64 (the FP64 path), anything else becomes
    float32."""
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch tensor on the GPU")
    if t.dim() == 6:
        if tuple(t.shape[:3]) != (1, 1, 1):
            raise ValueError(f"{name}: one tile per GPU, got tile grid {tuple(t.shape[:3])}")
        t = t[0, 0, 0]
    if t.dim() != 3:
        raise ValueError(f"{name}: expected a 3-D field, got shape {tuple(t.shape)}")
    if not t.is_cuda:
        raise ValueError(f"{name}: tensor must live on the GPU")
    want = torch.float64 if t.dtype == torch.float64 else torch.float32
    if t.dtype != want or not t.is_contiguous():
        t = t.to(want).contiguous()
    return t


def check_fields(desc, *fields, dtype=None):
    """Every field handed to a kernel must have the descriptor's extents and
    one common dtype (float32, or float64 for the FP64 path)."""
    dims = (desc.dims[0], desc.dims[1], desc.dims[2])
    for f in fields:
        if f is None:
            continue
        if dtype is None:
            dtype = f.dtype
        if (tuple(f.shape) != dims or f.dtype != dtype or f.dtype not in (torch.float32, torch.float64)
                or not f.is_contiguous()):
            raise ValueError(f"field of shape {tuple(f.shape)} ({f.dtype}) does not match the grid {dims} {dtype}")


def like_input(out3, template):
    return out3.view(1, 1, 1, *out3.shape) if template.dim() == 6 else out3


def new_field(shape3, device):
    return torch.empty(tuple(shape3), dtype=torch.float32, device=device)


# ------------------------------------------------------------- particles

def _planes(t, name):
    """Three contiguous (S, P) component planes of a (..., S, P, 3) tensor."""
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise ValueError(f"particles.{name} must be a CUDA tensor")
    if t.dim() == 6:
        if tuple(t.shape[:3]) != (1, 1, 1):
            raise ValueError(f"particles.{name}: one tile per GPU, got {tuple(t.shape[:3])}")
        t = t[0, 0, 0]
    if t.dim() != 3 or t.shape[-1] != 3:
        raise ValueError(f"particles.{name}: expected (S, P, 3), got {tuple(t.shape)}")
    if t.dtype not in (torch.float32, torch.float64):
        t = t.to(torch.float32)
    planes = [t[..., a] for a in range(3)]
    if not all(p.is_contiguous() for p in planes):
        soa = t.permute(2, 0, 1).contiguous()
        planes = [soa[a] for a in range(3)]
    return planes


def particle_planes(particles):
    xp = _planes(particles.x, "x")
    up = _planes(particles.u, "u")
    act = particles.active
    if act.dim() == 5:
        act = act[0, 0, 0]
    if act.dtype != torch.bool or not act.is_contiguous():
        act = act.to(torch.bool).contiguous()
    S, P = xp[0].shape
    if tuple(up[0].shape) != (S, P) or tuple(act.shape) != (S, P):
        raise ValueError("particles: x, u and active disagree on (species, slots)")
    if xp[0].dtype != up[0].dtype:
        raise ValueError("particles: x and u must share one dtype")
    return xp, up, act


def pic_particles(xp, up, act) -> N.PicParticles:
    pp = N.PicParticles()
    for a in range(3):
        pp.x[a] = xp[a].data_ptr()
        pp.u[a] = up[a].data_ptr()
    pp.active = act.data_ptr()
    return pp


def new_soa(S, P, device, dtype=torch.float32):
    """(3, S, P) storage and its (1, 1, 1, S, P, 3) view."""
    store = torch.empty((3, S, P), dtype=dtype, device=device)
    return store, soa_view(store)


def soa_view(store):
    return store.permute(1, 2, 0).unsqueeze(0).unsqueeze(0).unsqueeze(0)


def tiled_particles_from_arrays(x, u, active, device="cuda", dtype=torch.float32):
    """Build TiledParticles on the GPU from (S, P, 3) arrays (NumPy or torch);
    dtype float32, or float64 for the FP64 path."""
    from .particles.particle_class import TiledParticles
    x = torch.as_tensor(np.asarray(x) if not isinstance(x, torch.Tensor) else x)
    u = torch.as_tensor(np.asarray(u) if not isinstance(u, torch.Tensor) else u)
    a = torch.as_tensor(np.asarray(active) if not isinstance(active, torch.Tensor) else active)
    if x.dim() == 6:
        x, u, a = x[0, 0, 0], u[0, 0, 0], a[0, 0, 0]
    xs = x.permute(2, 0, 1).to(device=device, dtype=dtype).contiguous()
    us = u.permute(2, 0, 1).to(device=device, dtype=dtype).contiguous()
    act = a.to(device=device, dtype=torch.bool).contiguous()
    return TiledParticles(x=soa_view(xs), u=soa_view(us), active=act.unsqueeze(0).unsqueeze(0).unsqueeze(0))


def ceil_div(a, b):
    return (a + b - 1) // b


__all__ = ["make_desc", "shape_desc", "make_species", "as_field3", "is_f64", "like_input", "new_field",
           "particle_planes", "pic_particles", "new_soa", "soa_view", "tiled_particles_from_arrays"]
