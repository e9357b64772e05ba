"""Parameter builders and seeded synthetic plasmas for the benchmark workloads.

``make_parameters`` builds StaticParameters/DynamicParameters the way
PyPIC3D's kernel fixtures do (tests/kernel_fixtures.py:29-153) for one tile
covering the domain.  The plasma generators draw positions uniformly inside
each cell and Maxwellian velocities with a seeded torch generator on the GPU,
emitting particles in cell order (already cell-sorted).
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from .parameters import DynamicParameters, GridParameters, StaticParameters
from .particles.particle_class import SpeciesConfig, TiledParticles
from ._tensors import soa_view
from .utilities.grids import build_collocated_grid, build_tiled_yee_grids, build_yee_grid


def make_parameters(N, wind=None, d=None, dt=None, cfl=0.5, g=2, shape_factor=1, relativistic=True,
                    current_deposition="esirkepov", current_filter="none", C=1.0, eps=1.0, mu=1.0, kb=1.0,
                    alpha=1.0, bc=(0, 0, 0), pbc=(0, 0, 0), electrostatic=False, particle_pusher="boris", name="run"):
    N = tuple(int(n) for n in N)
    if d is None:
        d = tuple(1.0 for _ in N) if wind is None else tuple(float(w) / n for w, n in zip(wind, N))
    d = tuple(float(v) for v in d)
    if wind is None:
        wind = tuple(n * dd for n, dd in zip(N, d))
    wind = tuple(float(w) for w in wind)
    if dt is None:
        inv = sum(1.0 / dd for dd, n in zip(d, N) if n > 1)
        dt = cfl / (C * inv)
    static = StaticParameters(
        name=name, output_dir=".", Nt=0, verbose=False, GPUs=True, benchmark=False,
        solver="electrostatic" if electrostatic else "electrodynamic_yee", electrostatic=bool(electrostatic),
        relativistic=bool(relativistic), particle_pusher=particle_pusher, current_deposition=current_deposition,
        current_filter=current_filter, shape_factor=int(shape_factor), guard_cells=int(g), tile_shape=N,
        particle_tile_capacity_factor=1.0, pml_active=False, boundary_conditions=tuple(int(b) for b in bc),
        particle_boundary_conditions=tuple(int(b) for b in pbc), field_mesh=None,
    )
    dyn = DynamicParameters(
        dt=float(dt), dx=d[0], dy=d[1], dz=d[2], Nx=N[0], Ny=N[1], Nz=N[2],
        x_wind=wind[0], y_wind=wind[1], z_wind=wind[2], C=float(C), eps=float(eps), mu=float(mu), kb=float(kb),
        alpha=float(alpha), grids=GridParameters((), (), (), ()),
    )
    center, vertex = build_collocated_grid(dyn) if electrostatic else build_yee_grid(dyn)
    dyn = dyn._replace(grids=GridParameters(vertex=vertex, center=center, tiled_vertex_grid=(), tiled_center_grid=()))
    tc, tv = build_tiled_yee_grids(static, dyn)
    dyn = dyn._replace(grids=GridParameters(vertex=vertex, center=center, tiled_vertex_grid=tv, tiled_center_grid=tc))
    return static, dyn


def empty_fields(static, dynamic, device="cuda", dtype=torch.float32):
    """Zero (E, B, J, rho, phi, external_fields, None, overflow) in the tiled shape."""
    g = int(static.guard_cells)
    shape = (1, 1, 1, int(dynamic.Nx) + 2 * g, int(dynamic.Ny) + 2 * g, int(dynamic.Nz) + 2 * g)
    z = lambda: torch.zeros(shape, dtype=dtype, device=device)   # noqa: E731
    E = (z(), z(), z())
    B = (z(), z(), z())
    J = (z(), z(), z())
    ext = ((z(), z(), z()), (z(), z(), z()))
    return (E, B, J, z(), z(), ext, None, torch.zeros((), dtype=torch.bool, device=device))


@dataclass
class SpeciesSpec:
    ppc: int
    charge: float = -1.0
    mass: float = 1.0
    weight: float = 1.0
    vth: tuple = (0.01, 0.01, 0.01)
    drift: tuple = (0.0, 0.0, 0.0)
    update_x: bool = True
    update_u: bool = True
    perturb: float = 0.0      # sinusoidal density perturbation amplitude along x
    mode: int = 1


def plasma(static, dynamic, species, seed=0, device="cuda", dtype=torch.float32):
    """Cell-ordered particles: `ppc` per cell per species, uniform in the cell.
    A density perturbation (1 + A cos(2 pi m x / W)) along x is applied by the
    inverse-CDF map of each particle's x inside the domain.  dtype float32,
    or float64 for the FP64 path (the same draws, not rounded to FP32)."""
    N = (int(dynamic.Nx), int(dynamic.Ny), int(dynamic.Nz))
    d = (float(dynamic.dx), float(dynamic.dy), float(dynamic.dz))
    wind = (float(dynamic.x_wind), float(dynamic.y_wind), float(dynamic.z_wind))
    ncell = N[0] * N[1] * N[2]
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed))
    S = len(species)
    P = max(sp.ppc for sp in species) * ncell
    xs = torch.zeros((3, S, P), dtype=dtype, device=device)
    us = torch.zeros((3, S, P), dtype=dtype, device=device)
    act = torch.zeros((S, P), dtype=torch.bool, device=device)
    cells = torch.arange(ncell, device=device, dtype=torch.int64)
    ci = (cells // (N[1] * N[2]), (cells // N[2]) % N[1], cells % N[2])
    for s, sp in enumerate(species):
        n = sp.ppc * ncell
        for a in range(3):
            base = ci[a].repeat_interleave(sp.ppc).to(torch.float64)
            frac = torch.rand(n, generator=gen, device=device, dtype=torch.float64)
            pos = -wind[a] / 2 + (base + frac) * d[a]
            if a == 0 and sp.perturb != 0.0:
                pos = _perturbed(pos, wind[0], sp.perturb, sp.mode)
            if N[a] == 1:
                pos = (frac - 0.5) * d[a] * 0.8
            xs[a, s, :n] = pos.to(dtype)
            us[a, s, :n] = (sp.drift[a] + sp.vth[a] * torch.randn(n, generator=gen, device=device,
                                                                    dtype=torch.float64)).to(dtype)
        act[s, :n] = True
    cfg = SpeciesConfig(
        charge=np.array([sp.charge for sp in species], dtype=np.float64),
        mass=np.array([sp.mass for sp in species], dtype=np.float64),
        weight=np.array([sp.weight for sp in species], dtype=np.float64),
        update_x=np.array([[sp.update_x] * 3 for sp in species], dtype=bool),
        update_u=np.array([[sp.update_u] * 3 for sp in species], dtype=bool),
    )
    parts = TiledParticles(x=soa_view(xs), u=soa_view(us), active=act.view(1, 1, 1, S, P))
    return parts, cfg


def _perturbed(x, W, A, m):
    """Map uniform x in [-W/2, W/2) to density 1 + A cos(k (x + W/2)) by a few
    Newton steps on the CDF (positions stay in the domain)."""
    k = 2 * math.pi * m / W
    u = x + W / 2
    y = u.clone()
    for _ in range(6):
        F = y + (A / k) * torch.sin(k * y) - u
        y = y - F / (1 + A * torch.cos(k * y))
    y = torch.clamp(y, 0.0, W * (1 - 1e-12))
    return y - W / 2


def plasma_frequency_weight(ppc, dt, omega_dt=0.1, charge=-1.0, mass=1.0, eps=1.0, dV=1.0):
    """Macro-particle weight giving omega_p * dt = omega_dt."""
    wp = omega_dt / dt
    n = wp * wp * eps * mass / (charge * charge)
    return n * dV / ppc


# ------------------------------------------------------------------ configs

def config_two_stream_1d(device="cuda", seed=20261202, ppc=32):
    """64 x 1 x 1 periodic two-stream: beams at +-0.1c, vth 0.01c, 1e-3 mode-1
    perturbation, dx = Debye length of the beams, cfl 0.5."""
    vth = 0.01
    static, dyn = make_parameters((64, 1, 1), d=(1.0, 1.0, 1.0), cfl=0.5)
    # dx = Debye length  =>  omega_p = vth / dx
    w = plasma_frequency_weight(2 * ppc, dyn.dt, omega_dt=vth / 1.0 * dyn.dt)
    beams = [SpeciesSpec(ppc, weight=w, vth=(vth,) * 3, drift=(0.1, 0, 0), perturb=1e-3),
             SpeciesSpec(ppc, weight=w, vth=(vth,) * 3, drift=(-0.1, 0, 0), perturb=1e-3)]
    return static, dyn, beams


def config_weibel_3d(n=128, ppc=16, device="cuda"):
    """n^3 periodic Weibel: two electron populations, vth_x 0.1c, vth_y,z 0.01c."""
    static, dyn = make_parameters((n, n, n), cfl=0.5)
    w = plasma_frequency_weight(2 * ppc, dyn.dt, 0.1)
    sp = [SpeciesSpec(ppc, weight=w, vth=(0.1, 0.01, 0.01)), SpeciesSpec(ppc, weight=w, vth=(0.1, 0.01, 0.01))]
    return static, dyn, sp


def config_warm_3d(n=256, ppc=32, vth=0.05, current_deposition="esirkepov"):
    """n^3 periodic warm isotropic electron plasma (omega_p dt = 0.1)."""
    static, dyn = make_parameters((n, n, n), cfl=0.5, current_deposition=current_deposition)
    w = plasma_frequency_weight(ppc, dyn.dt, 0.1)
    return static, dyn, [SpeciesSpec(ppc, weight=w, vth=(vth,) * 3)]


def config_two_stream_2d(n=64, ppc=32):
    """n x n x 1 periodic two-stream (the differentiable case's grid)."""
    static, dyn = make_parameters((n, n, 1), cfl=0.5)
    w = plasma_frequency_weight(2 * ppc, dyn.dt, 0.1)
    return static, dyn, [SpeciesSpec(ppc, weight=w, vth=(0.01,) * 3, drift=(0.1, 0, 0), perturb=1e-3),
                         SpeciesSpec(ppc, weight=w, vth=(0.01,) * 3, drift=(-0.1, 0, 0), perturb=1e-3)]


def config_two_stream_ensemble(n_members=256, seed=20261202, device="cuda", ppc_range=(16, 128),
                               nx_choices=(64, 128, 256, 512), ny_choices=(1, 32), drift_range=(0.05, 0.2),
                               vth=0.01):
    """Independent 1-D / 2-D two-stream members of uneven size: Nx from
    nx_choices, Ny from ny_choices (Nz = 1), particles per cell per beam
    uniform in ppc_range, drift uniform in drift_range, thermal spread vth,
    dx = 1, cfl 0.5, omega_p dt = 0.1.  Returns a list of
    (particles, species_config, fields, static, dynamic)."""
    import random
    rng = random.Random(int(seed))
    members = []
    for m in range(n_members):
        nx = rng.choice(nx_choices)
        ny = rng.choice(ny_choices)
        ppc = rng.randint(*ppc_range)
        drift = rng.uniform(*drift_range)
        static, dyn = make_parameters((nx, ny, 1), cfl=0.5, name=f"member{m}")
        w = plasma_frequency_weight(2 * ppc, dyn.dt, 0.1)
        specs = [SpeciesSpec(ppc, weight=w, vth=(vth,) * 3, drift=(drift, 0, 0), perturb=1e-3),
                 SpeciesSpec(ppc, weight=w, vth=(vth,) * 3, drift=(-drift, 0, 0), perturb=1e-3)]
        parts, cfg = plasma(static, dyn, specs, seed=rng.randrange(1 << 30), device=device)
        members.append((parts, cfg, empty_fields(static, dyn, device=device), static, dyn))
    return members
