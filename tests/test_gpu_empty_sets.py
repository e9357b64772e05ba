"""Empty particle sets (no slots, or no species), as the reference's
particle-free runs use them (utils_test.py:286-316, pml_test.py:527-617):
every particle entry point accepts them and does nothing, the deposits give
zero currents and charge, and a step evolves the fields exactly as a step
with the same fields and only inactive particles does."""

import numpy as np
import pytest
import torch

from pypic3d_amd import workloads
from pypic3d_amd._tensors import tiled_particles_from_arrays
from pypic3d_amd.deposition.Esirkepov import Esirkepov_current
from pypic3d_amd.deposition.J_from_rhov import J_from_rhov
from pypic3d_amd.deposition.rho import compute_rho
from pypic3d_amd.evolve import time_loop_electrodynamic
from pypic3d_amd.particles.particle_class import SpeciesConfig, TiledParticles
from pypic3d_amd.particles.particle_tile_communication import (refresh_tiled_particle_tiles,
                                                               update_tiled_particle_positions)
from pypic3d_amd.pusher.particle_push import particle_push
from pypic3d_amd.utils import add_external_fields, compute_energy
from tests.gpu_helpers import gpu_field, oracle_params
from tests.helpers import random_field

pytestmark = pytest.mark.gpu


def _empty(S, P=0):
    z = torch.zeros((1, 1, 1, S, P, 3), dtype=torch.float32, device="cuda")
    parts = TiledParticles(x=z, u=z.clone(), active=torch.zeros((1, 1, 1, S, P), dtype=torch.bool, device="cuda"))
    cfg = SpeciesConfig(charge=np.full(S, -1.0), mass=np.ones(S), weight=np.ones(S),
                        update_x=np.ones((S, 3), bool), update_u=np.ones((S, 3), bool))
    return parts, cfg


def test_energy_of_external_fields_with_no_particles():
    """utils_test.py:286-316: 1^3 grid, eps = 2, mu = 4, C = 10, external
    Ex = 2 and By = 3 on the interior: U_E = 4, U_B = 1.125, kinetic 0."""
    static, dyn = workloads.make_parameters((1, 1, 1), d=(1.0, 1.0, 1.0), dt=0.1, C=10.0, eps=2.0, mu=4.0)
    E, B, *_ = workloads.empty_fields(static, dyn)
    g = static.guard_cells
    ext_E = [torch.zeros_like(c) for c in E]
    ext_B = [torch.zeros_like(c) for c in B]
    ext_E[0][0, 0, 0, g, g, g] = 2.0
    ext_B[1][0, 0, 0, g, g, g] = 3.0
    tE, tB = add_external_fields(E, B, (tuple(ext_E), tuple(ext_B)))
    for S in (1, 0):
        parts, cfg = _empty(S)
        ue, ub, ke = compute_energy(parts, tE, tB, static, dyn, species_config=cfg)
        assert float(ue) == pytest.approx(4.0, rel=1e-12)
        assert float(ub) == pytest.approx(1.125, rel=1e-12)
        assert float(ke) == 0.0


@pytest.mark.parametrize("S", [1, 0])
def test_particle_operators_on_empty_sets(S):
    static, dyn = workloads.make_parameters((8, 6, 4), cfl=0.5)
    p = oracle_params(static, dyn)
    E = tuple(gpu_field(random_field(p, 1 + c, scale=0.1)) for c in range(3))
    B = tuple(gpu_field(random_field(p, 4 + c, scale=0.1)) for c in range(3))
    _, _, J, rho, *_ = workloads.empty_fields(static, dyn)
    J = tuple(torch.full_like(c, 7.0) for c in J)           # templates: the result must not keep these values
    parts, cfg = _empty(S)
    pushed = particle_push(parts, cfg, E, B, static, dyn)
    assert pushed.u.shape == parts.u.shape
    moved = update_tiled_particle_positions(parts, cfg, float(dyn.dt))
    assert moved.x.shape == parts.x.shape
    refreshed, overflow = refresh_tiled_particle_tiles(parts, static, dyn)
    assert refreshed.active.shape == parts.active.shape and not bool(overflow)
    for Jn in (Esirkepov_current(parts, cfg, J, static, dyn), J_from_rhov(parts, cfg, J, static, dyn)):
        for c in Jn:
            assert float(c.abs().max()) == 0.0
    r = compute_rho(parts, cfg, torch.full_like(rho, 3.0), static, dyn)
    assert float(r.abs().max()) == 0.0


@pytest.mark.parametrize("dep", ["esirkepov", "direct"])
def test_particle_free_step_equals_step_with_inactive_particles(dep):
    static, dyn = workloads.make_parameters((10, 8, 6), cfl=0.5, current_deposition=dep)
    p = oracle_params(static, dyn)
    E = tuple(gpu_field(random_field(p, 11 + c, scale=0.1)) for c in range(3))
    B = tuple(gpu_field(random_field(p, 14 + c, scale=0.1)) for c in range(3))
    fields = (E, B) + workloads.empty_fields(static, dyn)[2:]
    parts, cfg = _empty(1)
    x = np.zeros((1, 5, 3), np.float32)
    inactive = tiled_particles_from_arrays(x, x, np.zeros((1, 5), bool))
    _, f_empty = time_loop_electrodynamic(parts, cfg, fields, static, dyn)
    _, f_inact = time_loop_electrodynamic(inactive, cfg, fields, static, dyn)
    for k in (0, 1, 2):
        for a, b in zip(f_empty[k], f_inact[k]):
            assert torch.equal(a, b), k
    assert float(max(c.abs().max() for c in f_empty[2])) == 0.0
    assert float(max(c.abs().max() for c in f_empty[0])) > 0.0


def test_pml_absorbs_the_particle_free_1d_wave():
    """pml_test.py:527-617 on one tile: 40 x 1 x 1, conducting x walls with
    8-cell PMLs (order 3, sigma_max 60), Ey = Bz = exp(-((x + 0.3)/0.04)^2)
    on the vertex line, update_E then update_B per step, field energy from
    compute_energy with no species.  The reference test expects < 0.65 of the
    initial energy after 60 steps, but its update_B advances B by dt/2
    (first_order_yee.py:116), so in this E-then-half-B loop the pulse moves at
    C/sqrt(2) and splits (energy 1.07 e0 at step 60, the oracle's value too);
    it crosses the +x layer by step 100.  Checked: the energy equals the FP64
    oracle's at 60 and 100 steps, and is below 0.65 e0 at 100."""
    from oracle import pic_oracle as O
    from pypic3d_amd.boundary_conditions.PML import initialize_tiled_pml_state, load_pml_from_toml
    from pypic3d_amd.solvers.first_order_yee import update_B, update_E
    walls = [{"wall": "-x", "thickness": 8, "order": 3.0, "sigma_max": 60.0},
             {"wall": "+x", "thickness": 8, "order": 3.0, "sigma_max": 60.0}]
    static, dyn = workloads.make_parameters((40, 1, 1), wind=(1.0, 1.0, 1.0), d=(1.0 / 40, 1.0, 1.0),
                                            dt=0.5 / 40, bc=(1, 0, 0))
    p = oracle_params(static, dyn)
    _, _, _, _, prof = load_pml_from_toml(walls, static, dyn)
    state = initialize_tiled_pml_state(static, dyn, prof)
    E, B, J, *_ = workloads.empty_fields(static, dyn)
    g = static.guard_cells
    xv = -0.5 + (np.arange(40) + 0.5) / 40
    pulse = np.exp(-((xv + 0.30) / 0.04) ** 2).astype(np.float32)
    E[1][0, 0, 0, g:-g, g, g] = torch.as_tensor(pulse, device="cuda")
    B[2][0, 0, 0, g:-g, g, g] = torch.as_tensor(pulse, device="cuda")
    # the same start in the oracle
    Eo = [np.zeros(p.shape) for _ in range(3)]
    Bo = [np.zeros(p.shape) for _ in range(3)]
    Jo = [np.zeros(p.shape) for _ in range(3)]
    Eo[1][g:-g, g, g] = pulse
    Bo[2][g:-g, g, g] = pulse
    oprof = [O.pml_ghosted(pr, p) for pr in O.pml_profiles(O.pml_layers(walls, p), p)]
    me = [np.zeros(p.N) for _ in range(6)]
    mb = [np.zeros(p.N) for _ in range(6)]
    parts, _ = _empty(0)
    cfg = SpeciesConfig(charge=np.zeros(0), mass=np.zeros(0), weight=np.zeros(0),
                        update_x=np.zeros((0, 3), bool), update_u=np.zeros((0, 3), bool))

    def energy():
        return sum(float(v) for v in compute_energy(parts, E, B, static, dyn, species_config=cfg)[:2])

    def oracle_energy():
        A = slice(g, -g)
        return 0.5 * sum(np.sum(c[A, A, A] ** 2) for c in Eo + Bo) * p.d[0]

    e0 = energy()
    assert e0 == pytest.approx(oracle_energy(), rel=1e-6)
    for step in range(1, 101):
        E, state = update_E(E, B, J, static, dyn, state)
        B, state = update_B(E, B, static, dyn, state)
        Eo, me = O.update_E(Eo, Bo, Jo, p, pml=(me, oprof))
        Bo, mb = O.update_B(Eo, Bo, p, pml=(mb, oprof))
        if step in (60, 100):
            e = energy()
            assert np.isfinite(e)
            assert abs(e - oracle_energy()) <= 1e-4 * e0, step
    assert e < 0.65 * e0


@pytest.mark.parametrize("S", [1, 0])
def test_rebin_and_gathered_launches_on_empty_sets(S):
    """The re-bin without its gather and the launches that read through its
    permutation accept an empty set: the ranges are all empty and nothing is
    read or written."""
    from pypic3d_amd import _native as N
    from pypic3d_amd._tensors import make_desc, make_species, particle_planes, pic_particles
    static, dyn = workloads.make_parameters((8, 8, 8), cfl=0.5)
    parts, cfg = _empty(S)
    desc, sp = make_desc(static, dyn), make_species(cfg, 0)
    assert int(N.lib().pic_tile_exact_grid(desc)) == 1
    xp, up, ap = particle_planes(parts)
    pp = pic_particles(xp, up, ap)
    tiles = int(N.lib().pic_tile_count(desc, 4))
    starts = torch.full((S * (tiles + 1) + 1,), -1, dtype=torch.int32, device="cuda")
    ws = torch.empty(max(1, int(N.lib().pic_tile_rebin_workspace_bytes(desc, sp, 4))), dtype=torch.uint8,
                     device="cuda")
    s = N.stream_handle()
    N.call("pic_tile_rebin", desc, sp, pp, 4, 0.0, None, N.ptr(starts), None, N.ptr(ws), ws.numel(), s)
    torch.cuda.synchronize()
    assert bool((starts == 0).all())
    E, B, J, *_ = workloads.empty_fields(static, dyn)
    f3 = lambda t: [c.view(c.shape[-3:]) for c in t]      # noqa: E731
    E3, B3, J3 = f3(E), f3(B), f3(J)
    rare = torch.zeros(1, dtype=torch.int32, device="cuda")
    for c in J3:
        c.fill_(0.0)
    N.call("pic_fused_push_esirkepov_tiled_gather", desc, sp, pp, None, N.p3(E3), N.p3(B3), None, None, pp,
           N.p3(J3), 4, N.ptr(starts), 4, 4, N.ptr(rare), s)
    N.call("pic_fused_push_direct_tiled_gather", desc, sp, pp, None, N.p3(E3), N.p3(B3), None, None, pp,
           N.p3(J3), 4, N.ptr(starts), 4, N.ptr(rare), s)
    torch.cuda.synchronize()
    assert all(float(c.abs().sum()) == 0.0 for c in J3)
