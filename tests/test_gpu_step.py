"""GPU parity of whole steps (evolve.py signatures) against the FP64 oracle, the
reference's single-particle known answers, and size-independent properties at
larger sizes (discrete continuity, total current)."""

import numpy as np
import pytest
import torch

from oracle import pic_oracle as O
from pypic3d_amd import workloads
from pypic3d_amd._tensors import tiled_particles_from_arrays
from pypic3d_amd.deposition.Esirkepov import Esirkepov_current
from pypic3d_amd.deposition.rho import compute_rho
from pypic3d_amd.evolve import time_loop_electrodynamic, time_loop_electrostatic
from pypic3d_amd.particles.particle_class import SpeciesConfig
from pypic3d_amd.particles.particle_tile_communication import (refresh_tiled_particle_tiles,
                                                               update_tiled_particle_positions)
from tests.gpu_helpers import TOL, gpu_field, host, host_particles, oracle_params, rel_err
from tests.helpers import random_field, random_particles, species

pytestmark = pytest.mark.gpu


def _cfg(sp):
    return SpeciesConfig(sp.charge, sp.mass, sp.weight, sp.update_x, sp.update_u)


def _fields(p, seed, with_external=False):
    E = [random_field(p, seed + c, scale=0.2) for c in range(3)]
    B = [random_field(p, seed + 3 + c, scale=0.2) for c in range(3)]
    z = np.zeros(p.shape)
    Ee = [random_field(p, seed + 6 + c, scale=0.1) for c in range(3)] if with_external else [z, z, z]
    Be = [random_field(p, seed + 9 + c, scale=0.1) for c in range(3)] if with_external else [z, z, z]
    gpu = (tuple(map(gpu_field, E)), tuple(map(gpu_field, B)), tuple(gpu_field(z) for _ in range(3)),
           gpu_field(z), gpu_field(z), (tuple(map(gpu_field, Ee)), tuple(map(gpu_field, Be))), None,
           torch.zeros((), dtype=torch.bool, device="cuda"))
    st = O.State(None, None, None, E, B, E_ext=Ee if with_external else None, B_ext=Be if with_external else None)
    return gpu, st


@pytest.mark.parametrize("dep", ["esirkepov", "direct"])
@pytest.mark.parametrize("N", [(16, 1, 1), (12, 10, 1), (10, 8, 6)])
@pytest.mark.parametrize("ext", [False, True])
def test_electrodynamic_step_matches_oracle(dep, N, ext):
    static, dyn = workloads.make_parameters(N, wind=tuple(0.5 * n for n in N), dt=0.05, current_deposition=dep,
                                            current_filter="bilinear" if dep == "direct" else "none",
                                            shape_factor=1 if dep == "esirkepov" else 2)
    p = oracle_params(static, dyn)
    x, u, act = random_particles(p, S=2, P=256, seed=21, vmax=0.03, inactive_frac=0.1)
    sp = species(2)
    gpu, st = _fields(p, 100, ext)
    st.x, st.u, st.active = x, u, act
    parts, fields = time_loop_electrodynamic(tiled_particles_from_arrays(x, u, act), _cfg(sp), gpu, static, dyn)
    want = O.time_loop_electrodynamic(st, sp, p)
    xg, ug, ag = host_particles(parts)
    np.testing.assert_array_equal(ag, want.active)
    assert np.max(np.abs(xg - want.x)) <= 1e-6 * max(p.wind)
    assert rel_err(ug, want.u) <= TOL
    for c in range(3):
        assert rel_err(host(fields[2][c]), want.J[c]) <= TOL, ("J", c)
        assert rel_err(host(fields[0][c]), want.E[c]) <= TOL, ("E", c)
        assert rel_err(host(fields[1][c]), want.B[c]) <= TOL, ("B", c)
    assert fields[6] is None and not bool(fields[7])


def test_electrostatic_step_matches_oracle():
    static, dyn = workloads.make_parameters((16, 12, 10), wind=(8.0, 6.0, 5.0), dt=0.05, electrostatic=True,
                                            current_deposition="direct")
    p = oracle_params(static, dyn)
    x, u, act = random_particles(p, S=2, P=800, seed=4, vmax=0.03)
    sp = species(2, charge=(-1.0, 1.0), mass=(1.0, 100.0))
    gpu, st = _fields(p, 300)
    st.x, st.u, st.active = x, u, act
    st.rho, st.phi = np.zeros(p.shape), np.zeros(p.shape)
    parts, fields = time_loop_electrostatic(tiled_particles_from_arrays(x, u, act), _cfg(sp), gpu, static, dyn)
    want = O.time_loop_electrostatic(st, sp, p, tol=1e-13, max_iter=5000)
    xg, ug, ag = host_particles(parts)
    assert np.max(np.abs(xg - want.x)) <= 1e-6 * max(p.wind)
    assert rel_err(ug, want.u) <= TOL
    assert rel_err(host(fields[3]), want.rho) <= TOL
    # the CG refines phi in FP64 to the reference's absolute tolerance
    phi_g, phi_w = host(fields[4]), want.phi
    assert rel_err(phi_g - phi_g[2:-2, 2:-2, 2:-2].mean(), phi_w - phi_w[2:-2, 2:-2, 2:-2].mean()) <= TOL
    for c in range(3):
        assert rel_err(host(fields[0][c]), want.E[c]) <= TOL


@pytest.mark.parametrize("dep", ["esirkepov", "direct"])
def test_single_particle_pipeline_known_answers(dep):
    # single_particle_pipeline_test.py:329-518: one step from zero fields gives
    # E = -dt J / eps and B_z from E_y (with the dt/2 half step).
    static, dyn = workloads.make_parameters((8, 1, 1), wind=(4.0, 1.0, 1.0), dt=0.05, relativistic=False,
                                            current_deposition=dep, shape_factor=1 if dep == "esirkepov" else 2)
    p = oracle_params(static, dyn)
    x = np.array([[[-1.32, 0.0, 0.0]]])
    u = np.array([[[0.08, 0.2, 0.0]]])
    sp = species(1, charge=(-1.0,), mass=(1.0,), weight=(0.5,))
    gpu, _ = _fields(p, 0)
    zero = tuple(gpu_field(np.zeros(p.shape)) for _ in range(3))
    gpu = (zero, tuple(gpu_field(np.zeros(p.shape)) for _ in range(3))) + gpu[2:]
    parts, fields = time_loop_electrodynamic(tiled_particles_from_arrays(x, u, np.ones((1, 1), bool)), _cfg(sp), gpu,
                                             static, dyn)
    E = [host(c) for c in fields[0]]
    J = [host(c) for c in fields[2]]
    A = (slice(2, -2),) * 3
    for c in range(3):
        assert np.max(np.abs(E[c][A] + p.dt * J[c][A] / p.eps)) <= 1e-6 * max(1e-30, np.max(np.abs(J[c])))
    Ey = E[1]
    Bz_want = -(p.dt / 2) * (Ey[3:-1, 2:-2, 2:-2] - Ey[2:-2, 2:-2, 2:-2]) / p.d[0]
    assert rel_err(host(fields[1][2])[2:-2, 2:-2, 2:-2], Bz_want) <= TOL
    if dep == "esirkepov":
        xg, _, _ = host_particles(parts)
        assert abs(xg[0, 0, 0] - (-1.32 + 0.08 * 0.05)) <= 1e-6


@pytest.mark.parametrize("N,ppc", [((48, 40, 32), 8), ((64, 64, 1), 32), ((256, 1, 1), 64)])
def test_esirkepov_continuity_and_total_current_at_size(N, ppc):
    """Size-independent properties: discrete continuity (rho_new - rho_old)/dt +
    div J at FP32 round-off of the stored positions, and sum(J) dV = sum q w v."""
    static, dyn = workloads.make_parameters(N, cfl=0.5)
    w = 0.01
    parts, cfg = workloads.plasma(static, dyn, [workloads.SpeciesSpec(ppc, weight=w, vth=(0.05,) * 3)], seed=3)
    J0 = tuple(torch.zeros((1, 1, 1) + tuple(n + 4 for n in N), device="cuda") for _ in range(3))
    rho_old = compute_rho(parts, cfg, J0[0], static, dyn)
    J = Esirkepov_current(parts, cfg, J0, static, dyn)
    moved = update_tiled_particle_positions(parts, cfg, dyn.dt)
    moved, _ = refresh_tiled_particle_tiles(moved, static, dyn)
    rho_new = compute_rho(moved, cfg, J0[0], static, dyn)
    A = (slice(2, -2),) * 3
    K = slice(1, -3)
    drho = (rho_new[0, 0, 0][A] - rho_old[0, 0, 0][A]).double() / dyn.dt
    div = torch.zeros_like(drho)
    Jd = [j[0, 0, 0].double() for j in J]
    if N[0] > 1:
        div += Jd[0][A] - Jd[0][K, 2:-2, 2:-2]
    if N[1] > 1:
        div += Jd[1][A] - Jd[1][2:-2, K, 2:-2]
    if N[2] > 1:
        div += Jd[2][A] - Jd[2][2:-2, 2:-2, K]
    resid = float((drho + div).abs().max())
    scale = float(drho.abs().max() + div.abs().max())
    # bound: each stored position carries <= ulp(|x|)/2 of rounding; rho moves by
    # q w / dV * ulp / d per particle, summed over the <= 2^dim cells it touches
    ulp_max = float(np.spacing(np.float32(max(N) / 2)))
    bound = ppc * 8 * w * ulp_max / dyn.dt
    assert resid <= max(bound, 1e-6 * scale), (resid, bound, scale)
    # total current: sum J dV = sum_p q w (x_new - x_old)/dt = sum q w v (Esirkepov)
    v = parts.u[0, 0, 0].double()
    act = parts.active[0, 0, 0]
    for a in range(3):
        if N[a] > 1:
            tot = (-1.0 * w * v[..., a][act]).sum().item()
            scale = w * v[..., a][act].abs().sum().item()
            assert abs(Jd[a][A].sum().item() - tot) <= 1e-5 * scale


@pytest.mark.parametrize("N,ppc,pusher", [((12, 10, 8), 24, "boris"), ((9, 7, 1), 40, "boris"),
                                          ((6, 5, 4), 70, "boris"), ((12, 10, 8), 24, "higuera_cary")])
def test_fused_step_on_cell_sorted_dense_plasma_matches_oracle(N, ppc, pusher):
    """Cell-sorted particles with long equal-cell runs exercise the wave
    segmented reduction of the fused deposit (runs longer than a wave too)."""
    static, dyn = workloads.make_parameters(N, cfl=0.5, particle_pusher=pusher)
    p = oracle_params(static, dyn)
    specs = [workloads.SpeciesSpec(ppc, weight=0.01, vth=(0.08,) * 3),
             workloads.SpeciesSpec(ppc // 2, charge=1.0, mass=50.0, weight=0.01, vth=(0.02,) * 3)]
    parts, cfg = workloads.plasma(static, dyn, specs, seed=11)
    x, u, act = host_particles(parts)
    gpu, st = _fields(p, 500)
    st.x, st.u, st.active = x, u, act
    sp = species(2, charge=(-1.0, 1.0), mass=(1.0, 50.0), weight=(0.01, 0.01))
    out_parts, fields = time_loop_electrodynamic(parts, cfg, gpu, static, dyn)
    want = O.time_loop_electrodynamic(st, sp, p)
    xg, ug, ag = host_particles(out_parts)
    np.testing.assert_array_equal(ag, want.active)
    assert np.max(np.abs(xg - want.x)) <= 1e-6 * max(p.wind)
    assert rel_err(ug, want.u) <= TOL
    for c in range(3):
        assert rel_err(host(fields[2][c]), want.J[c]) <= TOL, ("J", c)
        assert rel_err(host(fields[0][c]), want.E[c]) <= TOL, ("E", c)
        assert rel_err(host(fields[1][c]), want.B[c]) <= TOL, ("B", c)


@pytest.mark.parametrize("tile,rounds,min_blocks,dep", [(None, 2, 1, "esirkepov"), (8, 2, 3, "esirkepov"),
                                                       (4, 2, 1, "esirkepov"), (8, 4, 1, "esirkepov"),
                                                       (4, 4, 4, "esirkepov"), (4, 8, 4, "esirkepov"), (8, 4, 1, "direct"),
                                                       (4, 4, 1, "direct")])
def test_runner_matches_time_loop_and_sorting_is_transparent(tile, rounds, min_blocks, dep):
    """Slot-ordered and tile-ordered runners (re-sorting every 2 steps, the
    tile grid not dividing the domain) against the functional step."""
    from pypic3d_amd.runner import ElectrodynamicRunner
    static, dyn = workloads.make_parameters((16, 12, 10), cfl=0.5, current_deposition=dep,
                                            current_filter="bilinear" if dep == "direct" else "none")
    specs = [workloads.SpeciesSpec(16, weight=0.01, vth=(0.1,) * 3),
             workloads.SpeciesSpec(5, charge=1.0, mass=20.0, weight=0.03, vth=(0.02,) * 3)]
    parts, cfg = workloads.plasma(static, dyn, specs, seed=5)
    fields = workloads.empty_fields(static, dyn)
    p = oracle_params(static, dyn)
    E0 = tuple(gpu_field(random_field(p, 60 + c, scale=0.1)) for c in range(3))
    fields = (E0,) + fields[1:]
    runner = ElectrodynamicRunner(parts, cfg, fields, static, dyn, sort_every=2, tile=tile, rounds=rounds,
                                  min_blocks=min_blocks)
    ref_parts, ref_fields = parts, fields
    for _ in range(5):
        runner.step()
        ref_parts, ref_fields = time_loop_electrodynamic(ref_parts, cfg, ref_fields, static, dyn)
    r_parts, r_fields = runner.state()
    # the runner re-sorted its slots; compare particle sets ordered by position
    xa, ua, aa = host_particles(r_parts)
    xb, ub, ab = host_particles(ref_parts)
    for s_ in range(2):
        oa = np.lexsort((xa[s_, :, 2], xa[s_, :, 1], xa[s_, :, 0]))
        ob = np.lexsort((xb[s_, :, 2], xb[s_, :, 1], xb[s_, :, 0]))
        np.testing.assert_array_equal(aa[s_, oa], ab[s_, ob])
        assert np.max(np.abs(xa[s_, oa] - xb[s_, ob])) <= 1e-5
        assert rel_err(ua[s_, oa], ub[s_, ob]) <= 1e-5
    for k in (0, 1, 2):
        for c in range(3):
            assert rel_err(host(r_fields[k][c]), host(ref_fields[k][c])) <= 1e-5, (k, c)


@pytest.mark.parametrize("tile,sf", [(8, 1), (4, 1), (4, 2)])
def test_tile_ordered_rho_matches_oracle(tile, sf):
    """compute_rho's tile-ordered deposit on tile-sorted particles."""
    from pypic3d_amd import _native as N
    from pypic3d_amd.deposition.rho import deposit_rho_into
    from pypic3d_amd.runner import ElectrodynamicRunner
    static, dyn = workloads.make_parameters((14, 9, 11), cfl=0.5, shape_factor=sf, current_deposition="direct")
    specs = [workloads.SpeciesSpec(6, weight=0.02, vth=(0.1,) * 3),
             workloads.SpeciesSpec(3, charge=1.0, mass=30.0, weight=0.05, vth=(0.02,) * 3)]
    parts, cfg = workloads.plasma(static, dyn, specs, seed=9)
    r = ElectrodynamicRunner(parts, cfg, workloads.empty_fields(static, dyn), static, dyn, tile=tile)
    r.step()             # drift off the sorted cells before depositing
    rho = torch.empty_like(r.J[0])
    deposit_rho_into(r.desc, r.sp, r._pp(r.cur), rho, "none", N.stream_handle(), tile=r.tile, starts=r.starts)
    p = oracle_params(static, dyn)
    x, u, act = host_particles(r.state()[0])
    from tests.gpu_helpers import oracle_species
    want = O.compute_rho(x, act, oracle_species(cfg), p)
    assert rel_err(host(rho), want) <= TOL
