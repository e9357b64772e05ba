"""The FP64 path (float64 particles and fields, the reference's precision)
against the FP64 oracle at 1e-12 norm-wise relative error: each kernel over
every dimensionality and both shape orders, whole steps of both deposition
branches, and the 64-cell two-stream run for its 500 steps."""

import numpy as np
import pytest
import torch

from oracle import pic_oracle as O
from pypic3d_amd import workloads
from pypic3d_amd._tensors import tiled_particles_from_arrays
from pypic3d_amd.deposition.Esirkepov import Esirkepov_current
from pypic3d_amd.deposition.J_from_rhov import J_from_rhov
from pypic3d_amd.deposition.rho import compute_rho
from pypic3d_amd.evolve import time_loop_electrodynamic
from pypic3d_amd.particles.particle_class import SpeciesConfig
from pypic3d_amd.particles.particle_tile_communication import (refresh_tiled_particle_tiles,
                                                               update_tiled_particle_positions)
from pypic3d_amd.pusher.particle_push import particle_push
from pypic3d_amd.solvers.first_order_yee import update_B, update_E
from pypic3d_amd.utils import compute_energy
from tests.gpu_helpers import host, host_particles, oracle_params, oracle_species, rel_err
from tests.helpers import random_particles, species

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
DIMS = [(8, 1, 1), (1, 8, 1), (8, 8, 1), (1, 8, 8), (6, 6, 6), (12, 10, 8)]


def f64(a):
    return torch.as_tensor(np.asarray(a, dtype=np.float64)).cuda().view(1, 1, 1, *np.asarray(a).shape)


def _setup(N, sf=1, rel=True, dep="esirkepov", filt="none", bc=(0, 0, 0), pbc=(0, 0, 0), alpha=1.0,
           pusher="boris"):
    static, dyn = workloads.make_parameters(N, wind=tuple(0.5 * n for n in N), dt=0.05, shape_factor=sf,
                                            relativistic=rel, current_deposition=dep, current_filter=filt, bc=bc,
                                            pbc=pbc, alpha=alpha, particle_pusher=pusher)
    return static, dyn, oracle_params(static, dyn)


def _cfg(sp):
    return SpeciesConfig(sp.charge, sp.mass, sp.weight, sp.update_x, sp.update_u)


def _field(p, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    return O.refresh(rng.standard_normal(p.shape) * scale, p)


def _particles(p, seed, vmax=0.3, inactive=0.1, S=2, P=200):
    rng = np.random.default_rng(seed)
    x = np.zeros((S, P, 3))
    u = np.zeros((S, P, 3))
    for a in range(3):
        x[..., a] = rng.uniform(-p.wind[a] / 2, p.wind[a] / 2, (S, P)) if p.N[a] > 1 else \
            rng.uniform(-0.4, 0.4, (S, P)) * p.d[a]
        u[..., a] = rng.uniform(-vmax, vmax, (S, P)) * p.d[a] / p.dt
    return x, u, rng.random((S, P)) >= inactive


@pytest.mark.parametrize("N", DIMS)
@pytest.mark.parametrize("sf", [1, 2])
@pytest.mark.parametrize("pusher,rel", [("boris", True), ("boris", False), ("higuera_cary", True)])
def test_push_f64(N, sf, pusher, rel):
    static, dyn, p = _setup(N, sf=sf, rel=rel, pusher=pusher)
    x, u, act = _particles(p, 1, vmax=0.02)
    sp = species(2)
    E = [_field(p, 10 + c, 0.3) for c in range(3)]
    B = [_field(p, 20 + c, 0.3) for c in range(3)]
    parts = tiled_particles_from_arrays(x, u, act, dtype=torch.float64)
    out = particle_push(parts, _cfg(sp), tuple(map(f64, E)), tuple(map(f64, B)), static, dyn)
    want = O.particle_push(x, u, act, sp, E, B, p)
    assert out.u.dtype == torch.float64
    assert rel_err(host_particles(out)[1], want) <= TOL64


@pytest.mark.parametrize("N", DIMS)
@pytest.mark.parametrize("sf", [1, 2])
def test_deposits_f64(N, sf):
    static, dyn, p = _setup(N, sf=sf)
    x, u, act = _particles(p, 2, vmax=0.4)
    sp = species(2)
    parts = tiled_particles_from_arrays(x, u, act, dtype=torch.float64)
    z = [f64(np.zeros(p.shape)) for _ in range(3)]
    J = Esirkepov_current(parts, _cfg(sp), tuple(z), static, dyn)
    want = O.esirkepov_current(x, u, act, sp, p)
    for c in range(3):
        assert rel_err(host(J[c]), want[c]) <= TOL64, ("esirkepov", c)
    for filt in ("none", "bilinear", "digital"):
        static_f, dyn_f, p_f = _setup(N, sf=sf, dep="direct", filt=filt, alpha=0.7)
        Jd = J_from_rhov(parts, _cfg(sp), tuple(z), static_f, dyn_f)
        want_d = O.j_from_rhov(x, u, act, sp, p_f)
        for c in range(3):
            assert rel_err(host(Jd[c]), want_d[c]) <= TOL64, ("direct", filt, c)
        rho = compute_rho(parts, _cfg(sp), z[0], static_f, dyn_f)
        assert rel_err(host(rho), O.compute_rho(x, act, sp, p_f)) <= TOL64, ("rho", filt)


@pytest.mark.parametrize("pbc", [(0, 0, 0), (1, 2, 0)])
def test_positions_and_boundaries_f64(pbc):
    static, dyn, p = _setup((8, 6, 4), pbc=pbc)
    x, u, act = _particles(p, 3, vmax=3.0)
    sp = species(2)
    parts = tiled_particles_from_arrays(x, u, act, dtype=torch.float64)
    moved = update_tiled_particle_positions(parts, _cfg(sp), dyn.dt)
    out, _ = refresh_tiled_particle_tiles(moved, static, dyn)
    xw = O.update_positions(x, u, act, sp, p.dt)
    xw, uw, aw, _ = O.refresh_particles(xw, u, act, p)
    xg, ug, ag = host_particles(out)
    np.testing.assert_array_equal(ag, aw)
    np.testing.assert_array_equal(xg, xw)
    np.testing.assert_array_equal(ug, uw)


@pytest.mark.parametrize("N,bc,alpha", [((12, 10, 8), (0, 0, 0), 1.0), ((10, 9, 8), (1, 1, 1), 0.6),
                                         ((16, 12, 1), (0, 1, 0), 0.8), ((32, 1, 1), (0, 0, 0), 1.0)])
def test_yee_f64(N, bc, alpha):
    static, dyn, p = _setup(N, bc=bc, alpha=alpha)
    E = [_field(p, 30 + c) for c in range(3)]
    B = [_field(p, 40 + c) for c in range(3)]
    J = [_field(p, 50 + c) for c in range(3)]
    Bg, _ = update_B(tuple(map(f64, E)), tuple(map(f64, B)), static, dyn, do_filter=True)
    Bw = O.update_B(E, B, p, do_filter=True)
    Eg, _ = update_E(tuple(map(f64, E)), Bg, tuple(map(f64, J)), static, dyn)
    Ew = O.update_E(E, Bw, J, p)
    for c in range(3):
        assert rel_err(host(Bg[c]), Bw[c]) <= TOL64, ("B", c)
        assert rel_err(host(Eg[c]), Ew[c]) <= TOL64, ("E", c)


@pytest.mark.parametrize("dep", ["esirkepov", "direct"])
@pytest.mark.parametrize("N", [(16, 1, 1), (12, 10, 1), (10, 8, 6)])
def test_step_f64(dep, N):
    static, dyn, p = _setup(N, dep=dep, filt="bilinear" if dep == "direct" else "none",
                            sf=1 if dep == "esirkepov" else 2)
    x, u, act = _particles(p, 4, vmax=0.05)
    sp = species(2)
    E = [_field(p, 60 + c, 0.2) for c in range(3)]
    B = [_field(p, 70 + c, 0.2) for c in range(3)]
    Ee = [_field(p, 80 + c, 0.1) for c in range(3)]
    Be = [_field(p, 90 + c, 0.1) for c in range(3)]
    zero = np.zeros(p.shape)
    fields = (tuple(map(f64, E)), tuple(map(f64, B)), tuple(f64(zero) for _ in range(3)), f64(zero), f64(zero),
              (tuple(map(f64, Ee)), tuple(map(f64, Be))), None, torch.zeros((), dtype=torch.bool, device="cuda"))
    parts, out = time_loop_electrodynamic(tiled_particles_from_arrays(x, u, act, dtype=torch.float64), _cfg(sp),
                                          fields, static, dyn)
    want = O.time_loop_electrodynamic(O.State(x, u, act, E, B, E_ext=Ee, B_ext=Be), sp, p)
    xg, ug, ag = host_particles(parts)
    np.testing.assert_array_equal(ag, want.active)
    assert np.max(np.abs(xg - want.x)) <= 1e-14 * max(p.wind)
    assert rel_err(ug, want.u) <= TOL64
    for c in range(3):
        assert rel_err(host(out[2][c]), want.J[c]) <= TOL64, ("J", c)
        assert rel_err(host(out[0][c]), want.E[c]) <= TOL64, ("E", c)
        assert rel_err(host(out[1][c]), want.B[c]) <= TOL64, ("B", c)


def test_energy_f64():
    static, dyn, p = _setup((12, 10, 8))
    x, u, act = _particles(p, 5, vmax=0.02)
    sp = species(2)
    E = [_field(p, 100 + c) for c in range(3)]
    B = [_field(p, 110 + c) for c in range(3)]
    got = compute_energy(tiled_particles_from_arrays(x, u, act, dtype=torch.float64), tuple(map(f64, E)),
                         tuple(map(f64, B)), static, dyn, _cfg(sp))
    for g_, w_ in zip(got, O.compute_energy(u, act, sp, E, B, p)):
        assert abs(float(g_) - w_) <= 1e-13 * abs(w_)


@pytest.mark.timeout(300)
def test_two_stream_64_cells_fp64_500_steps():
    """The 64 x 1 x 1 two-stream configuration (beams at +-0.1c, vth 0.01c,
    32 ppc each, 1e-3 mode-1 density perturbation, dx = Debye length, cfl 0.5)
    in FP64 for 500 steps, the GPU and the oracle each running freely from the
    same initial state: every step's u, x, J, E and B within 1e-12, and the
    field-energy history within 1e-11."""
    static, dyn, specs = workloads.config_two_stream_1d()
    parts, cfg = workloads.plasma(static, dyn, specs, seed=20261202, dtype=torch.float64)
    fields = workloads.empty_fields(static, dyn, dtype=torch.float64)
    p = oracle_params(static, dyn)
    sp = oracle_species(cfg)
    x, u, a = host_particles(parts)
    st = O.State(x, u, a, [host(c) for c in fields[0]], [host(c) for c in fields[1]])
    worst = {}
    ue_g, ue_w = [], []
    for step in range(500):
        parts, fields = time_loop_electrodynamic(parts, cfg, fields, static, dyn)
        st = O.time_loop_electrodynamic(st, sp, p)
        xg, ugp, ag = host_particles(parts)
        np.testing.assert_array_equal(ag, st.active)
        errs = {"x": float(np.max(np.abs(xg - st.x))) / max(p.wind), "u": rel_err(ugp, st.u)}
        for c in range(3):
            for k, name in ((2, "J"), (0, "E"), (1, "B")):
                want = (st.J, st.E, st.B)[(2, 0, 1).index(k)][c]
                if np.max(np.abs(want)) > 0:
                    errs[f"{name}{c}"] = rel_err(host(fields[k][c]), want)
        for key, v in errs.items():
            worst[key] = max(worst.get(key, 0.0), v)
        assert max(errs.values()) <= TOL64, (step, errs)
        ue_g.append(float(compute_energy(parts, fields[0], fields[1], static, dyn, cfg)[0]))
        ue_w.append(O.compute_energy(st.u, st.active, sp, st.E, st.B, p)[0])
    ue_g, ue_w = np.array(ue_g), np.array(ue_w)
    print(f"\n[two-stream fp64] worst per-step errors over 500 steps {worst}; U_E(500) = {ue_w[-1]:.6e}, "
          f"max rel diff of the U_E history {float(np.max(np.abs(ue_g - ue_w)) / np.max(ue_w)):.3e}", flush=True)
    assert np.max(np.abs(ue_g - ue_w)) <= 1e-11 * np.max(ue_w)
