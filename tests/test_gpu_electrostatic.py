"""GPU parity of the electrostatic path against the FP64 oracle at 1e-5: the
refined CG, the tiled field solve with conducting walls and the digital
filter, the untiled calculate_electrostatic_fields, the reference's
known-answer cases, and the FP64 momentum and charge diagnostics."""

import numpy as np
import pytest
import torch

from oracle import pic_oracle as O
from pypic3d_amd import workloads
from pypic3d_amd._tensors import tiled_particles_from_arrays
from pypic3d_amd.deposition.rho import compute_rho
from pypic3d_amd.particles.particle_class import SpeciesConfig
from pypic3d_amd.solvers.electrostatic_yee import (calculate_electrostatic_fields,
                                                   calculate_tiled_electrostatic_fields, last_solve_info,
                                                   solve_poisson_with_conjugate_gradient)
from pypic3d_amd.utils import compute_total_charge, compute_total_momentum
from tests.gpu_helpers import TOL, gpu_field, host, oracle_params, rel_err
from tests.helpers import random_particles, species

pytestmark = pytest.mark.gpu


def _cfg(sp):
    return SpeciesConfig(sp.charge, sp.mass, sp.weight, sp.update_x, sp.update_u)


def _setup(N, wind, bc=(0, 0, 0), alpha=1.0):
    static, dyn = workloads.make_parameters(N, wind=wind, dt=0.05, electrostatic=True, current_deposition="direct",
                                            bc=bc, alpha=alpha)
    return static, dyn, oracle_params(static, dyn)


def _neutral_particles(p, P=600, seed=4):
    """Two species of opposite charge, equal weight and slot count, all active:
    zero net charge (the periodic / Neumann systems are then consistent)."""
    x, u, _ = random_particles(p, S=2, P=P, seed=seed, vmax=0.03)
    return x, u, np.ones((2, P), bool), species(2, charge=(-1.0, 1.0), mass=(1.0, 100.0), weight=(0.5, 0.5))


def test_refined_cg_reaches_the_fp64_solution_of_a_single_mode():
    static, dyn, p = _setup((24, 20, 16), (2 * np.pi,) * 3)
    xs = [O.tiled_axis(p, a, "center")[2:-2] for a in range(3)]
    X, Y, Z = np.meshgrid(*xs, indexing="ij")
    rhs = -O.laplacian_interior(O.refresh(np.pad(np.sin(X + Y + Z), 2), p), p)
    rho = np.zeros(p.shape)
    rho[2:-2, 2:-2, 2:-2] = rhs
    rho = rho.astype(np.float32).astype(np.float64)
    got = host(solve_poisson_with_conjugate_gradient(gpu_field(rho), gpu_field(np.zeros(p.shape)), static, dyn))
    assert last_solve_info["residual_norm"] <= 1e-12
    want = O.solve_poisson_cg(rho, np.zeros(p.shape), p, tol=1e-13, max_iter=5000)
    gi, wi = got[2:-2, 2:-2, 2:-2], want[2:-2, 2:-2, 2:-2]
    # FP32 output of an FP64-accurate potential
    assert rel_err(gi - gi.mean(), wi - wi.mean()) <= 1e-6


@pytest.mark.parametrize("N,bc,alpha", [((16, 12, 10), (0, 0, 0), 1.0), ((16, 12, 10), (1, 0, 0), 0.7),
                                         ((14, 10, 8), (1, 1, 0), 1.0), ((12, 10, 8), (1, 1, 1), 0.8),
                                         ((32, 1, 1), (0, 0, 0), 1.0), ((16, 12, 1), (1, 0, 0), 1.0)])
def test_tiled_electrostatic_fields_match_oracle(N, bc, alpha):
    static, dyn, p = _setup(N, tuple(0.5 * n for n in N), bc=bc, alpha=alpha)
    x, u, act, sp = _neutral_particles(p)
    parts = tiled_particles_from_arrays(x, u, act)
    z = gpu_field(np.zeros(p.shape))
    E, phi, rho = calculate_tiled_electrostatic_fields(static, dyn, parts, _cfg(sp), z, z.clone())
    rho_w = O.compute_rho(x, act, sp, p)
    phi_w = O.solve_poisson_cg(rho_w, np.zeros(p.shape), p, tol=1e-13, max_iter=5000)
    phi_w = O.phi_boundaries(phi_w, p)
    phi_w = O.phi_boundaries(O.digital_filter(phi_w, p.alpha, p.g), p)
    E_w = O.centered_gradient_E(phi_w, p)
    assert rel_err(host(rho), rho_w) <= TOL
    g = host(phi)
    A = (slice(2, -2),) * 3
    singular = all(b == 0 for b in bc) or bc == (1, 0, 0)
    if singular:          # constant null space: compare mean-free potentials
        assert rel_err(g[A] - g[A].mean(), phi_w[A] - phi_w[A].mean()) <= TOL
    else:
        assert rel_err(g, phi_w) <= TOL
    for c in range(3):
        assert rel_err(host(E[c]), E_w[c]) <= TOL, c


@pytest.mark.parametrize("bc", [(0, 0, 0), (1, 0, 0)])
def test_untiled_electrostatic_fields_match_oracle(bc):
    """calculate_electrostatic_fields (electrostatic_yee.py:159-209): CG on a
    given rho and the rolled interior gradient."""
    static, dyn, p = _setup((16, 12, 10), (8.0, 6.0, 5.0), bc=bc, alpha=0.8)
    # FP32 values and their negatives at shuffled cells: the total is exactly
    # zero in FP64, so the singular systems are consistent for the oracle's CG
    rng = np.random.default_rng(8)
    n = int(np.prod(p.N))
    half = rng.standard_normal(n // 2).astype(np.float32).astype(np.float64)
    rho = np.zeros(p.shape)
    rho[2:-2, 2:-2, 2:-2] = rng.permutation(np.concatenate([half, -half])).reshape(p.N)
    assert rho.sum() == 0.0
    E, phi, rho_out = calculate_electrostatic_fields(static, dyn, None, gpu_field(rho), gpu_field(np.zeros(p.shape)),
                                                     "electrostatic", "periodic")
    E_w, phi_w, _ = O.electrostatic_fields_from_rho(rho, np.zeros(p.shape), p, tol=1e-13)
    np.testing.assert_array_equal(host(rho_out), rho)
    g = host(phi)
    A = (slice(2, -2),) * 3
    assert rel_err(g[A] - g[A].mean(), phi_w[A] - phi_w[A].mean()) <= TOL
    for c in range(3):
        assert rel_err(host(E[c]), E_w[c]) <= TOL, c


def test_untiled_fields_from_zero_charge_are_zero():
    # electrostatic_yee_test.py:121-149
    static, dyn, p = _setup((16, 16, 16), (2 * np.pi,) * 3)
    z = gpu_field(np.zeros(p.shape))
    E, phi, rho = calculate_electrostatic_fields(static, dyn, None, z, z.clone(), "electrostatic", "periodic")
    assert float(phi.abs().max()) == 0.0
    for c in E:
        assert float(c.abs().max()) == 0.0


def test_zero_charge_particle_keeps_a_constant_potential():
    # electrostatic_yee_test.py:175-222
    static, dyn, p = _setup((8, 4, 4), (1.0, 1.0, 1.0), bc=(1, 0, 0))
    parts = tiled_particles_from_arrays(np.zeros((1, 1, 3)), np.zeros((1, 1, 3)), np.ones((1, 1), bool))
    cfg = SpeciesConfig(np.array([0.0]), np.array([1.0]), np.array([1.0]), np.ones((1, 3), bool),
                        np.ones((1, 3), bool))
    phi0 = np.zeros(p.shape)
    phi0[2:-2, 2:-2, 2:-2] = 3.0
    z = gpu_field(np.zeros(p.shape))
    E, phi, rho = calculate_tiled_electrostatic_fields(static, dyn, parts, cfg, z, gpu_field(phi0))
    ph = host(phi)
    assert float(rho.abs().max()) == 0.0
    np.testing.assert_allclose(ph[2:-2, 2:-2, 2:-2], 3.0, atol=1e-12)
    np.testing.assert_allclose(ph[:2], np.broadcast_to(ph[2:3], ph[:2].shape), atol=1e-12)
    np.testing.assert_allclose(ph[-2:], np.broadcast_to(ph[-3:-2], ph[-2:].shape), atol=1e-12)
    for c in E:
        assert float(c.abs().max()) == 0.0


def test_momentum_and_charge_diagnostics_are_fp64_accurate():
    static, dyn, p = _setup((12, 10, 8), (6.0, 5.0, 4.0))
    x, u, act = random_particles(p, S=2, P=3000, seed=12, vmax=0.03, inactive_frac=0.1)
    sp = species(2, charge=(-1.0, 2.0), mass=(1.0, 7.0), weight=(0.3, 0.2))
    parts = tiled_particles_from_arrays(x, u, act)
    got = float(compute_total_momentum(parts, _cfg(sp)))
    want = O.compute_total_momentum(u, act, sp)
    assert abs(got - want) <= 1e-12 * abs(want)
    rho = compute_rho(parts, _cfg(sp), gpu_field(np.zeros(p.shape)), static, dyn)
    q = float(compute_total_charge(rho, static, dyn))
    qw = float(np.sum(act * (sp.charge * sp.weight)[:, None]))
    # rho is FP32: its sum carries FP32 rounding of each cell value
    assert abs(q - qw) <= 1e-6 * float(np.sum(act * np.abs(sp.charge * sp.weight)[:, None]))
    assert abs(q - O.total_charge(host(rho), p)) <= 1e-12 * max(abs(q), 1.0)
