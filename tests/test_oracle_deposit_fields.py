"""Pins for the oracle's deposits, ghost cells, Yee, CG, sort and energy (CPU only)."""

import math

import numpy as np
import pytest

from oracle import pic_oracle as O
from tests.helpers import fp32, interior, random_particles, species

DIMS = [(8, 1, 1), (1, 8, 1), (1, 1, 8), (8, 8, 1), (8, 1, 8), (1, 8, 8), (6, 6, 6)]


def _params(N, sf, **kw):
    wind = tuple(float(n) * 0.5 for n in N)
    return O.Params(N=N, wind=wind, g=2, dt=0.05, shape_factor=sf, **kw)


@pytest.mark.parametrize("sf", [1, 2])
@pytest.mark.parametrize("N", DIMS)
def test_esirkepov_discrete_continuity(sf, N):
    # esirkepov_test.py:700-744 extended to every axis and dimensionality:
    # (rho_new - rho_old)/dt + div J = 0 on the interior.
    p = _params(N, sf)
    x, u, act = random_particles(p, S=2, P=40, seed=sum(N) + sf, vmax=0.45, inactive_frac=0.1)
    sp = species(2)
    rho_old = O.compute_rho(x, act, sp, p)
    J = O.esirkepov_current(x, u, act, sp, p)
    xn = O.update_positions(x, u, act, sp, p.dt)
    xn, _, actn, _ = O.refresh_particles(xn, u, act, p)
    rho_new = O.compute_rho(xn, actn, sp, p)
    g = p.g
    A = slice(g, -g)
    B_ = slice(g - 1, -g - 1)
    drho = (rho_new[A, A, A] - rho_old[A, A, A]) / p.dt
    div = np.zeros_like(drho)
    if p.N[0] > 1:
        div += (J[0][A, A, A] - J[0][B_, A, A]) / p.d[0]
    if p.N[1] > 1:
        div += (J[1][A, A, A] - J[1][A, B_, A]) / p.d[1]
    if p.N[2] > 1:
        div += (J[2][A, A, A] - J[2][A, A, B_]) / p.d[2]
    scale = max(1.0, np.max(np.abs(drho)) + np.max(np.abs(div)))
    assert np.max(np.abs(drho + div)) <= 1e-12 * scale


def _brute_esirkepov_single(xo, v, q, p):
    """Pure-Python restatement of Esirkepov (2001) in difference form:
    W = D_x (S0y S0z + (S0y Dz + Dy S0z)/2 + Dy Dz/3), with D = S1 - S0 on a
    common per-axis stencil, J_x = -q dx/(dV dt) * running sum of W along x.
    Written independently of the reference's 5-slot roll formulation; 3-D only."""
    g = p.g
    J = [np.zeros(p.shape) for _ in range(3)]
    S0, S1, base = [], [], []
    for a in range(3):
        line = O.tiled_axis(p, a, "center")
        org, sp_ = line[0], line[1] - line[0]
        xn = xo[a] + v[a] * p.dt

        def shape(pos):
            t = (pos - org) / sp_
            k = math.floor(t) if p.shape_factor == 1 else int(np.round(t))
            r = (pos - (k * sp_ + org)) / p.d[a]
            w = [0.0, 1 - r, r] if p.shape_factor == 1 else [0.5 * (0.5 - r) ** 2, 0.75 - r * r, 0.5 * (0.5 + r) ** 2]
            return {k - 1: w[0], k: w[1], k + 1: w[2]}

        s0, s1 = shape(xo[a]), shape(xn)
        lo = min(min(s0), min(s1))
        hi = max(max(s0), max(s1))
        idx = list(range(lo, hi + 1))
        S0.append([s0.get(i, 0.0) for i in idx])
        S1.append([s1.get(i, 0.0) for i in idx])
        base.append(lo)
    D = [[S1[a][i] - S0[a][i] for i in range(len(S0[a]))] for a in range(3)]
    dV = p.d[0] * p.d[1] * p.d[2]
    n = [len(S0[a]) for a in range(3)]
    for c in range(3):
        o1, o2 = [a for a in range(3) if a != c]
        for j in range(n[o1]):
            for k in range(n[o2]):
                run = 0.0
                for i in range(n[c]):
                    w = D[c][i] * (S0[o1][j] * S0[o2][k] + 0.5 * (S0[o1][j] * D[o2][k] + D[o1][j] * S0[o2][k])
                                   + D[o1][j] * D[o2][k] / 3.0)
                    run += w
                    idx = [0, 0, 0]
                    idx[c] = base[c] + i
                    idx[o1] = base[o1] + j
                    idx[o2] = base[o2] + k
                    J[c][tuple(idx)] += -q * p.d[c] / (dV * p.dt) * run
    return J


@pytest.mark.parametrize("sf", [1, 2])
def test_esirkepov_matches_bruteforce_difference_form(sf):
    p = _params((6, 6, 6), sf)
    rng = np.random.default_rng(7)
    for trial in range(6):
        xo = fp32(rng.uniform(-1.2, 1.2, 3))
        v = fp32(rng.uniform(-0.8, 0.8, 3) * 0.5 / p.dt)   # crossings included
        x = xo.reshape(1, 1, 3)
        u = v.reshape(1, 1, 3)
        act = np.ones((1, 1), bool)
        sp = species(1, charge=(-1.0,), weight=(0.5,))
        got = O.esirkepov_raw(x, u, act, sp, p)
        want = _brute_esirkepov_single(xo, v, -0.5, p)
        for c in range(3):
            np.testing.assert_allclose(got[c], want[c], atol=1e-12)


@pytest.mark.parametrize("sf", [1, 2])
@pytest.mark.parametrize("xpos", [-1.32, -0.03, 1.97])
def test_rho_and_direct_current_single_particle_moments(sf, xpos):
    # single_particle_pipeline_test.py:229-320 inputs; pins by invariants:
    # total charge = q*w (:429-431), total current = q*w*v, first moment of
    # rho on the collocated grid = q*w*x, of J_x on the staggered grid = q*w*vx*x
    # (CIC and TSC reproduce linears; the extrapolating face weights too).
    p = O.Params(N=(8, 1, 1), wind=(4.0, 1.0, 1.0), g=2, dt=0.05, shape_factor=sf)
    x = np.array([[[xpos, 0.0, 0.0]]])
    u = np.array([[[0.11, -0.17, 0.07]]])
    act = np.ones((1, 1), bool)
    sp = species(1, charge=(-1.0,), weight=(0.5,))
    dV = p.d[0] * p.d[1] * p.d[2]
    raw = O.rho_raw(x, act, sp, p)
    line_c = O.tiled_axis(p, 0, "center")
    line_v = O.tiled_axis(p, 0, "vertex")
    assert abs(raw.sum() * dV - (-0.5)) < 1e-12
    assert abs((raw.sum(axis=(1, 2)) * line_c).sum() * dV - (-0.5 * xpos)) < 1e-12
    rho = O.compute_rho(x, act, sp, p)
    assert abs(interior(rho, 2).sum() * dV - (-0.5)) < 1e-12
    Jr = O.j_from_rhov_raw(x, u, act, sp, p)
    for c in range(3):
        assert abs(Jr[c].sum() * dV - (-0.5 * u[0, 0, c])) < 1e-12
    assert abs((Jr[0].sum(axis=(1, 2)) * line_v).sum() * dV - (-0.5 * u[0, 0, 0] * xpos)) < 1e-12


def test_inactive_slots_with_noise_do_not_deposit():
    # rho_test.py:184-200: inactive slots may hold arbitrary data.
    p = _params((6, 6, 6), 1)
    x, u, act = random_particles(p, S=1, P=10, seed=9)
    act[0, 4:] = False
    x2 = x.copy(); x2[0, 4:] = 123.0
    sp = species(1)
    a = O.compute_rho(x, act, sp, p)
    b = O.compute_rho(x2, act, sp, p)
    np.testing.assert_array_equal(a, b)


def test_refresh_and_fold_slab_bookkeeping():
    p = O.Params(N=(4, 3, 1), wind=(4.0, 3.0, 1.0), g=2)
    f = np.arange(np.prod(p.shape), dtype=np.float64).reshape(p.shape)
    r = O.refresh(f, p)
    g = 2
    # periodic x: lower ghost <- upper interior, upper ghost <- lower interior
    np.testing.assert_array_equal(r[0:2, g:-g, g], f[4:6, g:-g, g])
    np.testing.assert_array_equal(r[6:8, g:-g, g], f[2:4, g:-g, g])
    # reduced z: every ghost equals interior[g]
    np.testing.assert_array_equal(r[:, :, 0], r[:, :, 2])
    np.testing.assert_array_equal(r[:, :, 4], r[:, :, 2])
    # fold conserves the total and moves a lone ghost deposit to its owner
    h = np.zeros(p.shape); h[0, 3, 2] = 1.0; h[7, 0, 4] = 2.0
    fo = O.fold(h, p)
    assert fo.sum() == 3.0
    assert fo[4, 3, 2] == 1.0       # x lower ghost 0 -> upper interior 4
    assert fo[3, 3, 2] == 2.0       # x upper ghost 7 -> 3, y lower ghost 0 -> 3, z ghost -> g
    # conducting fold subtracts into the adjacent interior slab
    pc = O.Params(N=(4, 3, 1), wind=(4.0, 3.0, 1.0), g=2, bc=(1, 0, 0))
    fc = O.fold(h, pc)
    assert fc[2, 3, 2] == -1.0 and fc[:2].sum() == 0.0


def test_update_E_from_zero_fields_is_minus_dt_J_over_eps():
    # esirkepov_test.py:510-529: J=4, eps=2, dt=0.25 -> E=-0.5
    p = O.Params(N=(4, 1, 1), wind=(4.0, 1.0, 1.0), g=2, dt=0.25, eps=2.0)
    E = [np.zeros(p.shape)] * 3
    B = [np.zeros(p.shape)] * 3
    J = [np.zeros(p.shape), np.zeros(p.shape), np.zeros(p.shape)]
    J[0][2:-2, 2:-2, 2:-2] = 4.0
    out = O.update_E(E, B, J, p)
    assert np.all(out[0][2:-2, 2:-2, 2:-2] == -0.5)


def test_update_B_second_order_convergence_with_half_step():
    # yee_convergence_test.py:77-98 (the B half step is dt/2, first_order_yee.py:116)
    errs = []
    for N in (32, 64, 128):
        p = O.Params(N=(N, 1, 1), wind=(2 * math.pi, 1.0, 1.0), g=2, dt=1e-3)
        E = [np.zeros(p.shape) for _ in range(3)]
        xc = O.tiled_axis(p, 0, "center")
        E[2][2:-2, 2:-2, 2:-2] = np.sin(xc[2:-2])[:, None, None]
        B = O.update_B(E, [np.zeros(p.shape)] * 3, p)
        xv = O.tiled_axis(p, 0, "vertex")
        exact = 0.5 * p.dt * np.cos(xv[2:-2])
        errs.append(np.sqrt(np.mean((B[1][2:-2, 2, 2] - exact) ** 2)))
    assert math.log2(errs[0] / errs[1]) > 1.8 and math.log2(errs[1] / errs[2]) > 1.8


def test_discrete_curl_of_gradient_vanishes_and_div_of_curl_vanishes():
    # Yee identities: B update from E = grad(phi) is zero; the E update
    # changes div E by exactly -dt * div J / eps (div curl B = 0).
    p = O.Params(N=(6, 5, 4), wind=(6.0, 5.0, 4.0), g=2, dt=0.1, eps=1.5, C=0.9)
    rng = np.random.default_rng(11)
    phi = O.refresh(rng.standard_normal(p.shape), p)
    A = slice(2, -2); F = slice(3, -1)
    E = [np.zeros(p.shape) for _ in range(3)]
    E[0][A, A, A] = (phi[F, A, A] - phi[A, A, A]) / p.d[0]
    E[1][A, A, A] = (phi[A, F, A] - phi[A, A, A]) / p.d[1]
    E[2][A, A, A] = (phi[A, A, F] - phi[A, A, A]) / p.d[2]
    B0 = [O.refresh(rng.standard_normal(p.shape), p) for _ in range(3)]
    B1 = O.update_B(E, B0, p)
    for c in range(3):
        np.testing.assert_allclose(B1[c][A, A, A], B0[c][A, A, A], atol=1e-12)
    J = [O.refresh(rng.standard_normal(p.shape), p) for _ in range(3)]
    E0 = [O.refresh(rng.standard_normal(p.shape), p) for _ in range(3)]
    E1 = O.update_E(E0, B0, J, p)
    K = slice(1, -3)

    def div(V):
        return ((V[0][A, A, A] - V[0][K, A, A]) / p.d[0] + (V[1][A, A, A] - V[1][A, K, A]) / p.d[1]
                + (V[2][A, A, A] - V[2][A, A, K]) / p.d[2])
    np.testing.assert_allclose(div(E1) - div(E0), -p.dt * div(J) / p.eps, atol=1e-11)


def test_filters():
    p = O.Params(N=(6, 5, 4), wind=(6.0, 5.0, 4.0), g=2)
    rng = np.random.default_rng(12)
    f = O.refresh(rng.standard_normal(p.shape), p)
    np.testing.assert_array_equal(O.digital_filter(f, 1.0, 2), f)
    np.testing.assert_allclose(O.bilinear_filter(np.full(p.shape, 2.0), 2), 2.0)
    # the digital filter conserves the periodic sum for any alpha
    out = O.digital_filter(f, 0.4, 2)
    np.testing.assert_allclose(interior(out, 2).sum(), interior(f, 2).sum(), rtol=1e-12)


def test_cg_single_mode_and_sparse_solve():
    # electrostatic_yee_test.py:92-119
    import scipy.sparse as sps
    import scipy.sparse.linalg as spl
    N = (12, 10, 8)
    p = O.Params(N=N, wind=(2 * math.pi,) * 3, g=2, electrostatic=True)
    xs = [O.tiled_axis(p, a, "center")[2:-2] for a in range(3)]
    X, Y, Z = np.meshgrid(*xs, indexing="ij")
    phi_true = np.sin(X + Y + Z)
    phi_pad = O.refresh(np.pad(phi_true, 2), p)
    rhs = -O.laplacian_interior(phi_pad, p)
    rho = np.zeros(p.shape); rho[2:-2, 2:-2, 2:-2] = rhs * p.eps
    phi, iters = O.solve_poisson_cg(rho, np.zeros(p.shape), p, tol=1e-10, max_iter=4000, return_iterations=True)
    num = interior(phi, 2) - interior(phi, 2).mean()
    np.testing.assert_allclose(num, phi_true - phi_true.mean(), atol=1e-7)
    # scipy direct solve of the same periodic 7-point system (mean-zero gauge)
    def lap1(n, d):
        m = sps.diags([np.ones(n - 1), -2 * np.ones(n), np.ones(n - 1)], [-1, 0, 1]).tolil()
        m[0, n - 1] = 1; m[n - 1, 0] = 1
        return m.tocsr() / (d * d)
    I = [sps.identity(n) for n in N]
    L = (sps.kron(sps.kron(lap1(N[0], p.d[0]), I[1]), I[2]) + sps.kron(sps.kron(I[0], lap1(N[1], p.d[1])), I[2])
         + sps.kron(sps.kron(I[0], I[1]), lap1(N[2], p.d[2])))
    b = (rho[2:-2, 2:-2, 2:-2] / p.eps).reshape(-1)
    sol = spl.lsqr(-L, b, atol=1e-14, btol=1e-14, iter_lim=20000)[0].reshape(N)
    np.testing.assert_allclose(num, sol - sol.mean(), atol=1e-7)


def test_sort_is_stable_bucket_order():
    p = O.Params(N=(4, 3, 2), wind=(4.0, 3.0, 2.0), g=2)
    x, _, _ = random_particles(p, S=1, P=200, seed=13)
    keys = O.cell_keys(x, p)[0]
    perm = O.sort_permutation(keys)
    buckets = [[] for _ in range(24)]
    for i, k in enumerate(keys):
        buckets[k].append(i)
    assert perm.tolist() == [i for b in buckets for i in b]
    # key = (ix*Ny + iy)*Nz + iz from floor((x + W/2)/d), clipped
    assert O.cell_keys(np.array([[[2.0, -1.5, 0.99]]]), p)[0, 0] == (3 * 3 + 0) * 2 + 1


def test_particle_boundaries():
    # particle_refresh_test.py:127-170 and wrap_periodic_position's +W/2 rule
    p = O.Params(N=(8, 1, 1), wind=(4.0, 1.0, 1.0), particle_bc=(0, 0, 0))
    x = np.array([[[2.1, 0.0, 0.0], [2.0, 0.0, 0.0], [-2.25, 0.0, 0.0]]])
    u = np.ones((1, 3, 3))
    xo, _, _, _ = O.refresh_particles(x, u, np.ones((1, 3), bool), p)
    np.testing.assert_allclose(xo[0, :, 0], [-1.9, 2.0, 1.75], atol=1e-15)
    pr = O.Params(N=(8, 1, 1), wind=(4.0, 1.0, 1.0), particle_bc=(1, 0, 0))
    xo, uo, _, _ = O.refresh_particles(x, u, np.ones((1, 3), bool), pr)
    np.testing.assert_allclose(xo[0, :, 0], [1.9, 2.0, -1.75], atol=1e-15)
    assert uo[0, 0, 0] == -1 and uo[0, 1, 0] == -1 and uo[0, 2, 0] == -1
    pa = O.Params(N=(8, 1, 1), wind=(4.0, 1.0, 1.0), particle_bc=(2, 0, 0))
    xo, uo, ao, _ = O.refresh_particles(x, u, np.ones((1, 3), bool), pa)
    assert ao.tolist() == [[False, True, False]]
    assert np.all(xo[0, 0] == 0) and np.all(uo[0, 2] == 0)


def test_wrap_periodic_known_answers():
    # grid_and_stencil_test.py:27-33: W = 4
    x = np.array([-2.5, -2.0, 1.75, 2.0, 2.25, 6.5])
    np.testing.assert_array_equal(O.wrap_periodic(x, 4.0), [1.5, -2.0, 1.75, 2.0, -1.75, -1.5])


def test_energy_closed_forms():
    p = O.Params(N=(4, 3, 2), wind=(2.0, 1.5, 1.0), g=2, eps=2.0, mu=0.5, C=1.0)
    E = [np.full(p.shape, 1.0), np.zeros(p.shape), np.zeros(p.shape)]
    B = [np.zeros(p.shape), np.full(p.shape, 2.0), np.zeros(p.shape)]
    u = np.zeros((1, 1, 3)); u[0, 0, 0] = 0.6
    sp = species(1, mass=(2.0,), weight=(3.0,))
    uE, uB, ke = O.compute_energy(u, np.ones((1, 1), bool), sp, E, B, p)
    vol = 2.0 * 1.5 * 1.0
    assert abs(uE - 0.5 * 2.0 * vol) < 1e-12
    assert abs(uB - 0.5 / 0.5 * 4.0 * vol) < 1e-12
    assert abs(ke - (1.25 - 1.0) * 6.0) < 1e-12          # (gamma - 1) m c^2, gamma = 1.25
    assert abs(O.courant_condition(0.5, O.Params(N=(8, 8, 1), wind=(8.0, 4.0, 1.0))) - 0.5 / 3.0) < 1e-15


@pytest.mark.parametrize("dep", ["esirkepov", "direct"])
def test_single_particle_step_E_is_minus_dt_J_and_B_from_Ey(dep):
    # single_particle_pipeline_test.py:329-518 with the dt/2 B half step.
    p = O.Params(N=(8, 1, 1), wind=(4.0, 1.0, 1.0), g=2, dt=0.05, shape_factor=1 if dep == "esirkepov" else 2,
                 relativistic=False, current_deposition=dep)
    x = np.array([[[-1.32, 0.0, 0.0]]])
    u = np.array([[[0.08, 0.2, 0.0]]])
    z = [np.zeros(p.shape) for _ in range(3)]
    st = O.State(x, u, np.ones((1, 1), bool), z, [c.copy() for c in z])
    sp = species(1, charge=(-1.0,), mass=(1.0,), weight=(0.5,))
    out = O.time_loop_electrodynamic(st, sp, p)
    for c in range(3):
        np.testing.assert_allclose(out.E[c][2:-2, 2:-2, 2:-2], -p.dt * out.J[c][2:-2, 2:-2, 2:-2] / p.eps, atol=1e-14)
    Ey = out.E[1]
    expected_Bz = -(p.dt / 2) * (Ey[3:-1, 2:-2, 2:-2] - Ey[2:-2, 2:-2, 2:-2]) / p.d[0]
    np.testing.assert_allclose(out.B[2][2:-2, 2:-2, 2:-2], expected_Bz, atol=1e-14)
    if dep == "esirkepov":
        assert abs(out.x[0, 0, 0] - (-1.32 + 0.08 * 0.05)) < 1e-14
