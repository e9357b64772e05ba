"""Pins for oracle functions that the reference's own continuity and
conservation checks leave open (CPU only): the centred electrostatic gradient,
the conducting-wall ghost helpers and the phi boundary sequence, the
extrapolating face weights of J_from_rhov, the Esirkepov currents along
reduced axes, and JAX's ``mode="drop"`` scatter.  Each check is a closed form
or a known answer from the reference's tests, chosen so that a plausible
mistake (a bracketed instead of extrapolated face stencil, a swapped axis, a
wrapped instead of dropped index, a 1/2-1/2 instead of 1/3-1/6 weighting)
fails it."""

import math

import numpy as np
import pytest

from oracle import pic_oracle as O
from tests.helpers import species


# ---------------------------------------------------------------- electrostatics

def test_centered_gradient_of_a_single_mode_is_exact():
    # phi = sin(x + y + z) on the collocated interior: the centred difference of
    # sin is cos(theta) sin(d)/d exactly, with a different d per axis
    p = O.Params(N=(12, 10, 8), wind=(2 * math.pi, 2 * math.pi, 2 * math.pi), g=2, electrostatic=True)
    xs = [O.tiled_axis(p, a, "center") for a in range(3)]
    X, Y, Z = np.meshgrid(*xs, indexing="ij")
    theta = X + Y + Z
    phi = np.sin(theta)
    E = O.centered_gradient_E(phi, p)
    A = (slice(2, -2),) * 3
    for a in range(3):
        want = -np.cos(theta[A]) * math.sin(p.d[a]) / p.d[a]
        np.testing.assert_allclose(E[a][A], want, atol=1e-13)
    # ghosts are the periodic refresh of the interior
    np.testing.assert_allclose(E[0][0:2], E[0][12:14], atol=0)


def test_centered_gradient_with_a_conducting_axis_sees_constant_ghosts():
    # phi = a x with conducting x: interior E_x = -a, except on the two wall
    # planes where the ghost copies the wall value (a one-sided half difference)
    p = O.Params(N=(8, 4, 4), wind=(8.0, 4.0, 4.0), g=2, electrostatic=True, bc=(1, 0, 0))
    xc = O.tiled_axis(p, 0, "center")
    phi = np.broadcast_to((0.75 * xc)[:, None, None], p.shape).copy()
    E = O.centered_gradient_E(phi, p)
    Ex = E[0][2:-2, 2:-2, 2:-2]
    np.testing.assert_allclose(Ex[1:-1], -0.75, atol=1e-14)
    np.testing.assert_allclose(Ex[0], -0.75 / 2, atol=1e-14)
    np.testing.assert_allclose(Ex[-1], -0.75 / 2, atol=1e-14)
    np.testing.assert_allclose(E[1][2:-2, 2:-2, 2:-2], 0.0, atol=1e-14)


def test_zero_boundary_zeroes_global_tangential_faces():
    # ghost_cells_test.py:97-131: tile (2,2,2), g = 1, conducting everywhere
    p = O.Params(N=(2, 2, 2), wind=(2.0, 2.0, 2.0), g=1, bc=(1, 1, 1))
    Ex, Ey, Ez = (np.ones(p.shape) for _ in range(3))
    Ey = O.zero_boundary(Ey, p, 0)
    Ez = O.zero_boundary(Ez, p, 0)
    Ex = O.zero_boundary(Ex, p, 1)
    Ez = O.zero_boundary(Ez, p, 1)
    Ex = O.zero_boundary(Ex, p, 2)
    Ey = O.zero_boundary(Ey, p, 2)
    assert np.all(Ey[1, :, :] == 0.0) and np.all(Ez[1, :, :] == 0.0)
    assert np.all(Ex[:, 1, :] == 0.0) and np.all(Ez[:, 1, :] == 0.0)
    assert np.all(Ex[:, :, 1] == 0.0) and np.all(Ey[:, :, 1] == 0.0)
    # the last interior plane is the upper wall: zero too; the untouched
    # component keeps its interior (one cell here: index (2, 2, 2) of Ex)
    assert np.all(Ey[2, :, :] == 0.0)
    assert Ex[2, 2, 2] == 0.0 and Ez[2, 2, 2] == 0.0 and Ey[2, 2, 2] == 0.0


def test_zero_boundary_on_a_periodic_axis_is_the_refresh():
    # ghost_cells_test.py:122-131
    p = O.Params(N=(2, 2, 2), wind=(2.0, 2.0, 2.0), g=1)
    f = np.arange(64, dtype=float).reshape(p.shape)
    ref = O.refresh(f, p)
    np.testing.assert_array_equal(O.zero_boundary(ref, p, 0), ref)


def test_constant_boundary_copies_adjacent_interior_into_wall_ghosts():
    # ghost_cells_test.py:133-149, then an element-wise statement of the result
    p = O.Params(N=(2, 2, 2), wind=(2.0, 2.0, 2.0), g=1, bc=(1, 0, 0))
    f = np.arange(64, dtype=float).reshape(p.shape)
    out = O.constant_boundary(f, p, 0)
    np.testing.assert_array_equal(out[0], out[1])
    np.testing.assert_array_equal(out[-1], out[-2])
    np.testing.assert_array_equal(out[1:-1, 1:-1, 1:-1], f[1:-1, 1:-1, 1:-1])
    # element-wise: y, z wrap periodically over the 2-cell interior, x clamps
    L = 4
    for i in range(L):
        for j in range(L):
            for k in range(L):
                ii = min(max(i, 1), 2)
                jj = 1 + (j - 1) % 2
                kk = 1 + (k - 1) % 2
                assert out[i, j, k] == f[ii, jj, kk], (i, j, k)


def test_constant_boundary_on_a_periodic_axis_is_the_refresh():
    # ghost_cells_test.py:151-160
    p = O.Params(N=(2, 2, 2), wind=(2.0, 2.0, 2.0), g=1)
    f = np.arange(64, dtype=float).reshape(p.shape)
    np.testing.assert_array_equal(O.constant_boundary(f, p, 0), O.refresh(f, p))


def test_phi_refresh_for_a_conducting_axis():
    # electrostatic_yee_test.py:150-173: 8x4x4, g = 2, conducting x
    p = O.Params(N=(8, 4, 4), wind=(1.0, 1.0, 1.0), g=2, bc=(1, 0, 0), electrostatic=True)
    phi = np.zeros(p.shape)
    inner = np.arange(128, dtype=float).reshape(8, 4, 4)
    phi[2:-2, 2:-2, 2:-2] = inner
    out = O.phi_boundaries(phi, p)
    np.testing.assert_array_equal(out[:2], np.broadcast_to(out[2:3], out[:2].shape))
    np.testing.assert_array_equal(out[-2:], np.broadcast_to(out[-3:-2], out[-2:].shape))
    np.testing.assert_array_equal(out[2:-2, 2:-2, 2:-2], inner)


def test_phi_boundaries_sequence_on_two_conducting_axes():
    # _apply_tiled_phi_constant_boundaries applies apply_tiled_constant_boundary
    # per conducting axis in x, y, z order, and each call first refreshes every
    # axis, which zeroes the ghosts of non-periodic axes: only the last
    # conducting axis keeps copied ghosts.  Element-wise on a 3x2x2 interior.
    p = O.Params(N=(3, 2, 2), wind=(3.0, 2.0, 2.0), g=1, bc=(1, 1, 0), electrostatic=True)
    f = np.zeros(p.shape)
    f[1:-1, 1:-1, 1:-1] = np.arange(1.0, 13.0).reshape(3, 2, 2)
    out = O.phi_boundaries(f, p)
    X, Y, Z = p.shape
    for i in range(X):
        for j in range(Y):
            for k in range(Z):
                if i in (0, X - 1):
                    want = 0.0                               # x ghosts: zeroed by the y call's refresh
                else:
                    jj = min(max(j, 1), Y - 2)               # y ghosts: copies of the adjacent plane
                    kk = 1 + (k - 1) % 2                     # z periodic
                    want = f[i, jj, kk]
                assert out[i, j, k] == want, (i, j, k)


def test_tiled_electrostatic_zero_charge_keeps_a_constant_potential():
    # electrostatic_yee_test.py:175-222: one active particle of zero charge,
    # phi = 3 in the interior, conducting x: rho = 0, phi stays 3 with copied
    # wall ghosts (zero wall ghosts would make lapl(phi) != 0 and move phi),
    # E = 0 everywhere
    p = O.Params(N=(8, 4, 4), wind=(1.0, 1.0, 1.0), g=2, bc=(1, 0, 0), electrostatic=True)
    x = np.zeros((1, 1, 3))
    sp = species(1, charge=(0.0,), mass=(1.0,), weight=(1.0,))
    phi = np.zeros(p.shape)
    phi[2:-2, 2:-2, 2:-2] = 3.0
    E, phi_out, rho = O.electrostatic_fields(x, np.ones((1, 1), bool), sp, np.zeros(p.shape), phi, p)
    assert np.all(rho == 0.0)
    np.testing.assert_allclose(phi_out[2:-2, 2:-2, 2:-2], 3.0, atol=1e-12)
    np.testing.assert_allclose(phi_out[:2], np.broadcast_to(phi_out[2:3], phi_out[:2].shape), atol=1e-12)
    np.testing.assert_allclose(phi_out[-2:], np.broadcast_to(phi_out[-3:-2], phi_out[-2:].shape), atol=1e-12)
    for c in E:
        np.testing.assert_allclose(c, 0.0, atol=1e-12)


# ---------------------------------------------------------------- J_from_rhov face weights

def _single(xpos, v=(0.11, -0.17, 0.07)):
    return np.array([[[xpos, 0.0, 0.0]]]), np.array([[list(v)]], dtype=float)


@pytest.mark.parametrize("xpos", [-1.32, -1.05, -0.03, 0.6, 1.97])
def test_direct_current_face_weights_extrapolate_cic(xpos):
    """J_from_rhov.py:139-141 takes the face offset from the NODE anchor,
    df = dn - d/2, so for dn < d/2 the CIC face weights [1 - r, r] at the
    faces a, a+1 extrapolate (r < 0).  Weights that bracket the particle have
    the same moments 0 and 1 but a second central moment r'(1 - r') d^2 > 0;
    the extrapolating ones give r (1 - r) d^2 with r = df/d, negative here."""
    p = O.Params(N=(8, 1, 1), wind=(4.0, 1.0, 1.0), g=2, dt=0.05, shape_factor=1)
    x, u = _single(xpos)
    sp = species(1, charge=(-1.0,), weight=(0.5,))
    J = O.j_from_rhov_raw(x, u, np.ones((1, 1), bool), sp, p)
    d = p.d[0]
    org = O.tiled_axis(p, 0, "center")[0]
    anchor = math.floor((xpos - org) / d)
    r = ((xpos - org) - (anchor + 0.5) * d) / d
    xv = O.tiled_axis(p, 0, "vertex")
    qwv = -0.5 * u[0, 0, 0]
    dV = p.d[0] * p.d[1] * p.d[2]
    jx = J[0].sum(axis=(1, 2)) * dV
    assert abs(jx.sum() - qwv) < 1e-12
    m2 = float((jx * (xv - xpos) ** 2).sum())
    assert abs(m2 - qwv * r * (1 - r) * d * d) < 1e-12
    dn = (xpos - org) - anchor * d
    if dn < d / 2:
        assert r < 0 and m2 / qwv < 0           # extrapolation: negative "variance"
    # the transverse components use the (bracketing) node weights: variance rn (1 - rn) d^2 >= 0
    xc = O.tiled_axis(p, 0, "center")
    rn = dn / d
    jy = J[1].sum(axis=(1, 2)) * dV
    assert abs(float((jy * (xc - xpos) ** 2).sum()) - (-0.5 * u[0, 0, 1]) * rn * (1 - rn) * d * d) < 1e-12


@pytest.mark.parametrize("xpos", [-1.32, -1.05, -0.03, 0.6, 1.97])
def test_direct_current_face_weights_extrapolate_tsc(xpos):
    """TSC: the face offset r = df/d lies in [-1, 0] from the rounded node
    anchor; the quadratic-spline weights reproduce moments 0, 1 and 2 (variance
    d^2/4) for any r, and the third central moment is r (1/4 - r^2) d^3, which
    differs from the bracketed stencil's (r + 1)(1/4 - (r + 1)^2) d^3."""
    p = O.Params(N=(8, 1, 1), wind=(4.0, 1.0, 1.0), g=2, dt=0.05, shape_factor=2)
    x, u = _single(xpos)
    sp = species(1, charge=(-1.0,), weight=(0.5,))
    J = O.j_from_rhov_raw(x, u, np.ones((1, 1), bool), sp, p)
    d = p.d[0]
    org = O.tiled_axis(p, 0, "center")[0]
    anchor = int(np.round((xpos - org) / d))
    r = ((xpos - org) - (anchor + 0.5) * d) / d
    assert -1.0 <= r <= 0.0
    xv = O.tiled_axis(p, 0, "vertex")
    qwv = -0.5 * u[0, 0, 0]
    dV = p.d[0] * p.d[1] * p.d[2]
    jx = J[0].sum(axis=(1, 2)) * dV
    dx_ = xv - xpos
    assert abs(jx.sum() - qwv) < 1e-12
    assert abs((jx * dx_).sum()) < 1e-12
    assert abs((jx * dx_ ** 2).sum() - qwv * d * d / 4) < 1e-12
    assert abs((jx * dx_ ** 3).sum() - qwv * r * (0.25 - r * r) * d ** 3) < 1e-12


# ---------------------------------------------------------------- Esirkepov along reduced axes

@pytest.mark.parametrize("null", [0, 1, 2])
@pytest.mark.parametrize("sf", [1, 2])
def test_esirkepov_2d_null_axis_current_moments(null, sf):
    """2-D Esirkepov (Esirkepov.py:438-504): the current along the reduced axis
    is q w v_null W_null / dV with W_null = (S1 S1 + S0 S0)/3 + (S1 S0 + S0 S1)/6.
    Since the shapes reproduce linears, sum J dV = q w v and
    sum X Y J dV = q w v (x0 y0 + x1 y1)/3 + (x0 y1 + x1 y0)/6), i.e. the
    time average of x(t) y(t) along the straight path.  Continuity does not
    constrain this component."""
    N = [6, 6, 6]
    N[null] = 1
    p = O.Params(N=tuple(N), wind=tuple(0.5 * n for n in N), g=2, dt=0.05, shape_factor=sf)
    a1, a2 = [a for a in range(3) if a != null]
    rng = np.random.default_rng(31 + null + 3 * sf)
    sp = species(1, charge=(-1.0,), weight=(0.5,))
    for _ in range(4):
        xo = np.zeros(3)
        v = np.zeros(3)
        for a in (a1, a2):
            xo[a] = rng.uniform(-1.0, 1.0)
            v[a] = rng.uniform(-0.7, 0.7) * p.d[a] / p.dt
        v[null] = rng.uniform(-1, 1)
        J = O.esirkepov_raw(xo.reshape(1, 1, 3), v.reshape(1, 1, 3), np.ones((1, 1), bool), sp, p)
        dV = p.d[0] * p.d[1] * p.d[2]
        X1 = O.tiled_axis(p, a1, "center")
        X2 = O.tiled_axis(p, a2, "center")
        Jn = J[null].sum(axis=null) * dV                 # (L1, L2) plane
        qwv = -0.5 * v[null]
        assert abs(Jn.sum() - qwv) < 1e-12
        x0, x1 = xo[a1], xo[a1] + v[a1] * p.dt
        y0, y1 = xo[a2], xo[a2] + v[a2] * p.dt
        want = qwv * ((x0 * y0 + x1 * y1) / 3 + (x0 * y1 + x1 * y0) / 6)
        got = float((Jn * X1[:, None] * X2[None, :]).sum())
        assert abs(got - want) < 1e-12, (got, want)
        # the reduced-axis index is g
        assert np.all(np.delete(J[null], p.g, axis=null) == 0.0)


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("sf", [1, 2])
def test_esirkepov_1d_transverse_current_moments(axis, sf):
    """1-D Esirkepov (Esirkepov.py:412-435): a transverse component is
    q w v_t (S1 + S0)/2 / dV, so sum J dV = q w v_t and
    sum X J dV = q w v_t (x0 + x1)/2."""
    N = [1, 1, 1]
    N[axis] = 8
    p = O.Params(N=tuple(N), wind=tuple(0.5 * n for n in N), g=2, dt=0.05, shape_factor=sf)
    rng = np.random.default_rng(41 + axis + 3 * sf)
    sp = species(1, charge=(-1.0,), weight=(0.5,))
    X = O.tiled_axis(p, axis, "center")
    for _ in range(4):
        xo = np.zeros(3)
        v = rng.uniform(-1, 1, 3)
        xo[axis] = rng.uniform(-1.5, 1.5)
        v[axis] = rng.uniform(-0.7, 0.7) * p.d[axis] / p.dt
        J = O.esirkepov_raw(xo.reshape(1, 1, 3), v.reshape(1, 1, 3), np.ones((1, 1), bool), sp, p)
        dV = p.d[0] * p.d[1] * p.d[2]
        x0, x1 = xo[axis], xo[axis] + v[axis] * p.dt
        others = tuple(a for a in range(3) if a != axis)
        for c in others:
            line = J[c].sum(axis=others) * dV
            qwv = -0.5 * v[c]
            assert abs(line.sum() - qwv) < 1e-12
            assert abs((line * X).sum() - qwv * (x0 + x1) / 2) < 1e-12


# ---------------------------------------------------------------- mode="drop"

def test_rho_deposit_drops_past_the_upper_end_and_wraps_negative_indices():
    """TSC with g = 1: a particle 0.1 d below the upper face has anchor L-1 and
    a stencil point at index L, which is dropped (weight (1/2 - 0.1)^2 / 2 =
    0.08 of its charge); one 0.6 d below the lower face has a point at index
    -1, which NumPy-style normalisation sends to L-1 (weight 0.005)."""
    p = O.Params(N=(4, 1, 1), wind=(4.0, 1.0, 1.0), g=1, shape_factor=2)
    L = p.shape[0]
    sp = species(1, charge=(-1.0,), weight=(0.5,))
    act = np.ones((1, 1), bool)
    dV = p.d[0] * p.d[1] * p.d[2]
    hi = O.rho_raw(np.array([[[2.0 - 0.1, 0.0, 0.0]]]), act, sp, p)
    assert abs(hi.sum() * dV - (-0.5) * (1 - 0.08)) < 1e-12
    lo = O.rho_raw(np.array([[[-2.0 - 0.6, 0.0, 0.0]]]), act, sp, p)
    assert abs(lo.sum() * dV - (-0.5)) < 1e-12
    assert abs(lo[L - 1].sum() * dV - (-0.5) * 0.005) < 1e-12
    assert abs(lo[0].sum() * dV - (-0.5) * 0.59) < 1e-12        # 3/4 - 0.4^2 at the anchor
    assert abs(lo[1].sum() * dV - (-0.5) * 0.405) < 1e-12       # (1/2 + 0.4)^2 / 2


def test_esirkepov_transverse_deposit_drops_past_the_upper_end():
    """1-D TSC Esirkepov with g = 1 near the upper face: the transverse current
    (S1 + S0)/2 at index L is dropped, so the total falls short of q w v_y by
    the dropped mean weight."""
    p = O.Params(N=(4, 1, 1), wind=(4.0, 1.0, 1.0), g=1, shape_factor=2, dt=0.05)
    sp = species(1, charge=(-1.0,), weight=(0.5,))
    x0 = 2.0 - 0.1
    vx = 0.02 / p.dt                      # x1 = 1.92: r = -0.08 at the same anchor
    u = np.array([[[vx, 0.3, 0.0]]])
    J = O.esirkepov_raw(np.array([[[x0, 0.0, 0.0]]]), u, np.ones((1, 1), bool), sp, p)
    dropped = 0.5 * (0.5 * (0.5 - 0.1) ** 2 + 0.5 * (0.5 - 0.08) ** 2)
    dV = p.d[0] * p.d[1] * p.d[2]
    assert abs(J[1].sum() * dV - (-0.5 * 0.3) * (1 - dropped)) < 1e-12


def test_rolled_gradient_wraps_the_interior_on_a_conducting_axis():
    # calculate_electrostatic_fields differentiates the interior with jnp.roll:
    # phi = a x gives -a inside and a (n - 2)/2 on the two wall planes
    p = O.Params(N=(8, 4, 4), wind=(8.0, 4.0, 4.0), g=2, electrostatic=True, bc=(1, 0, 0))
    xc = O.tiled_axis(p, 0, "center")
    phi = np.broadcast_to((0.75 * xc)[:, None, None], p.shape).copy()
    E = O.roll_gradient_E(phi, p)
    Ex = E[0][2:-2, 2:-2, 2:-2]
    np.testing.assert_allclose(Ex[1:-1], -0.75, atol=1e-14)
    np.testing.assert_allclose(Ex[0], 0.75 * 6 / 2, atol=1e-13)
    np.testing.assert_allclose(Ex[-1], 0.75 * 6 / 2, atol=1e-13)
    # conducting x ghosts of E are zero after the field refresh
    assert np.all(E[0][:2] == 0.0) and np.all(E[0][-2:] == 0.0)


def test_untiled_electrostatic_fields_from_zero_charge():
    # electrostatic_yee_test.py:121-149: rho = phi = 0 gives phi = 0 and E = 0
    p = O.Params(N=(16, 16, 16), wind=(2 * math.pi,) * 3, g=2, electrostatic=True)
    E, phi, rho = O.electrostatic_fields_from_rho(np.zeros(p.shape), np.zeros(p.shape), p)
    assert np.all(phi == 0.0) and np.all(rho == 0.0)
    for c in E:
        assert np.all(c == 0.0)


def test_total_momentum_and_charge_closed_forms():
    p = O.Params(N=(6, 1, 1), wind=(3.0, 1.0, 1.0), g=2, shape_factor=1)
    u = np.array([[[0.3, 0.4, 0.0], [0.0, 0.0, 0.5]], [[1.0, 0.0, 0.0], [0.2, 0.0, 0.0]]])
    act = np.array([[True, True], [False, True]])
    sp = species(2, charge=(-1.0, 2.0), mass=(2.0, 5.0), weight=(0.5, 3.0))
    assert abs(O.compute_total_momentum(u, act, sp) - (1.0 * (0.5 + 0.5) + 15.0 * 0.2)) < 1e-15
    x = np.array([[[-1.3, 0, 0], [0.7, 0, 0]], [[0.1, 0, 0], [1.4, 0, 0]]])
    rho = O.compute_rho(x, act, sp, p)
    assert abs(O.total_charge(rho, p) - (-1.0 * 0.5 * 2 + 2.0 * 3.0 * 1)) < 1e-12
