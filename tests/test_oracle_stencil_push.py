"""Pins for the oracle's grids, shapes, gather and pushers (CPU only).

Each check is an independent property (closed form, exact reproduction,
conservation law, reference known-answer), not a retyping of the formula.
"""

import math

import numpy as np
import pytest

from oracle import pic_oracle as O
from tests.helpers import fp32


def test_tiled_axis_index_map():
    # grids.py:80-111: local l <-> physical index l-g; center collocated at
    # -W/2 + i d, vertex staggered at -W/2 + (i+1/2) d.
    p = O.Params(N=(8, 6, 4), wind=(4.0, 3.0, 2.0), g=2)
    c = O.tiled_axis(p, 0, "center")
    v = O.tiled_axis(p, 0, "vertex")
    assert len(c) == 12
    assert c[2] == -2.0 and c[3] == -1.5 and c[-3] == 1.5
    assert v[2] == -1.75 and v[-3] == 1.75
    # electrostatic solver: both lines collocated (grids.py:14-39)
    pe = O.Params(N=(8, 6, 4), wind=(4.0, 3.0, 2.0), g=2, electrostatic=True)
    assert np.array_equal(O.tiled_axis(pe, 0, "vertex"), c)


def test_anchor_rounding_rules():
    # CIC floors, TSC rounds half-to-even (jnp.round semantics).
    assert O.compute_anchor(np.array([2.5, 3.5, -0.5]), 0.0, 1.0, 2).tolist() == [2, 4, 0]
    assert O.compute_anchor(np.array([2.5, 3.999, -0.5]), 0.0, 1.0, 1).tolist() == [2, 3, -1]


@pytest.mark.parametrize("sf", [1, 2])
def test_shape_partition_of_unity_and_first_moment(sf):
    rng = np.random.default_rng(1)
    d = 0.37
    delta = rng.uniform(-0.5 * d, 0.5 * d, 100) if sf == 2 else rng.uniform(0, d, 100)
    w = O.shape_weights(delta, d, sf)
    np.testing.assert_allclose(w.sum(axis=0), 1.0, atol=1e-14)
    # first moment: sum_k w_k * k*d == delta (both shapes reproduce linears)
    np.testing.assert_allclose((w * (np.array([-1, 0, 1])[:, None] * d)).sum(axis=0), delta, atol=1e-14)
    w0 = O.shape_weights(np.zeros(1), d, sf)[:, 0]
    assert w0.tolist() == ([0.0, 1.0, 0.0] if sf == 1 else [0.125, 0.75, 0.125])


@pytest.mark.parametrize("sf", [1, 2])
@pytest.mark.parametrize("N", [(8, 6, 4), (8, 1, 1), (6, 6, 1)])
def test_gather_reproduces_constant_and_linear_fields(sf, N):
    # boris_test.py:69-129: constants are reproduced exactly; linear fields
    # are reproduced exactly by CIC and TSC away from the periodic seam.
    p = O.Params(N=N, wind=tuple(float(n) for n in N), g=2, shape_factor=sf)
    lines = O.component_lines(p)
    rng = np.random.default_rng(2)
    npart = 50
    pos = [rng.uniform(-p.wind[a] / 2 + 1.0, p.wind[a] / 2 - 1.0, npart) if p.N[a] > 1
           else np.zeros(npart) for a in range(3)]
    for name, ln in lines.items():
        const = np.full(p.shape, 3.25)
        np.testing.assert_allclose(O.interpolate(const, *pos, ln, p), 3.25, rtol=1e-14)
        a = 0
        lin = np.broadcast_to((2.0 * ln[a] + 1.0)[:, None, None], p.shape).copy()
        np.testing.assert_allclose(O.interpolate(lin, *pos, ln, p), 2.0 * pos[a] + 1.0, atol=1e-12)


def test_gather_reduced_axis_matches_collapsed_1d():
    # boris_test.py:92-129: a reduced axis collapses to index g with summed weights.
    p = O.Params(N=(8, 1, 1), wind=(4.0, 1.0, 1.0), g=2, shape_factor=2)
    rng = np.random.default_rng(3)
    f = rng.standard_normal(p.shape)
    x = rng.uniform(-1.5, 1.5, 20)
    y = rng.uniform(-0.3, 0.3, 20)
    z = rng.uniform(-0.3, 0.3, 20)
    ln = O.component_lines(p)["Ey"]
    got = O.interpolate(f, x, y, z, ln, p)
    line = f[:, 2, 2]
    origin, sp = ln[0][0], ln[0][1] - ln[0][0]
    a = O.compute_anchor(x, origin, sp, 2)
    w = O.shape_weights(x - (a * sp + origin), sp, 2)
    want = sum(line[(a + k) % len(line)] * w[k + 1] for k in (-1, 0, 1))
    np.testing.assert_allclose(got, want, rtol=1e-13)


def test_boris_pure_magnetic_rotation_and_pure_electric_kick():
    rng = np.random.default_rng(4)
    v = rng.standard_normal((3, 10)) * 0.1
    B = np.zeros((3, 10)); B[2] = 2.0
    E = np.zeros((3, 10))
    q, m, dt = 1.0, 1.0, 0.1
    out = O.boris(v, E, B, q, m, dt)
    np.testing.assert_allclose(np.linalg.norm(out, axis=0), np.linalg.norm(v, axis=0), rtol=1e-14)
    # rotation angle about z is 2*atan(q B dt / 2m), clockwise for q>0
    ang = np.arctan2(out[1], out[0]) - np.arctan2(v[1], v[0])
    ang = (ang + np.pi) % (2 * np.pi) - np.pi
    np.testing.assert_allclose(ang, -2 * math.atan(q * 2.0 * dt / (2 * m)), atol=1e-12)
    np.testing.assert_allclose(out[2], v[2], atol=1e-15)
    E[0] = 0.7
    out = O.boris(v, E, np.zeros((3, 10)), q, m, dt)
    np.testing.assert_allclose(out[0], v[0] + q * 0.7 * dt / m, atol=1e-14)


def _rel_step(v, E, B, dt, steps, C=1.0, Efun=None):
    v = np.asarray(v, dtype=np.float64).reshape(3, 1)
    for k in range(steps):
        e = Efun(k) if Efun else E
        v = O.relativistic_boris(v, np.asarray(e, dtype=np.float64).reshape(3, 1),
                                 np.asarray(B, dtype=np.float64).reshape(3, 1), 1.0, 1.0, dt, C)
    return v[:, 0]


def _gamma(v):
    return 1.0 / math.sqrt(1.0 - float(np.dot(v, v)))


def test_relativistic_boris_schmitz_constant_B_preserves_energy():
    # pusher_schmitz_test.py:69-84
    for gamma0 in (1.001, 10.0):
        u0 = math.sqrt(gamma0 ** 2 - 1.0)
        v0 = np.array([u0 / gamma0, 0.0, 0.0])
        v = _rel_step(v0, (0, 0, 0), (0, 0, 1.0), 0.2 * math.pi * gamma0, 10)
        assert abs(_gamma(v) - gamma0) < 2e-12


def test_relativistic_boris_schmitz_force_free_weakness_bounds():
    # pusher_schmitz_test.py:85-158 (Boris half)
    gamma0 = 10.0
    v0 = math.sqrt(gamma0 ** 2 - 1.0) / gamma0
    E, B, dt = (0.0, v0, 0.0), (0.0, 0.0, 1.0), 0.2 * math.pi * gamma0
    v = np.array([v0, 0.0, 0.0])
    ang = en = 0.0
    for _ in range(20):
        v = _rel_step(v, E, B, dt, 1)
        ang = max(ang, abs(math.atan2(v[1], v[0])))
        en = max(en, abs((_gamma(v) - gamma0) / (gamma0 - 1.0)))
    assert 1e-3 < ang < 2e-1 and en > 1e-2


def test_relativistic_boris_schmitz_boosted_gyration():
    # pusher_schmitz_test.py:120-183: exact lab-frame velocity of a boosted orbit
    gp, gf = 10.0, 5.0
    bf = math.sqrt(1.0 - 1.0 / gf ** 2)
    vp = math.sqrt(1.0 - 1.0 / gp ** 2)
    dt = 1.0e-3 * 2.0 * math.pi * gp * gf
    v = _rel_step((bf, vp / gf, 0.0), (0.0, bf * gf, 0.0), (0.0, 0.0, gf), dt, 20)
    up = math.sqrt(gp ** 2 - 1.0)
    omega, radius = 1.0 / gp, (up / gp) / (1.0 / gp)
    t_lab = 20 * dt
    tm = t_lab / gf
    for _ in range(8):
        res = gf * (tm + bf * radius * (1.0 - math.cos(omega * tm))) - t_lab
        der = gf * (1.0 + bf * vp * math.sin(omega * tm))
        tm -= res / der
    vxm, vym = vp * math.sin(omega * tm), vp * math.cos(omega * tm)
    den = 1.0 + bf * vxm
    exact = np.array([(vxm + bf) / den, vym / (gf * den), 0.0])
    assert np.linalg.norm(v - exact) < 1e-3


def test_relativistic_boris_schmitz_oscillating_parallel_E():
    # pusher_schmitz_test.py:185-211
    gperp = 1.1
    up = math.sqrt(gperp ** 2 - 1.0)
    v0 = np.array([up / gperp, 0.0, 0.0])
    omega0 = 0.5 * (1.0 / gperp)
    E0 = 10.0 * omega0
    dt = (2 * math.pi / omega0) / 100
    v = _rel_step(v0, None, (0, 0, 1.0), dt, 500,
                  Efun=lambda k: (0.0, 0.0, E0 * math.cos(omega0 * (k + 0.5) * dt)))
    assert abs(v[2]) < 1e-10 and abs(_gamma(v) - gperp) < 1e-10


def test_particle_push_respects_active_and_update_flags():
    # pusher_test.py:293-332: inactive slots and frozen components keep v.
    p = O.Params(N=(8, 6, 4), wind=(8.0, 6.0, 4.0), g=2, dt=0.1)
    rng = np.random.default_rng(5)
    x = fp32(rng.uniform(-1.5, 1.5, (2, 8, 3)))
    u = fp32(rng.uniform(-0.2, 0.2, (2, 8, 3)))
    active = np.ones((2, 8), bool); active[0, 3] = False
    sp = O.Species(np.array([-1.0, 1.0]), np.array([1.0, 2.0]), np.array([1.0, 1.0]),
                   np.ones((2, 3), bool), np.array([[True, True, True], [True, False, True]]))
    E = [np.full(p.shape, 0.5)] * 3
    B = [np.full(p.shape, 0.25)] * 3
    out = O.particle_push(x, u, active, sp, E, B, p)
    assert np.array_equal(out[0, 3], u[0, 3])
    assert np.array_equal(out[1, :, 1], u[1, :, 1])
    assert not np.allclose(out[1, :, 0], u[1, :, 0])


def _hc_step(v, E, B, dt, steps, C=1.0, Efun=None, q=1.0, m=1.0):
    v = np.asarray(v, dtype=np.float64).reshape(3, 1)
    for k in range(steps):
        e = Efun(k) if Efun else E
        v = O.higuera_cary(v, np.asarray(e, dtype=np.float64).reshape(3, 1),
                           np.asarray(B, dtype=np.float64).reshape(3, 1), q, m, dt, C)
    return v[:, 0]


def test_higuera_cary_schmitz_constant_B_preserves_energy():
    # pusher_schmitz_test.py:69-84 (Higuera-Cary half)
    for gamma0 in (1.001, 10.0):
        u0 = math.sqrt(gamma0 ** 2 - 1.0)
        v = _hc_step((u0 / gamma0, 0.0, 0.0), (0, 0, 0), (0, 0, 1.0), 0.2 * math.pi * gamma0, 10)
        assert abs(_gamma(v) - gamma0) < 2e-12


def test_higuera_cary_schmitz_force_free_is_exact():
    # pusher_schmitz_test.py:85-120: E = -v0 x B keeps the straight orbit to
    # machine precision (a dropped or mis-signed term in the gamma expression
    # or the rotation breaks this by many orders of magnitude).
    gamma0 = 10.0
    v0 = math.sqrt(gamma0 ** 2 - 1.0) / gamma0
    E, B, dt = (0.0, v0, 0.0), (0.0, 0.0, 1.0), 0.2 * math.pi * gamma0
    v = np.array([v0, 0.0, 0.0])
    ang = en = 0.0
    for _ in range(20):
        v = _hc_step(v, E, B, dt, 1)
        ang = max(ang, abs(math.atan2(v[1], v[0])))
        en = max(en, abs((_gamma(v) - gamma0) / (gamma0 - 1.0)))
    assert ang < 1e-13 and en < 1e-12


def test_higuera_cary_schmitz_boosted_gyration():
    # pusher_schmitz_test.py:120-183 (Higuera-Cary bound 1e-6)
    gp, gf = 10.0, 5.0
    bf = math.sqrt(1.0 - 1.0 / gf ** 2)
    vp = math.sqrt(1.0 - 1.0 / gp ** 2)
    dt = 1.0e-3 * 2.0 * math.pi * gp * gf
    v = _hc_step((bf, vp / gf, 0.0), (0.0, bf * gf, 0.0), (0.0, 0.0, gf), dt, 20)
    up = math.sqrt(gp ** 2 - 1.0)
    omega, radius = 1.0 / gp, (up / gp) / (1.0 / gp)
    t_lab = 20 * dt
    tm = t_lab / gf
    for _ in range(8):
        res = gf * (tm + bf * radius * (1.0 - math.cos(omega * tm))) - t_lab
        der = gf * (1.0 + bf * vp * math.sin(omega * tm))
        tm -= res / der
    vxm, vym = vp * math.sin(omega * tm), vp * math.cos(omega * tm)
    den = 1.0 + bf * vxm
    exact = np.array([(vxm + bf) / den, vym / (gf * den), 0.0])
    assert np.linalg.norm(v - exact) < 1e-6


def test_higuera_cary_schmitz_oscillating_parallel_E():
    # pusher_schmitz_test.py:185-211 (Higuera-Cary half)
    gperp = 1.1
    up = math.sqrt(gperp ** 2 - 1.0)
    omega0 = 0.5 * (1.0 / gperp)
    E0 = 10.0 * omega0
    dt = (2 * math.pi / omega0) / 100
    v = _hc_step((up / gperp, 0.0, 0.0), None, (0, 0, 1.0), dt, 500,
                 Efun=lambda k: (0.0, 0.0, E0 * math.cos(omega0 * (k + 0.5) * dt)))
    assert abs(v[2]) < 1e-10 and abs(_gamma(v) - gperp) < 1e-10


def test_higuera_cary_single_particle_known_answers():
    # boris_test.py:185-247 with C = 10.
    C = 10.0
    v = _hc_step((0.2, -0.1, 0.05), (0, 0, 0), (0, 0, 0), 0.1, 1, C=C)
    np.testing.assert_allclose(v, [0.2, -0.1, 0.05], rtol=1e-14)
    v = _hc_step((0.2, 0.0, 0.0), (0, 0, 0), (0, 0, 1.0), 0.1, 1, C=C)
    assert abs(np.linalg.norm(v) - 0.2) < 1e-12 and v[1] < 0
    v = _hc_step((0.0, 0.0, 0.0), (1.0, 0, 0), (0, 0, 0), 0.1, 1, C=C)
    assert v[0] > 0 and v[1] == 0 and v[2] == 0
    # pure E from rest: u = q E dt / m exactly, v = u / sqrt(1 + u^2/C^2)
    np.testing.assert_allclose(v[0], 0.1 / math.sqrt(1 + 0.01 / C ** 2), rtol=1e-15)
    q = np.array([1.0, 2.0])
    out = O.higuera_cary(np.zeros((3, 2)), np.array([[1.0, 1.0], [0, 0], [0, 0]]), np.zeros((3, 2)),
                         q, np.ones(2), 0.1, C)
    assert out[0, 1] > out[0, 0] and not out[1:].any()


def test_higuera_cary_nonrelativistic_limit_matches_boris():
    # As C -> inf both relativistic maps reduce to the classical Boris rotation.
    rng = np.random.default_rng(9)
    v = rng.standard_normal((3, 16)) * 0.1
    E = rng.standard_normal((3, 16)); B = rng.standard_normal((3, 16))
    a = O.higuera_cary(v, E, B, -1.0, 1.0, 0.05, 1e8)
    b = O.boris(v, E, B, -1.0, 1.0, 0.05)
    np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-12)
