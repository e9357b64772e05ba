"""Oracle pins for the PML (boundary_conditions/PML.py): profile values at
hand-computed indices and the load-time validation, the closed form of the
ADE memory, sigma = 0 as the identity, and absorption of a 1-D pulse."""

import math

import numpy as np
import pytest

from oracle import pic_oracle as O


def test_profiles_ramp_outwards_from_the_interface_with_hand_computed_values():
    p = O.Params(N=(8, 3, 2), wind=(8.0, 3.0, 2.0))
    layers = O.pml_layers([{"wall": "-x", "thickness": 3, "order": 2.0, "sigma_max": 9.0},
                           {"wall": "+x", "thickness": 3, "order": 2.0, "sigma_max": 9.0}], p)
    sx, sy, sz = O.pml_profiles(layers, p)
    assert sx.shape == (10, 5, 4)
    # sigma_max ((i+1)/T)^2 = 1, 4, 9 from the interface to the outer wall
    np.testing.assert_array_equal(sx[:, 1, 1], [0, 9, 4, 1, 0, 0, 1, 4, 9, 0])
    assert not sy.any() and not sz.any()
    # the one-ghost profile sits at offset g - 1 of a field-shaped array
    gx = O.pml_ghosted(sx, p)
    assert gx.shape == p.shape
    np.testing.assert_array_equal(gx[1:11, 2, 2], sx[:, 1, 1])
    # interior cell c of the field array (index g + c) reads profile index c + 1
    np.testing.assert_array_equal(gx[2:10, 2, 2], [9, 4, 1, 0, 0, 1, 4, 9])


def test_default_sigma_max_and_validation():
    p = O.Params(N=(10, 4, 4), wind=(5.0, 4.0, 4.0), C=2.0)
    (_, axis, T, order, smax), = O.pml_layers({"wall": "+x", "thickness": 4}, p)
    assert (axis, T, order) == (0, 4, 3.0)
    assert math.isclose(smax, -(4.0 * 2.0 * math.log(1e-8)) / (2.0 * 4 * 0.5), rel_tol=1e-15)
    with pytest.raises(ValueError, match="Invalid PML wall"):
        O.pml_layers({"wall": "x", "thickness": 2}, p)
    with pytest.raises(ValueError, match="Duplicate PML wall"):
        O.pml_layers([{"wall": "-y", "thickness": 2}, {"wall": "-y", "thickness": 1}], p)
    with pytest.raises(ValueError, match="exceeds active cells"):
        O.pml_layers({"wall": "-z", "thickness": 5}, p)
    with pytest.raises(ValueError, match="must be positive"):
        O.pml_layers({"wall": "-z", "thickness": 0}, p)


def test_memory_of_a_constant_derivative_follows_the_closed_form():
    rng = np.random.default_rng(0)
    d = rng.standard_normal(50)
    sigma = rng.uniform(0, 40, 50)
    dt = 0.03
    m = np.zeros(50)
    b = np.exp(-sigma * dt)
    for n in range(1, 12):
        out, m = O.pml_stretch(d, m, sigma, dt)
        np.testing.assert_allclose(m, (b ** n - 1.0) * d, rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(out, b ** n * d, rtol=1e-12, atol=1e-15)


def test_zero_profiles_leave_the_yee_updates_unchanged():
    p = O.Params(N=(6, 5, 4), wind=(3.0, 2.5, 2.0), dt=0.05)
    rng = np.random.default_rng(2)
    E = [O.refresh(rng.standard_normal(p.shape), p) for _ in range(3)]
    B = [O.refresh(rng.standard_normal(p.shape), p) for _ in range(3)]
    J = [O.refresh(rng.standard_normal(p.shape), p) for _ in range(3)]
    zeros = [np.zeros(p.shape) for _ in range(3)]
    mem = [np.zeros(p.N) for _ in range(6)]
    Bp, mb = O.update_B(E, B, p, pml=(mem, zeros))
    Ep, me = O.update_E(E, B, J, p, pml=(mem, zeros))
    for a, b in zip(Bp, O.update_B(E, B, p)):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(Ep, O.update_E(E, B, J, p)):
        np.testing.assert_array_equal(a, b)
    assert not any(m.any() for m in mb + me)


def _pulse_energy(with_pml, steps=120):
    """The particle-free 1-D setup of pml_test.py:527-617: Nx = 40, dx = 1/40,
    dt = 0.5/40, conducting x walls, Ey = Bz = exp(-((x + 0.3)/0.04)^2) on the
    vertex line, PML of 8 cells, order 3, sigma_max 60 at both ends, one
    update_E and one (half-step) update_B per step as that test steps.  The
    pulse travels in +x and enters the +x layer after about 70 steps, so the
    check runs 120 steps (the reference test stops at 60)."""
    p = O.Params(N=(40, 1, 1), wind=(1.0, 1.0, 1.0), d=(1.0 / 40, 1.0, 1.0), dt=0.5 / 40, bc=(1, 0, 0))
    g = p.g
    xv = -0.5 + (np.arange(40) + 0.5) / 40          # vertex line of the interior cells
    pulse = np.exp(-((xv + 0.30) / 0.04) ** 2)
    E = [np.zeros(p.shape) for _ in range(3)]
    B = [np.zeros(p.shape) for _ in range(3)]
    J = [np.zeros(p.shape) for _ in range(3)]
    E[1][g:-g, g, g] = pulse
    B[2][g:-g, g, g] = pulse
    E = [O.refresh(c, p) for c in E]
    B = [O.refresh(c, p) for c in B]
    prof = [O.pml_ghosted(pr, p) for pr in O.pml_profiles(
        O.pml_layers([{"wall": "-x", "thickness": 8, "order": 3.0, "sigma_max": 60.0},
                      {"wall": "+x", "thickness": 8, "order": 3.0, "sigma_max": 60.0}], p), p)]
    me = [np.zeros(p.N) for _ in range(6)]
    mb = [np.zeros(p.N) for _ in range(6)]

    def energy(E, B):
        A = slice(g, -g)
        dV = p.d[0] * p.d[1] * p.d[2]
        return 0.5 * sum(np.sum(c[A, A, A] ** 2) for c in E) * dV + 0.5 * sum(np.sum(c[A, A, A] ** 2) for c in B) * dV

    e0 = energy(E, B)
    for _ in range(steps):
        if with_pml:
            E, me = O.update_E(E, B, J, p, pml=(me, prof))
            B, mb = O.update_B(E, B, p, pml=(mb, prof))
        else:
            E = O.update_E(E, B, J, p)
            B = O.update_B(E, B, p)
    return e0, energy(E, B)


def test_pml_absorbs_a_pulse_that_conducting_walls_alone_keep():
    e0, e_pml = _pulse_energy(True)
    _, e_wall = _pulse_energy(False)
    assert np.isfinite(e_pml) and e_pml < 0.1 * e0
    assert e_wall > 0.9 * e0
