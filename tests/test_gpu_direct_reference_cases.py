"""J_from_rhov on the particle sets of the reference's direct-deposition tests
(direct_deposition_test.py), each on one tile, against the FP64 oracle.

The reference compares its multi-tile deposit with its one-tile deposit on
these inputs; here the same inputs (positions, velocities, active flags,
species, grid, guard cells, shape factor, filter) go through the HIP path
and the oracle.  The cases cover second-order shapes with two guard cells,
reduced axes, the bilinear and digital filters, the active mask, particles
just outside the periodic faces, conducting field walls and two species.
"""

import numpy as np
import pytest

from oracle import pic_oracle as O
from pypic3d_amd import workloads
from pypic3d_amd._tensors import tiled_particles_from_arrays
from pypic3d_amd.deposition.J_from_rhov import J_from_rhov
from pypic3d_amd.particles.particle_class import SpeciesConfig
from tests.gpu_helpers import TOL, gpu_field, host, oracle_params, rel_err

pytestmark = pytest.mark.gpu

WIND = (4.0, 3.0, 2.0)      # direct_deposition_test.py:141-160


def _particles(slots, S):
    """(species, x, u, active) rows -> (S, P, 3) arrays, unused slots inactive."""
    P = max(sum(1 for r in slots if r[0] == s) for s in range(S))
    x = np.zeros((S, P, 3))
    u = np.zeros((S, P, 3))
    act = np.zeros((S, P), bool)
    fill = [0] * S
    for s, xx, uu, a in slots:
        k = fill[s]
        x[s, k], u[s, k], act[s, k] = xx, uu, a
        fill[s] += 1
    return x.astype(np.float32).astype(np.float64), u.astype(np.float32).astype(np.float64), act


_SIX = [(0, (-1.55, -1.10, -0.70), (0.18, 0.03, -0.06), True),       # :372-383
        (0, (-0.52, -0.55, -0.04), (-0.11, 0.17, 0.24), True),
        (0, (-0.03, -0.03, 0.03), (0.07, -0.22, 0.11), True),
        (0, (0.49, 0.02, 0.31), (-0.04, 0.19, -0.14), True),
        (0, (0.55, 0.52, 0.49), (0.21, -0.08, 0.05), True),
        (0, (1.45, 1.05, 0.72), (-0.16, 0.12, -0.19), True)]
_LINE = [(0, (-1.95, 0.0, 0.0), (0.18, 0.03, -0.06), True),          # :402-413
         (0, (-1.51, 0.0, 0.0), (-0.11, 0.17, 0.24), True),
         (0, (-1.02, 0.0, 0.0), (0.07, -0.22, 0.11), True),
         (0, (-0.48, 0.0, 0.0), (-0.04, 0.19, -0.14), True),
         (0, (0.02, 0.0, 0.0), (0.21, -0.08, 0.05), True),
         (0, (0.47, 0.0, 0.0), (-0.16, 0.12, -0.19), True),
         (0, (1.04, 0.0, 0.0), (0.09, -0.15, 0.16), True),
         (0, (1.88, 0.0, 0.0), (-0.13, 0.05, -0.07), True)]
_FOUR = [(0, (-1.25, -1.0, -0.65), (0.2, 0.0, -0.05), True),         # :809-815, :915-921
         (0, (-0.25, -0.25, -0.15), (-0.1, 0.15, 0.25), True),
         (0, (0.65, 0.35, 0.25), (0.05, -0.2, 0.1), True),
         (0, (1.45, 1.05, 0.75), (0.3, 0.1, -0.15), True)]
_MASK = [(0, r[1], r[2], a) for r, a in zip(_FOUR, (True, False, True, False))]
_TWO_SPECIES = [(0, (-1.25, -1.0, -0.65), (0.2, 0.0, -0.05), True),  # :724-732
                (1, (-1.65, 1.15, 0.35), (-0.1, 0.3, 0.1), True),
                (0, (-0.25, -0.25, -0.15), (-0.1, 0.15, 0.25), True),
                (1, (0.15, -0.75, -0.45), (0.2, -0.05, 0.05), True),
                (0, (0.65, 0.35, 0.25), (0.05, -0.2, 0.1), True),
                (1, (1.75, 0.45, 0.85), (-0.25, 0.15, -0.2), True),
                (0, (1.45, 1.05, 0.75), (0.3, 0.1, -0.15), True)]


def _crossing():                                                      # :931-948
    dx = WIND[0] / 10
    return [(0, (-WIND[0] / 2 - 0.2 * dx, 0.0, 0.0), (-0.25, -0.2, 0.15), True),
            (0, (WIND[0] / 2 + 0.1 * dx, 0.0, 0.0), (0.5, 0.1, 0.0), True)]


def _conducting():                                                    # :950-990
    d = (WIND[0] / 8, WIND[1] / 6, WIND[2] / 4)
    return [(0, (-WIND[0] / 2 + 0.1 * d[0], -WIND[1] / 2 + 0.1 * d[1], 0.0), (0.5, 0.1, -0.15), True),
            (0, (WIND[0] / 2 - 0.1 * d[0], 0.0, -WIND[2] / 2 + 0.1 * d[2]), (-0.25, -0.2, 0.35), True),
            (0, (0.0, WIND[1] / 2 - 0.1 * d[1], WIND[2] / 2 - 0.1 * d[2]), (0.15, 0.3, -0.1), True)]


# name: (N, g, shape factor, dt, filter, alpha, field bc, slots, (charges, masses, weights))
CASES = {
    "quadratic_two_guards": ((8, 6, 4), 2, 2, 0.05, "none", 1.0, (0, 0, 0), _SIX, ([-1.0], [1.0], [0.5])),
    "quadratic_reduced_axes": ((20, 1, 1), 2, 2, 0.05, "none", 1.0, (0, 0, 0), _LINE, ([-1.0], [1.0], [0.5])),
    "quadratic_reduced_axes_bilinear": ((20, 1, 1), 2, 2, 0.05, "bilinear", 1.0, (0, 0, 0), _LINE,
                                        ([-1.0], [1.0], [0.5])),
    "digital_filter": ((8, 6, 4), 1, 1, 0.05, "digital", 0.6, (0, 0, 0), _FOUR, ([-1.0], [1.0], [0.5])),
    "active_mask": ((8, 6, 4), 1, 1, 0.05, "none", 1.0, (0, 0, 0), _MASK, ([1.0], [1.0], [1.0])),
    "two_species": ((8, 6, 4), 1, 1, 0.05, "none", 1.0, (0, 0, 0), _TWO_SPECIES,
                    ([-1.0, 2.0], [1.0, 4.0], [0.5, 0.25])),
    "periodic_crossing": ((10, 6, 4), 1, 1, 0.0, "none", 1.0, (0, 0, 0), _crossing(), ([1.0], [1.0], [1.0])),
    "conducting_walls": ((8, 6, 4), 1, 1, 0.0, "none", 1.0, (1, 1, 1), _conducting(), ([1.0], [1.0], [1.0])),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_reference_direct_deposition_case(name):
    N_, g, sf, dt, filt, alpha, bc, slots, (q, m, w) = CASES[name]
    static, dyn = workloads.make_parameters(N_, wind=WIND, dt=dt, g=g, shape_factor=sf, current_deposition="direct",
                                            current_filter=filt, alpha=alpha, bc=bc)
    p = oracle_params(static, dyn)
    S = len(q)
    x, u, act = _particles(slots, S)
    cfg = SpeciesConfig(charge=np.array(q), mass=np.array(m), weight=np.array(w),
                        update_x=np.ones((S, 3), bool), update_u=np.ones((S, 3), bool))
    sp = O.Species(np.array(q), np.array(m), np.array(w), np.ones((S, 3), bool), np.ones((S, 3), bool))
    J0 = tuple(gpu_field(np.zeros(p.shape)) for _ in range(3))
    got = J_from_rhov(tiled_particles_from_arrays(x, u, act), cfg, J0, static, dyn)
    want = O.j_from_rhov(x, u, act, sp, p)
    assert max(float(np.max(np.abs(c))) for c in want) > 0.0
    for c in range(3):
        assert rel_err(host(got[c]), want[c]) <= TOL, (name, c)
    if name == "active_mask":
        keep = act.copy()
        want_active = O.j_from_rhov(x[:, keep[0]], u[:, keep[0]], act[:, keep[0]], sp, p)
        for c in range(3):
            np.testing.assert_allclose(want[c], want_active[c], rtol=0, atol=1e-15)
