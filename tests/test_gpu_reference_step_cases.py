"""One electrodynamic step on the reference's own single-particle cases
(evolve_test.py:75-157, 217-292) and the periodic wrap known answers of
grid_and_stencil_test.py:27-33, through the HIP path."""

import numpy as np
import pytest
import torch

from pypic3d_amd import workloads
from pypic3d_amd._tensors import tiled_particles_from_arrays
from pypic3d_amd.evolve import time_loop_electrodynamic
from pypic3d_amd.particles.particle_class import SpeciesConfig
from pypic3d_amd.particles.particle_tile_communication import refresh_tiled_particle_tiles
from tests.gpu_helpers import host

pytestmark = pytest.mark.gpu


def _one(x, u, active=True):
    cfg = SpeciesConfig(charge=np.array([1.0]), mass=np.array([1.0]), weight=np.array([1.0]),
                        update_x=np.ones((1, 3), bool), update_u=np.ones((1, 3), bool))
    parts = tiled_particles_from_arrays(np.array([[x]], np.float32), np.array([[u]], np.float32),
                                        np.array([[active]]))
    return parts, cfg


def test_external_electric_field_pushes_without_evolving_the_fields():
    """3^3 cells of 1/3, dt 0.1, direct deposit, non-relativistic Boris, a
    resting unit particle at the origin, external Ex = 1 everywhere: after
    one step vx = q E dt / m = 0.1, vy = vz = 0, the external field is
    unchanged and nothing overflows."""
    static, dyn = workloads.make_parameters((3, 3, 3), wind=(1.0, 1.0, 1.0), dt=0.1, relativistic=False,
                                            current_deposition="direct")
    fields = workloads.empty_fields(static, dyn)
    ext_E = (torch.ones_like(fields[0][0]), torch.zeros_like(fields[0][1]), torch.zeros_like(fields[0][2]))
    ext_B = tuple(torch.zeros_like(c) for c in fields[1])
    fields = fields[:5] + ((ext_E, ext_B),) + fields[6:]
    parts, cfg = _one((0.0, 0.0, 0.0), (0.0, 0.0, 0.0))
    parts, out = time_loop_electrodynamic(parts, cfg, fields, static, dyn)
    u = host(parts.u).reshape(-1, 3)
    assert u[0, 0] == pytest.approx(0.1, rel=1e-6)
    assert u[0, 1] == 0.0 and u[0, 2] == 0.0
    assert torch.equal(out[5][0][0], ext_E[0])
    assert out[6] is None
    assert not bool(out[7])


def test_absorbing_wall_deactivates_a_particle_within_the_step():
    """3 x 1 x 1, absorbing x walls, a particle at x = 0.49 moving at 0.2
    with dt 0.1 crosses x = 0.5 during the step and comes out inactive."""
    static, dyn = workloads.make_parameters((3, 1, 1), wind=(1.0, 1.0, 1.0), dt=0.1, relativistic=False,
                                            current_deposition="direct", pbc=(2, 0, 0))
    parts, cfg = _one((0.49, 0.0, 0.0), (0.2, 0.0, 0.0))
    parts, _ = time_loop_electrodynamic(parts, cfg, workloads.empty_fields(static, dyn), static, dyn)
    assert parts.x.reshape(-1, 3).shape[0] == 1
    assert host(parts.active).reshape(-1).tolist() == [False]


def test_periodic_wrap_known_answers():
    static, dyn = workloads.make_parameters((8, 1, 1), wind=(4.0, 1.0, 1.0), dt=0.1)
    xs = [-2.5, -2.0, 1.75, 2.0, 2.25, 6.5]
    x = np.zeros((1, len(xs), 3), np.float32)
    x[0, :, 0] = xs
    parts = tiled_particles_from_arrays(x, np.zeros_like(x), np.ones((1, len(xs)), bool))
    out, overflow = refresh_tiled_particle_tiles(parts, static, dyn)
    np.testing.assert_array_equal(host(out.x).reshape(-1, 3)[:, 0], [1.5, -2.0, 1.75, 2.0, -1.75, -1.5])
    assert not bool(overflow)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_wrap_periodic_position_known_answers(dtype):
    # grid_and_stencil_test.py:27-33: W = 4
    from pypic3d_amd.boundary_conditions.grid_and_stencil import wrap_periodic_position
    x = torch.tensor([-2.5, -2.0, 1.75, 2.0, 2.25, 6.5], dtype=dtype, device="cuda")
    got = wrap_periodic_position(x, 4.0).cpu().numpy()
    np.testing.assert_array_equal(got, [1.5, -2.0, 1.75, 2.0, -1.75, -1.5])
    assert wrap_periodic_position(x, 4.0).dtype == dtype
