"""GPU parity: every HIP kernel (through the C ABI) against the FP64 oracle on
identical FP32-representable inputs.

Tolerance: norm-wise relative error max|gpu - oracle| / max|oracle| <= 1e-5
(TOL) for floating point; cell keys and sort permutations bit-exact; ghost
refresh (a copy) bit-exact.
"""

import numpy as np
import pytest
import torch

from oracle import pic_oracle as O
from pypic3d_amd import workloads
from pypic3d_amd._tensors import tiled_particles_from_arrays
from pypic3d_amd.boundary_conditions import ghost_cells as gc
from pypic3d_amd.deposition.Esirkepov import Esirkepov_current
from pypic3d_amd.deposition.J_from_rhov import J_from_rhov
from pypic3d_amd.deposition.rho import compute_rho
from pypic3d_amd.particles.particle_tile_communication import (refresh_tiled_particle_tiles,
                                                               update_tiled_particle_positions)
from pypic3d_amd.pusher.particle_push import particle_push
from pypic3d_amd.solvers.electrostatic_yee import solve_poisson_with_conjugate_gradient
from pypic3d_amd.solvers.first_order_yee import update_B, update_E
from pypic3d_amd.sorting import cell_keys, sort_particles_by_cell
from pypic3d_amd.utilities.filters import bilinear_filter, digital_filter
from pypic3d_amd.utils import compute_energy
from tests.gpu_helpers import TOL, gpu_field, host, host_particles, oracle_params, oracle_species, rel_err
from tests.helpers import random_field, random_particles, species

pytestmark = pytest.mark.gpu

DIMS = [(8, 1, 1), (1, 8, 1), (1, 1, 8), (8, 8, 1), (8, 1, 8), (1, 8, 8), (6, 6, 6), (12, 10, 8)]


def _setup(N, sf=1, rel=True, dep="esirkepov", filt="none", bc=(0, 0, 0), pbc=(0, 0, 0), alpha=1.0, wind=None,
           electrostatic=False, pusher="boris"):
    wind = wind or tuple(0.5 * n for n in N)
    static, dyn = workloads.make_parameters(N, wind=wind, dt=0.05, shape_factor=sf, relativistic=rel,
                                            current_deposition=dep, current_filter=filt, bc=bc, pbc=pbc,
                                            alpha=alpha, electrostatic=electrostatic, particle_pusher=pusher)
    return static, dyn, oracle_params(static, dyn)


def _cfg(sp):
    from pypic3d_amd.particles.particle_class import SpeciesConfig
    return SpeciesConfig(sp.charge, sp.mass, sp.weight, sp.update_x, sp.update_u)


# ---------------------------------------------------------------- fields

@pytest.mark.parametrize("N,bc,alpha", [((24, 20, 16), (0, 0, 0), 1.0), ((32, 1, 1), (0, 0, 0), 1.0),
                                         ((16, 12, 1), (0, 0, 0), 0.7), ((12, 10, 8), (1, 0, 1), 1.0),
                                         ((10, 9, 8), (1, 1, 1), 0.6)])
def test_yee_update_B_and_E(N, bc, alpha):
    static, dyn, p = _setup(N, bc=bc, alpha=alpha)
    E = [random_field(p, 1 + c, refresh=False) for c in range(3)]
    B = [random_field(p, 4 + c, refresh=False) for c in range(3)]
    J = [random_field(p, 7 + c, scale=0.3) for c in range(3)]
    for do_filter in (False, True):
        got, st = update_B(tuple(gpu_field(e) for e in E), tuple(gpu_field(b) for b in B), static, dyn,
                           do_filter=do_filter)
        want = O.update_B(E, B, p, do_filter=do_filter)
        assert st is None
        for c in range(3):
            assert rel_err(host(got[c]), want[c]) <= TOL
    got, _ = update_E(tuple(gpu_field(e) for e in E), tuple(gpu_field(b) for b in B), tuple(gpu_field(j) for j in J),
                      static, dyn)
    want = O.update_E(E, B, J, p)
    for c in range(3):
        assert rel_err(host(got[c]), want[c]) <= TOL


def test_yee_on_the_128_cubed_weibel_grid():
    static, dyn, p = _setup((128, 128, 128), wind=(128.0, 128.0, 128.0))
    E = [random_field(p, 11 + c) for c in range(3)]
    B = [random_field(p, 14 + c) for c in range(3)]
    J = [random_field(p, 17 + c, scale=0.1) for c in range(3)]
    gB, _ = update_B(tuple(map(gpu_field, E)), tuple(map(gpu_field, B)), static, dyn)
    gE, _ = update_E(tuple(map(gpu_field, E)), gB, tuple(map(gpu_field, J)), static, dyn)
    wB = O.update_B(E, B, p)
    wE = O.update_E(E, wB, J, p)
    for c in range(3):
        assert rel_err(host(gB[c]), wB[c]) <= TOL
        assert rel_err(host(gE[c]), wE[c]) <= TOL


@pytest.mark.parametrize("N", [(10, 9, 8), (12, 1, 1), (9, 7, 1)])
@pytest.mark.parametrize("bc_type", [0, 1])
def test_ghost_refresh_and_fold(N, bc_type):
    static, dyn, p = _setup(N, bc=(0, 1, 0), pbc=(0, 0, 2) if bc_type else (0, 0, 0))
    f = random_field(p, 3, refresh=False)
    bcs = p.bc if bc_type == 0 else p.particle_bc
    got = host(gc.update_tiled_ghost_cells(gpu_field(f), static, 2, bc_type))
    np.testing.assert_array_equal(got, O.refresh(f, p, bcs))
    got = host(gc.fold_tiled_ghost_cells(gpu_field(f), static, 2, bc_type))
    assert rel_err(got, O.fold(f, p, bcs)) <= TOL


def test_filters_and_walls():
    static, dyn, p = _setup((10, 9, 8), bc=(1, 0, 1))
    f = random_field(p, 5)
    assert rel_err(host(digital_filter(gpu_field(f), 0.3, 2)), O.digital_filter(f, 0.3, 2)) <= TOL
    assert rel_err(host(bilinear_filter(gpu_field(f), 2)), O.bilinear_filter(f, 2)) <= TOL
    for axis in range(3):
        assert rel_err(host(gc.apply_tiled_zero_boundary(gpu_field(f), static, axis)),
                       O.zero_boundary(f, p, axis)) <= TOL
        np.testing.assert_array_equal(host(gc.apply_tiled_constant_boundary(gpu_field(f), static, axis)),
                                      O.constant_boundary(f, p, axis))


# ------------------------------------------------------------- particles

@pytest.mark.parametrize("N", DIMS)
@pytest.mark.parametrize("sf", [1, 2])
@pytest.mark.parametrize("rel", [True, False])
def test_push_gather_boris(N, sf, rel):
    static, dyn, p = _setup(N, sf=sf, rel=rel)
    # |v_i| < 0.3 so |v| < C = 1 for the relativistic gamma
    x, u, act = random_particles(p, S=2, P=300, seed=sum(N) + 10 * sf, vmax=0.03, inactive_frac=0.1)
    sp = species(2)
    sp.update_u[1, 1] = False
    E = [random_field(p, 20 + c, scale=0.5) for c in range(3)]
    B = [random_field(p, 30 + c, scale=0.5) for c in range(3)]
    parts = tiled_particles_from_arrays(x, u, act)
    out = particle_push(parts, _cfg(sp), tuple(map(gpu_field, E)), tuple(map(gpu_field, B)), static, dyn)
    want = O.particle_push(x, u, act, sp, E, B, p)
    assert rel_err(host(out.u[0, 0, 0]), want) <= TOL
    assert out.x is parts.x and out.active is parts.active


@pytest.mark.parametrize("N", DIMS)
@pytest.mark.parametrize("sf", [1, 2])
def test_push_gather_higuera_cary(N, sf):
    # pusher_test.py:240-291: particle_push with particle_pusher="higuera_cary"
    static, dyn, p = _setup(N, sf=sf, pusher="higuera_cary")
    # |v_i| < 0.03 d / dt keeps |v| < C = 1 for the relativistic gamma
    x, u, act = random_particles(p, S=2, P=300, seed=sum(N) + 3 * sf, vmax=0.03, inactive_frac=0.1)
    sp = species(2)
    sp.update_u[0, 2] = False
    E = [random_field(p, 40 + c, scale=0.5) for c in range(3)]
    B = [random_field(p, 50 + c, scale=0.5) for c in range(3)]
    parts = tiled_particles_from_arrays(x, u, act)
    out = particle_push(parts, _cfg(sp), tuple(map(gpu_field, E)), tuple(map(gpu_field, B)), static, dyn)
    want = O.particle_push(x, u, act, sp, E, B, p)
    assert rel_err(host(out.u[0, 0, 0]), want) <= TOL


@pytest.mark.parametrize("N", DIMS)
@pytest.mark.parametrize("sf", [1, 2])
def test_esirkepov_current(N, sf):
    static, dyn, p = _setup(N, sf=sf)
    x, u, act = random_particles(p, S=2, P=400, seed=7 + sf, vmax=0.45, inactive_frac=0.1)
    sp = species(2)
    parts = tiled_particles_from_arrays(x, u, act)
    J0 = tuple(gpu_field(np.zeros(p.shape)) for _ in range(3))
    got = Esirkepov_current(parts, _cfg(sp), J0, static, dyn)
    want = O.esirkepov_current(x, u, act, sp, p)
    for c in range(3):
        assert rel_err(host(got[c]), want[c]) <= TOL, c


@pytest.mark.parametrize("N", DIMS)
@pytest.mark.parametrize("sf", [1, 2])
@pytest.mark.parametrize("filt", ["none", "bilinear", "digital"])
def test_direct_current_and_rho(N, sf, filt):
    static, dyn, p = _setup(N, sf=sf, dep="direct", filt=filt, alpha=0.6)
    x, u, act = random_particles(p, S=2, P=400, seed=9 + sf, vmax=0.45, inactive_frac=0.1)
    sp = species(2)
    parts = tiled_particles_from_arrays(x, u, act)
    J0 = tuple(gpu_field(np.zeros(p.shape)) for _ in range(3))
    got = J_from_rhov(parts, _cfg(sp), J0, static, dyn)
    want = O.j_from_rhov(x, u, act, sp, p)
    for c in range(3):
        assert rel_err(host(got[c]), want[c]) <= TOL
    rho = compute_rho(parts, _cfg(sp), gpu_field(np.zeros(p.shape)), static, dyn)
    assert rel_err(host(rho), O.compute_rho(x, act, sp, p)) <= TOL


@pytest.mark.parametrize("pbc", [(0, 0, 0), (1, 1, 1), (2, 2, 2), (0, 1, 2)])
def test_positions_and_particle_boundaries(pbc):
    static, dyn, p = _setup((8, 6, 4), pbc=pbc)
    x, u, act = random_particles(p, S=2, P=500, seed=3, vmax=4.0, inactive_frac=0.1)   # many leave the box
    sp = species(2)
    sp.update_x[0, 2] = False
    parts = tiled_particles_from_arrays(x, u, act)
    moved = update_tiled_particle_positions(parts, _cfg(sp), dyn.dt)
    xw = O.update_positions(x, u, act, sp, p.dt)
    xg, _, _ = host_particles(moved)
    assert np.max(np.abs(xg - xw)) <= 1e-6 * max(p.wind)
    ref, ovf = refresh_tiled_particle_tiles(moved, static, dyn)
    assert not bool(ovf)
    xo, uo, ao, _ = O.refresh_particles(xg, u, act, p)   # boundaries act on the stored FP32 positions
    xg2, ug2, ag2 = host_particles(ref)
    np.testing.assert_array_equal(ag2, ao)
    assert np.max(np.abs(xg2 - xo)) <= 1e-6 * max(p.wind)
    np.testing.assert_array_equal(ug2, uo)


def test_cell_keys_and_stable_sort_are_bit_exact():
    static, dyn, p = _setup((16, 12, 10), wind=(16.0, 12.0, 10.0))
    x, u, act = random_particles(p, S=2, P=5000, seed=5, inactive_frac=0.05)
    # put particles exactly on faces and on the +W/2 edge
    x[0, :40, 0] = np.round(x[0, :40, 0])
    x[1, :10, 1] = p.wind[1] / 2
    sp = species(2)
    parts = tiled_particles_from_arrays(x, u, act)
    keys = cell_keys(parts, _cfg(sp), static, dyn).cpu().numpy()
    ncell = 16 * 12 * 10
    want = O.cell_keys(x.astype(np.float32).astype(np.float64), p)
    want = np.where(act, want, ncell) + np.arange(2)[:, None] * (ncell + 1)
    np.testing.assert_array_equal(keys, want)
    sorted_parts, perm = sort_particles_by_cell(parts, _cfg(sp), static, dyn)
    np.testing.assert_array_equal(perm.cpu().numpy(), O.sort_permutation(want.reshape(-1)))
    xs, us, as_ = host_particles(sorted_parts)
    flat = lambda a: a.reshape(-1, *a.shape[2:])   # noqa: E731
    np.testing.assert_array_equal(flat(xs), flat(x)[O.sort_permutation(want.reshape(-1))])


@pytest.mark.parametrize("tile", [4, 8, 3])
def test_tile_keys_and_ranges_are_exact(tile):
    """pic_tile_keys: key = s*(span+1) + tile*T^3 + cell-in-tile with the
    cell clip(floor((x + W/2) * (1/d)), 0, N-1) in FP64 (the same operations
    in NumPy give the same bits), inactive slots at s*(span+1)+span; the tile
    ranges are the left insertion points of each (species, tile) key."""
    from pypic3d_amd import _native as N
    from pypic3d_amd._tensors import make_desc, make_species, particle_planes, pic_particles
    static, dyn, p = _setup((18, 12, 9), wind=(9.0, 6.0, 4.5))
    S, P = 2, 6000
    x, u, act = random_particles(p, S=S, P=P, seed=11, inactive_frac=0.05)
    x[0, :50, 0] = np.round(x[0, :50, 0] * 2.0) / 2.0            # on cell faces
    x[1, :20, 1] = p.wind[1] / 2                                  # on the +W/2 edge (clipped)
    x[1, 20:40, 2] = -p.wind[2] / 2
    parts = tiled_particles_from_arrays(x, u, act)
    xp, up, ap = particle_planes(parts)
    desc, sp = make_desc(static, dyn), make_species(_cfg(species(S)), P)
    keys = torch.empty(S * P, dtype=torch.int32, device="cuda")
    N.call("pic_tile_keys", desc, sp, pic_particles(xp, up, ap), tile, N.ptr(keys), N.stream_handle())
    n = np.array(p.N)
    nt = (n + tile - 1) // tile
    vol = tile ** 3
    span = int(nt.prod()) * vol
    c = np.empty((S, P, 3), dtype=np.int64)
    for a in range(3):
        t = np.floor((x[..., a].astype(np.float32).astype(np.float64) + 0.5 * p.wind[a]) * (1.0 / p.d[a]))
        c[..., a] = np.clip(t, 0, n[a] - 1)
    tl = ((c[..., 0] // tile) * nt[1] + c[..., 1] // tile) * nt[2] + c[..., 2] // tile
    loc = ((c[..., 0] % tile) * tile + c[..., 1] % tile) * tile + c[..., 2] % tile
    want = np.where(act, tl * vol + loc, span) + np.arange(S)[:, None] * (span + 1)
    np.testing.assert_array_equal(keys.cpu().numpy().reshape(S, P), want)
    # ranges on the sorted keys
    tiles = int(N.lib().pic_tile_count(desc, tile))
    assert tiles == int(nt.prod())
    ks = torch.as_tensor(np.sort(want.reshape(-1)).astype(np.int32)).cuda()
    starts = torch.empty(S * (tiles + 1) + 1, dtype=torch.int32, device="cuda")
    N.call("pic_tile_ranges", desc, sp, tile, N.ptr(ks), N.ptr(starts), N.stream_handle())
    b = np.arange(S * (tiles + 1))
    targets = (b // (tiles + 1)) * (span + 1) + (b % (tiles + 1)) * vol
    want_starts = np.append(np.searchsorted(np.sort(want.reshape(-1)), targets, side="left"), S * P)
    np.testing.assert_array_equal(starts.cpu().numpy(), want_starts)


def test_poisson_cg_single_mode():
    static, dyn, p = _setup((24, 20, 16), wind=(2 * np.pi,) * 3, electrostatic=True)
    xs = [O.tiled_axis(p, a, "center")[2:-2] for a in range(3)]
    X, Y, Z = np.meshgrid(*xs, indexing="ij")
    phi_true = np.sin(X + Y + Z)
    rhs = -O.laplacian_interior(O.refresh(np.pad(phi_true, 2), p), p)
    rho = np.zeros(p.shape)
    rho[2:-2, 2:-2, 2:-2] = rhs
    rho = rho.astype(np.float32).astype(np.float64)
    got = host(solve_poisson_with_conjugate_gradient(gpu_field(rho), gpu_field(np.zeros(p.shape)), static, dyn,
                                                     tol=1e-10, max_iter=4000))
    want = O.solve_poisson_cg(rho, np.zeros(p.shape), p, tol=1e-11, max_iter=4000)
    gi, wi = got[2:-2, 2:-2, 2:-2], want[2:-2, 2:-2, 2:-2]
    assert rel_err(gi - gi.mean(), wi - wi.mean()) <= TOL


def test_energy_is_fp64_accurate():
    static, dyn, p = _setup((12, 10, 8), wind=(6.0, 5.0, 4.0))
    x, u, act = random_particles(p, S=2, P=1000, seed=2, vmax=0.03, inactive_frac=0.1)
    sp = species(2)
    E = [random_field(p, 40 + c) for c in range(3)]
    B = [random_field(p, 50 + c) for c in range(3)]
    parts = tiled_particles_from_arrays(x, u, act)
    got = compute_energy(parts, tuple(map(gpu_field, E)), tuple(map(gpu_field, B)), static, dyn, _cfg(sp))
    want = O.compute_energy(u, act, sp, E, B, p)
    for g_, w_ in zip(got, want):
        assert abs(float(g_) - w_) <= 1e-10 * abs(w_)
