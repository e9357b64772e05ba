"""Fused direct-branch tile kernel (gather + push, half advances, particle
boundaries and the J_from_rhov deposit in one launch) against the separate
kernels of the same step: pic_push, pic_advance_positions,
pic_particle_boundaries, pic_deposit_direct, pic_advance_positions,
pic_particle_boundaries (each checked against the FP64 oracle in
test_gpu_kernels / test_gpu_step).  The two launches compile the same push
expressions in different kernels, so the compiler's multiply-add fusion may
round a velocity differently by an ulp of its largest term: velocities are
compared to 2e-6 of the largest speed, positions to 1e-6 of the box,
activity flags exactly."""

import pytest
import torch

from pypic3d_amd import _native as N
from pypic3d_amd import workloads
from pypic3d_amd._tensors import pic_particles
from pypic3d_amd.runner import ElectrodynamicRunner
from tests.gpu_helpers import gpu_field, oracle_params
from tests.helpers import random_field

pytestmark = pytest.mark.gpu


def _runner(N_, wind, pbc, tile, dt=None, seed=7, d=None, specs=None):
    static, dyn = workloads.make_parameters(N_, wind=wind, d=d, cfl=0.5, dt=dt, pbc=pbc,
                                            current_deposition="direct")
    specs = specs or [workloads.SpeciesSpec(20, weight=0.01, vth=(0.12,) * 3),
                      workloads.SpeciesSpec(6, charge=1.0, mass=20.0, weight=0.03, vth=(0.05,) * 3)]
    parts, cfg = workloads.plasma(static, dyn, specs, seed=seed)
    p = oracle_params(static, dyn)
    fields = workloads.empty_fields(static, dyn)
    E0 = tuple(gpu_field(random_field(p, 70 + c, scale=0.05)) for c in range(3))
    B0 = tuple(gpu_field(random_field(p, 73 + c, scale=0.05)) for c in range(3))
    Ee = tuple(gpu_field(random_field(p, 76 + c, scale=0.02)) for c in range(3))
    Be = tuple(gpu_field(random_field(p, 79 + c, scale=0.02)) for c in range(3))
    fields = (E0, B0) + fields[2:5] + ((Ee, Be),) + fields[6:]
    return ElectrodynamicRunner(parts, cfg, fields, static, dyn, sort_every=0, tile=tile)


def _separate(r):
    """The unfused sequence into the inactive buffer set; returns its result."""
    d, s = r.desc, N.stream_handle()
    src, dst = r.cur, 1 - r.cur
    for c in r.J:
        c.zero_()
    uo = [r.vel[dst][a] for a in range(3)]
    N.call("pic_push", d, r.sp, r._pp(src), N.p3(r.E[r.ecur]), N.p3(r.B), r.Ee, r.Be, N.p3(uo), s)
    r.act[dst].copy_(r.act[src])
    half = pic_particles([r.pos[src][a] for a in range(3)], uo, r.act[dst])
    xd = [r.pos[dst][a] for a in range(3)]
    N.call("pic_advance_positions", d, r.sp, half, 0.5, N.p3(xd), s)
    N.call("pic_particle_boundaries", d, r.sp, r._pp(dst), s)
    N.call("pic_deposit_direct", d, r.sp, r._pp(dst), N.p3(r.J), s)
    N.call("pic_advance_positions", d, r.sp, r._pp(dst), 0.5, N.p3(xd), s)
    N.call("pic_particle_boundaries", d, r.sp, r._pp(dst), s)
    torch.cuda.synchronize()
    return (torch.stack(r.J).clone(), r.pos[dst].clone(), r.vel[dst].clone(),
            r.act[dst].clone())


def _fused(r, min_blocks=1):
    src, dst = r.cur, 1 - r.cur
    for c in r.J:
        c.zero_()
    for t in (r.pos[dst], r.vel[dst]):
        t.fill_(-3.0)
    r.act[dst].fill_(True)
    r.tile_ws[0] = -7
    N.call("pic_fused_push_direct_tiled", r.desc, r.sp, r._pp(src), N.p3(r.E[r.ecur]), N.p3(r.B), r.Ee, r.Be,
           r._pp(dst), N.p3(r.J), r.tile, N.ptr(r.starts), min_blocks, N.ptr(r.tile_ws), N.stream_handle())
    torch.cuda.synchronize()
    return (torch.stack(r.J).clone(), r.pos[dst].clone(), r.vel[dst].clone(),
            r.act[dst].clone(), int(r.tile_ws[0]))


@pytest.mark.parametrize("tile,pbc,dt,steps", [(4, (0, 0, 0), None, 4), (8, (0, 0, 0), None, 4),
                                               (4, (1, 2, 0), None, 4), (4, (0, 0, 0), 3.0, 1)])
def test_fused_direct_launch_matches_separate_kernels(tile, pbc, dt, steps):
    """Same particles, J equal up to FP32 summation order; drifted particles,
    reflecting and absorbing walls and steps longer than a cell (dt = 3 on a
    unit grid, one step before the comparison since the Yee update is
    unstable there) take the rare-list route."""
    r = _runner((20, 16, 12), (20.0, 16.0, 12.0), pbc, tile, dt=dt)
    assert r.direct_fused
    for _ in range(steps):      # drift away from the sorted order (no re-sort)
        r.step()
    J1, x1, u1, a1, nr = _fused(r)
    J4, x4, u4, a4, nr4 = _fused(r, 4)          # 4 blocks per CU (tile 4: two slots per thread): same arithmetic
    assert torch.equal(x1, x4) and torch.equal(u1, u4) and torch.equal(a1, a4) and nr4 == nr
    assert float((J4 - J1).norm() / J1.norm()) <= 1e-6
    J0, x0, u0, a0 = _separate(r)
    n = a0.numel()
    assert 0 < nr < n
    if dt is None and pbc == (0, 0, 0):
        assert nr < n // 10
    assert torch.equal(a0, a1)
    assert bool(torch.isfinite(x0).all() and torch.isfinite(u0).all())
    assert float((x0 - x1).abs().max()) <= 1e-6 * 20.0
    assert float((u0 - u1).abs().max()) <= 2e-6 * float(u0.abs().max())
    if pbc[1] == 2:
        assert int(a0.sum()) < n     # the absorbing wall removed particles
    assert float((J0 - J1).norm() / J0.norm()) <= 1e-6


@pytest.mark.parametrize("tile,min_blocks,pbc,dt", [(4, 4, (0, 0, 0), None), (4, 1, (1, 2, 0), None),
                                                    (8, 1, (0, 0, 0), None), (8, 4, (0, 0, 0), None),
                                                    (4, 4, (0, 0, 0), 3.0)])
def test_gathered_direct_launch_equals_permute_then_launch(tile, min_blocks, pbc, dt):
    """A re-bin that leaves its gather to the fused direct launch
    (pic_fused_push_direct_tiled_gather reads slot i from perm[i]) against
    the permuted particles through pic_fused_push_direct_tiled: identical
    particles and rare lists, J equal up to FP32 summation order; the four
    launch shapes (16-byte group loads, 8-byte groups, slot-by-slot loads)
    and the rare-list route (dt = 3: 5 % of the particles) all read through
    perm."""
    r = _runner((20, 16, 12), (20.0, 16.0, 12.0), pbc, tile, dt=dt)
    assert r.direct_fused and r.defer_permute
    for _ in range(3 if dt is None else 1):
        r.step()
    r.sort()
    assert r._pending
    src, dst = r._cur, 1 - r._cur
    for c in r.J:
        c.zero_()
    r.tile_ws[0] = -7
    N.call("pic_fused_push_direct_tiled_gather", r.desc, r.sp, r._pp(src), N.ptr(r.perm), N.p3(r.E[r.ecur]),
           N.p3(r.B), r.Ee, r.Be, r._pp(dst), N.p3(r.J), r.tile, N.ptr(r.starts), min_blocks, N.ptr(r.tile_ws),
           N.stream_handle())
    torch.cuda.synchronize()
    Ja, xa, ua, aa, nra = (torch.stack(r.J).clone(), r.pos[dst].clone(), r.vel[dst].clone(), r.act[dst].clone(),
                           int(r.tile_ws[0]))
    Jb, xb, ub, ab, nrb = _fused(r, min_blocks)      # r.cur applies perm (into dst), then the plain launch
    assert not r._pending
    assert nra == nrb and nra > 0
    if dt is not None:
        assert nra > aa.numel() // 50
    bits = lambda t: t.view(torch.int32)            # noqa: E731
    assert torch.equal(bits(xa), bits(xb)) and torch.equal(bits(ua), bits(ub)) and torch.equal(aa, ab)
    assert bool(torch.isfinite(Jb).all())
    assert float((Ja - Jb).norm() / Jb.norm()) <= 1e-6


def test_fused_direct_with_frozen_ions_and_drifting_beam():
    """Three species: a drifting beam, a thermal species and frozen ions
    (update_x = update_u = False, no push and no motion, but a deposit)."""
    specs = [workloads.SpeciesSpec(8, weight=0.01, vth=(0.02,) * 3, drift=(0.3, -0.1, 0.05)),
             workloads.SpeciesSpec(8, weight=0.01, vth=(0.1,) * 3),
             workloads.SpeciesSpec(4, charge=1.0, mass=1836.0, weight=0.04, vth=(0.01,) * 3,
                                   update_x=False, update_u=False)]
    r = _runner((16, 12, 8), (16.0, 12.0, 8.0), (0, 0, 0), 4, specs=specs)
    assert r.direct_fused
    for _ in range(3):
        r.step()
    J1, x1, u1, a1, nr = _fused(r)
    J0, x0, u0, a0 = _separate(r)
    assert torch.equal(a0, a1)
    assert float((x0 - x1).abs().max()) <= 1e-6 * 16.0
    assert float((u0 - u1).abs().max()) <= 2e-6 * float(u0.abs().max())
    assert torch.equal(x1[:, 2], r.pos[r.cur][:, 2]) and torch.equal(u1[:, 2], r.vel[r.cur][:, 2])
    assert float((J0 - J1).norm() / J0.norm()) <= 1e-6


def test_non_exact_or_unequal_cells_keep_the_separate_kernels():
    r = _runner((12, 10, 8), (1.3, 1.1, 0.7), (0, 0, 0), 4)
    assert not r.direct_fused
    r.step()
