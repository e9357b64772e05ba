"""Tile-ordered fused kernel on exact grids (FP32 anchors by bracketing exact
node coordinates, register-summed deposits, rare-list second kernel) against
the strided tile kernel on the same launch, and whole runs against the FP64
oracle, including particles that leave their tile and steps longer than a
cell (both handled by the rare-list kernel)."""

import numpy as np
import pytest
import torch

from pypic3d_amd import _native as N
from pypic3d_amd import workloads
from pypic3d_amd.evolve import time_loop_electrodynamic
from pypic3d_amd.runner import ElectrodynamicRunner
from tests.gpu_helpers import gpu_field, host, host_particles, oracle_params, rel_err
from tests.helpers import random_field

pytestmark = pytest.mark.gpu


def _runner(N_, wind, specs, tile, rounds, min_blocks, dt=None, ext=True, seed=3):
    static, dyn = workloads.make_parameters(N_, wind=wind, cfl=0.5, dt=dt)
    parts, cfg = workloads.plasma(static, dyn, specs, seed=seed)
    p = oracle_params(static, dyn)
    fields = workloads.empty_fields(static, dyn)
    E0 = tuple(gpu_field(random_field(p, 40 + c, scale=0.05)) for c in range(3))
    B0 = tuple(gpu_field(random_field(p, 43 + c, scale=0.05)) for c in range(3))
    if ext:
        Ee = tuple(gpu_field(random_field(p, 46 + c, scale=0.02)) for c in range(3))
        Be = tuple(gpu_field(random_field(p, 49 + c, scale=0.02)) for c in range(3))
        fields = (E0, B0) + fields[2:5] + ((Ee, Be),) + fields[6:]
    else:
        fields = (E0, B0) + fields[2:]
    r = ElectrodynamicRunner(parts, cfg, fields, static, dyn, sort_every=0, tile=tile, rounds=rounds,
                             min_blocks=min_blocks)
    return r, parts, cfg, fields, static, dyn


def _launch(r, variant):
    src, dst = r.cur, 1 - r.cur
    r.desc.variant = variant
    for c in r.J:
        c.zero_()
    r.tile_ws[0] = -7
    N.call("pic_fused_push_esirkepov_tiled", r.desc, r.sp, r._pp(src), N.p3(r.E[r.ecur]), N.p3(r.B), r.Ee, r.Be,
           r._pp(dst), N.p3(r.J), r.tile, N.ptr(r.starts), r.rounds, r.min_blocks, N.ptr(r.tile_ws),
           N.stream_handle())
    torch.cuda.synchronize()
    r.desc.variant = 0
    return (torch.stack(r.J).clone(), [t.clone() for t in r.pos[dst]], [t.clone() for t in r.vel[dst]],
            r.act[dst].clone(), int(r.tile_ws[0]))


@pytest.mark.parametrize("tile,rounds,min_blocks", [(4, 4, 4), (4, 2, 1), (8, 4, 3), (4, 8, 4), (8, 8, 1)])
def test_exact_tile_kernel_agrees_with_strided_tile_kernel(tile, rounds, min_blocks):
    """Same sorted input, one launch each: identical particles, J equal up to
    FP32 summation order; some particles have drifted out of their tile or
    wrapped through the periodic faces and take the rare list."""
    specs = [workloads.SpeciesSpec(24, weight=0.01, vth=(0.2,) * 3),
             workloads.SpeciesSpec(7, charge=1.0, mass=20.0, weight=0.03, vth=(0.05,) * 3)]
    r, *_ = _runner((20, 16, 12), (10.0, 8.0, 6.0), specs, tile, rounds, min_blocks)
    for _ in range(6):          # drift away from the sorted order (no re-sort)
        r.step()
    J1, x1, u1, a1, nr1 = _launch(r, 1)
    J0, x0, u0, a0, nr0 = _launch(r, 0)
    assert nr1 == -7                                  # strided kernel leaves the workspace alone
    assert 0 < nr0 < a0.numel() // 10
    for p, q in zip(x0 + u0, x1 + u1):
        assert torch.equal(p, q)
    assert torch.equal(a0, a1)
    assert float((J0 - J1).norm() / J1.norm()) <= 1e-6


@pytest.mark.parametrize("tile,rounds,min_blocks,dt", [(4, 4, 4, None), (8, 2, 1, None), (4, 8, 4, 3.0)])
def test_gathered_read_equals_permute_then_push(tile, rounds, min_blocks, dt):
    """A re-bin that leaves its gather to the push (pic_tile_rebin without
    out, then pic_fused_push_esirkepov_tiled_gather reading slot i from
    perm[i]) against the permuted particles pushed by
    pic_fused_push_esirkepov_tiled: identical particles and rare lists, J
    equal up to FP32 summation order.  dt = 3 sends most particles through
    the rare-list kernel, which reads through perm as well."""
    specs = [workloads.SpeciesSpec(24, weight=0.01, vth=(0.2,) * 3),
             workloads.SpeciesSpec(7, charge=1.0, mass=20.0, weight=0.03, vth=(0.05,) * 3)]
    r, *_ = _runner((20, 16, 12), (20.0, 16.0, 12.0), specs, tile, rounds, min_blocks, dt=dt)
    assert r.defer_permute
    # drift, then a re-bin whose gather is deferred (one step at dt = 3, which
    # is past the field solver's CFL limit)
    for _ in range(4 if dt is None else 1):
        r.step()
    r.sort()
    assert r._pending
    src, dst = r._cur, 1 - r._cur
    for c in r.J:
        c.zero_()
    r.tile_ws[0] = -7
    N.call("pic_fused_push_esirkepov_tiled_gather", r.desc, r.sp, r._pp(src), N.ptr(r.perm), N.p3(r.E[r.ecur]),
           N.p3(r.B), r.Ee, r.Be, r._pp(dst), N.p3(r.J), r.tile, N.ptr(r.starts), r.rounds, r.min_blocks,
           N.ptr(r.tile_ws), N.stream_handle())
    torch.cuda.synchronize()
    Ja, xa, ua = torch.stack(r.J).clone(), r.pos[dst].clone(), r.vel[dst].clone()
    aa, nra = r.act[dst].clone(), int(r.tile_ws[0])
    Jb, xb, ub, ab, nrb = _launch(r, 0)             # r.cur applies perm (into dst), then the plain launch
    assert not r._pending
    assert nra == nrb and nra >= 0
    if dt is not None:
        assert nra > aa.numel() // 4
    bits = lambda t: t.view(torch.int32)            # noqa: E731
    assert torch.equal(bits(xa), bits(torch.stack(xb))) and torch.equal(bits(ua), bits(torch.stack(ub)))
    assert torch.equal(aa, ab)
    assert bool(torch.isfinite(Jb).all())
    assert float((Ja - Jb).norm() / Jb.norm()) <= 1e-6


def test_pending_permutation_is_applied_before_the_strided_kernel():
    """A re-bin leaves its permutation pending; switching to the strided
    comparison kernel (variant 1, which has no gathered read) before the next
    step applies the permutation first, giving the particles of a runner that
    read its state in between."""
    specs = [workloads.SpeciesSpec(12, weight=0.01, vth=(0.15,) * 3)]
    runs = []
    for read_between in (False, True):
        r, *_ = _runner((16, 12, 10), (16.0, 12.0, 10.0), specs, 4, 4, 1)
        r.step()
        r.sort()
        assert r._pending
        if read_between:
            r.state()
            assert not r._pending
        r.desc.variant = 1
        r.step()
        r.desc.variant = 0
        assert not r._pending
        parts, _ = r.state()
        x = parts.x.reshape(-1, 3)
        u = parts.u.reshape(-1, 3)
        act = parts.active.reshape(-1)
        x, u = x[act], u[act]
        order = torch.argsort(x[:, 0])      # the two re-bins may order equal keys differently
        runs.append((x[order].clone(), u[order].clone(), int(act.sum())))
    (xa, ua, na), (xb, ub, nb) = runs
    # J of the first step sums in atomic order, so E and the second push agree to FP32 rounding
    assert na == nb
    assert float((xa - xb).abs().max()) <= 1e-5
    assert float((ua - ub).abs().max()) <= 1e-5 * float(ua.abs().max())


def test_gathered_read_needs_an_exact_grid():
    specs = [workloads.SpeciesSpec(8, weight=0.01, vth=(0.1,) * 3)]
    r, *_ = _runner((12, 10, 8), (1.3, 1.1, 0.7), specs, 4, 4, 1, ext=False)
    assert not r.defer_permute and not r._pending
    src, dst = r.cur, 1 - r.cur
    perm = torch.arange(r.S * r.P, dtype=torch.int32, device="cuda")
    rc = N.lib().pic_fused_push_esirkepov_tiled_gather(
        r.desc, r.sp, r._pp(src), N.ptr(perm), N.p3(r.E[r.ecur]), N.p3(r.B), r.Ee, r.Be, r._pp(dst), N.p3(r.J),
        r.tile, N.ptr(r.starts), r.rounds, r.min_blocks, N.ptr(r.tile_ws), N.stream_handle())
    assert rc == -1


def test_new_positions_next_to_a_face_take_the_crossing_route():
    """The fast loop tests "stays in its cell" on the FP32 offset
    (x - node) + v dt against a band of 2^-19 spacings at each face, twice the
    FP64 test's 2^-20: a particle whose FP64 new position lies between the
    two bands stays in its cell but goes through the crossing queue.  Frozen
    velocities (update_u = False) put new positions 1.5 * 2^-20 inside a face
    on purpose; the kernel's particles equal the strided kernel's bit for bit
    and J agrees to FP32 summation order."""
    specs = [workloads.SpeciesSpec(16, weight=0.01, vth=(0.1,) * 3, update_u=False)]
    r, *_ = _runner((20, 16, 12), (20.0, 16.0, 12.0), specs, 4, 4, 4, ext=False)
    k = r.cur
    dt, tol = float(r.desc.dt), 2.0 ** -20
    act = r.act[k].view(-1)
    idx = torch.nonzero(act).view(-1)[::37][:200]
    for a in range(3):
        org, sp = float(r.desc.org_c[a]), float(r.desc.sp_c[a])
        assert sp == 1.0
        x = r.pos[k][a].view(-1)
        node = torch.floor((x[idx].double() - org) / sp) * sp + org
        low = (torch.arange(idx.numel(), device="cuda") + a) % 2 == 0
        x0 = node + torch.where(low, 0.25, 0.75)                  # exact in FP32, same cell
        target = torch.where(low, node + 1.5 * tol, node + sp - 1.5 * tol)
        x[idx] = x0.float()
        r.vel[k][a].view(-1)[idx] = ((target - x0) / dt).float()
    J1, x1, u1, a1, _ = _launch(r, 1)
    J0, x0_, u0, a0, nr0 = _launch(r, 0)
    # the planted particles stayed in their cells, inside the band between the tests
    for a in range(3):
        org = float(r.desc.org_c[a])
        xn = x0_[a].view(-1)[idx].double()
        off = xn - (torch.floor(xn - org) + org)
        assert bool(((off > 0.5 * tol) & (off < 3 * tol) | (off > 1 - 3 * tol) & (off < 1 - 0.5 * tol)).all())
    for p, q in zip(x0_ + u0, x1 + u1):
        assert torch.equal(p, q)
    assert torch.equal(a0, a1)
    assert 0 <= nr0 < a0.numel() // 10
    assert float((J0 - J1).norm() / J1.norm()) <= 1e-6


def test_non_exact_grid_uses_the_strided_kernel():
    specs = [workloads.SpeciesSpec(8, weight=0.01, vth=(0.1,) * 3)]
    r, *_ = _runner((12, 10, 8), (1.3, 1.1, 0.7), specs, 4, 4, 1, ext=False)
    r.step()
    J0, x0, u0, a0, nr0 = _launch(r, 0)
    J1, x1, u1, a1, nr1 = _launch(r, 1)
    assert nr0 == -7 and nr1 == -7
    for p, q in zip(x0 + u0, x1 + u1):
        assert torch.equal(p, q)
    assert float((J0 - J1).norm() / J1.norm()) <= 1e-6


@pytest.mark.parametrize("dt,steps", [(None, 6), (3.0, 2)])
def test_exact_tile_runner_matches_functional_step(dt, steps):
    """Tile runner without re-sorting against time_loop_electrodynamic (the
    slot-ordered functional step, itself checked against the FP64 oracle);
    dt = 3 on a unit grid moves the fast particles more than a cell per step,
    through the rare-list kernel's general route."""
    specs = [workloads.SpeciesSpec(12, weight=0.01, vth=(0.15,) * 3),
             workloads.SpeciesSpec(5, charge=1.0, mass=20.0, weight=0.03, vth=(0.05,) * 3)]
    r, parts, cfg, fields, static, dyn = _runner((16, 12, 10), (16.0, 12.0, 10.0), specs, 4, 4, 1, dt=dt)
    ref_parts, ref_fields = parts, fields
    for _ in range(steps):
        r.step()
        ref_parts, ref_fields = time_loop_electrodynamic(ref_parts, cfg, ref_fields, static, dyn)
    if dt is not None:
        assert int(r.tile_ws[0]) > 0
    r_parts, r_fields = r.state()
    xa, ua, aa = host_particles(r_parts)
    xb, ub, ab = host_particles(ref_parts)
    # the runner sorted once (set_tile): its slot i holds reference slot perm[i]
    perm = r.perm.cpu().numpy().astype(np.int64)
    flat = lambda a: a.reshape(-1, *a.shape[2:])    # noqa: E731
    xa, ua, aa = flat(xa), flat(ua), flat(aa)
    xb, ub, ab = flat(xb)[perm], flat(ub)[perm], flat(ab)[perm]
    np.testing.assert_array_equal(aa, ab)
    assert np.max(np.abs(xa - xb)) <= 1e-5
    assert rel_err(ua, ub) <= 1e-5
    for k in (0, 1, 2):
        for c in range(3):
            assert rel_err(host(r_fields[k][c]), host(ref_fields[k][c])) <= 1e-5, (k, c)
