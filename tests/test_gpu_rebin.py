"""pic_tile_rebin (counting re-bin into tile-key order) against the tile keys
computed in NumPy: the output is a permutation of the input (bit-equal
planes at the reported source slots), its keys are non-decreasing, every
key's slots are contiguous, and the (species, tile) ranges equal the left
insertion points of the sorted keys -- for unsorted input and for the nearly
sorted input of a running simulation."""

import numpy as np
import pytest
import torch

from pypic3d_amd import _native as N
from pypic3d_amd import workloads
from pypic3d_amd._tensors import (make_desc, make_species, new_soa, particle_planes, pic_particles,
                                  tiled_particles_from_arrays)
from pypic3d_amd.particles.particle_class import SpeciesConfig
from tests.gpu_helpers import oracle_params
from tests.helpers import random_particles, species

pytestmark = pytest.mark.gpu


def _want_keys(x, act, p, tile):
    S, P = act.shape
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
    return np.where(act, tl * vol + loc, span) + np.arange(S)[:, None] * (span + 1), int(nt.prod()), span, vol


def _rebin(desc, sp, parts, tile, ahead=0.0, gather=True):
    """gather=False: the re-bin stops at starts and perm (out NULL); the
    returned planes are then the input planes taken through perm here."""
    xp, up, ap = particle_planes(parts)
    S, P = xp[0].shape
    xs, xv = new_soa(S, P, "cuda")
    us, uv = new_soa(S, P, "cuda")
    ao = torch.empty_like(ap)
    tiles = int(N.lib().pic_tile_count(desc, tile))
    starts = torch.full((S * (tiles + 1) + 1,), -1, dtype=torch.int32, device="cuda")
    perm = torch.full((S * P,), -1, dtype=torch.int32, device="cuda")
    ws = torch.empty(int(N.lib().pic_tile_rebin_workspace_bytes(desc, sp, tile)), dtype=torch.uint8, device="cuda")
    for t in (xs, us):
        t.fill_(-1.0)
    ao.fill_(True)
    out = pic_particles([xs[0], xs[1], xs[2]], [us[0], us[1], us[2]], ao) if gather else None
    N.call("pic_tile_rebin", desc, sp, pic_particles(xp, up, ap), tile, ahead, out, N.ptr(starts), N.ptr(perm),
           N.ptr(ws), ws.numel(), N.stream_handle())
    torch.cuda.synchronize()
    if not gather:
        assert bool((xs == -1.0).all()) and bool((us == -1.0).all()) and bool(ao.all())    # left alone
        p64 = perm.long()
        for a in range(3):
            xs[a].view(-1).copy_(xp[a].reshape(-1)[p64])
            us[a].view(-1).copy_(up[a].reshape(-1)[p64])
        ao.view(-1).copy_(ap.reshape(-1)[p64])
    return xs, us, ao, starts.cpu().numpy(), perm.cpu().numpy().astype(np.int64)


def _check(x, u, act, xs, us, ao, starts, perm, p, tile):
    S, P = act.shape
    want, tiles, span, vol = _want_keys(x, act, p, tile)
    flat = want.reshape(-1)
    assert np.array_equal(np.sort(perm), np.arange(S * P))               # a permutation
    xin = np.moveaxis(x, -1, 0).reshape(3, -1).astype(np.float32)
    uin = np.moveaxis(u, -1, 0).reshape(3, -1).astype(np.float32)
    np.testing.assert_array_equal(xs.reshape(3, -1).cpu().numpy(), xin[:, perm])
    np.testing.assert_array_equal(us.reshape(3, -1).cpu().numpy(), uin[:, perm])
    np.testing.assert_array_equal(ao.reshape(-1).cpu().numpy(), act.reshape(-1)[perm])
    k = flat[perm]
    assert np.all(k[1:] >= k[:-1])                                         # grouped by ascending key
    b = np.arange(S * (tiles + 1))
    targets = (b // (tiles + 1)) * (span + 1) + (b % (tiles + 1)) * vol
    want_starts = np.append(np.searchsorted(np.sort(flat), targets, side="left"), S * P)
    np.testing.assert_array_equal(starts, want_starts)


@pytest.mark.parametrize("tile,gather", [(4, True), (8, True), (4, False)])
def test_rebin_of_unsorted_particles(tile, gather):
    static, dyn = workloads.make_parameters((18, 12, 9), wind=(9.0, 6.0, 4.5), dt=0.05)
    p = oracle_params(static, dyn)
    S, P = 2, 7000
    x, u, act = random_particles(p, S=S, P=P, seed=5, inactive_frac=0.05)
    x[0, :50, 0] = np.round(x[0, :50, 0] * 2.0) / 2.0            # on cell faces
    x[1, :20, 1] = p.wind[1] / 2                                  # on the +W/2 edge (clipped)
    sp = species(S)
    cfg = SpeciesConfig(sp.charge, sp.mass, sp.weight, sp.update_x, sp.update_u)
    parts = tiled_particles_from_arrays(x, u, act)
    desc, spk = make_desc(static, dyn), make_species(cfg, P)
    _check(x, u, act, *_rebin(desc, spk, parts, tile, gather=gather), p, tile)


def test_rebin_in_a_running_simulation():
    """Dense cell-ordered plasma stepped 6 times: nearly sorted input with long
    equal-key runs and particles that changed cell or tile."""
    from pypic3d_amd.runner import ElectrodynamicRunner
    from tests.gpu_helpers import host_particles
    static, dyn = workloads.make_parameters((16, 12, 8), cfl=0.5)
    p = oracle_params(static, dyn)
    specs = [workloads.SpeciesSpec(40, weight=0.01, vth=(0.2,) * 3),
             workloads.SpeciesSpec(9, charge=1.0, mass=20.0, weight=0.03, vth=(0.05,) * 3)]
    parts, cfg = workloads.plasma(static, dyn, specs, seed=8)
    r = ElectrodynamicRunner(parts, cfg, workloads.empty_fields(static, dyn), static, dyn, tile=4)
    for _ in range(6):
        r.step()
    cur, _ = r.state()
    x, u, act = host_particles(cur)
    _check(x, u, act, *_rebin(r.desc, r.sp, cur, 4), p, 4)
    # the runner's own re-bin leaves the state the kernels expect
    r.sort()
    x2, _, a2 = host_particles(r.state()[0])
    k2, *_ = _want_keys(x2, a2, p, 4)
    k2 = k2.reshape(-1)
    assert np.all(k2[1:] >= k2[:-1])


def test_rebin_with_lookahead_sorts_by_the_predicted_position():
    """lookahead t: keys at x + t u (one FP32 fused multiply-add; a species'
    frozen axes stay put).  Checked against the keys of the predicted
    positions, leaving out the few particles whose predicted position lies
    within a few ulps of a cell face (where the FP32 rounding decides)."""
    static, dyn = workloads.make_parameters((16, 12, 8), wind=(8.0, 6.0, 4.0), dt=0.05)
    p = oracle_params(static, dyn)
    S, P = 2, 6000
    x, u, act = random_particles(p, S=S, P=P, seed=9, inactive_frac=0.05, vmax=0.9)
    sp = species(S)
    sp.update_x[1, 2] = False                                     # species 1 frozen along z
    cfg = SpeciesConfig(sp.charge, sp.mass, sp.weight, sp.update_x, sp.update_u)
    parts = tiled_particles_from_arrays(x, u, act)
    desc, spk = make_desc(static, dyn), make_species(cfg, P)
    ahead = 0.1
    xs, us, ao, starts, perm = _rebin(desc, spk, parts, 4, ahead)
    x32, u32 = x.astype(np.float32).astype(np.float64), u.astype(np.float32).astype(np.float64)
    xp = x32 + np.float64(np.float32(ahead)) * u32 * sp.update_x[:, None, :]
    xp = xp.astype(np.float32).astype(np.float64)
    near = np.zeros((S, P), bool)
    for a in range(3):
        t = (xp[..., a] + 0.5 * p.wind[a]) / p.d[a]
        near |= np.abs(t - np.round(t)) < 1e-5
    want, tiles, span, vol = _want_keys(xp, act, p, 4)
    k = want.reshape(-1)[perm]
    ok = ~near.reshape(-1)[perm]
    assert np.array_equal(np.sort(perm), np.arange(S * P))
    kk = k[ok]
    assert np.all(kk[1:] >= kk[:-1])
    # every range holds exactly the particles of its (species, tile)
    for b in range(S * (tiles + 1)):
        sl = slice(starts[b], starts[b + 1])
        lo = (b // (tiles + 1)) * (span + 1) + (b % (tiles + 1)) * vol
        hi = lo + (vol if b % (tiles + 1) < tiles else 1)
        kb, okb = k[sl], ok[sl]
        assert np.all((kb[okb] >= lo) & (kb[okb] < hi)), b
    np.testing.assert_array_equal(xs.reshape(3, -1).cpu().numpy(),
                                  np.moveaxis(x, -1, 0).reshape(3, -1).astype(np.float32)[:, perm])
