"""pic_energy_noalloc (partial sums in a static device array) against
pic_energy (partial sums in a caller workspace) through the C ABI: two calls
in a row, the second on a smaller grid and particle set, each equal to the
workspace entry point's result (same kernels and launch shapes, no atomics:
bit-identical FP64 sums) and to the FP64 oracle."""


import numpy as np
import pytest
import torch

from oracle import pic_oracle as O
from pypic3d_amd import _native as N
from pypic3d_amd import workloads
from pypic3d_amd._tensors import make_desc, make_species, particle_planes, pic_particles, tiled_particles_from_arrays
from pypic3d_amd.particles.particle_class import SpeciesConfig
from tests.gpu_helpers import gpu_field, oracle_params
from tests.helpers import random_field, random_particles, species

pytestmark = pytest.mark.gpu


def _case(N_, P, seed):
    static, dyn = workloads.make_parameters(N_, wind=tuple(0.5 * n for n in N_), dt=0.05)
    p = oracle_params(static, dyn)
    x, u, act = random_particles(p, S=2, P=P, seed=seed, vmax=0.03, inactive_frac=0.1)
    sp = species(2)
    E = [random_field(p, seed + c) for c in range(3)]
    B = [random_field(p, seed + 3 + c) for c in range(3)]
    cfg = SpeciesConfig(sp.charge, sp.mass, sp.weight, sp.update_x, sp.update_u)
    parts = tiled_particles_from_arrays(x, u, act)
    want = O.compute_energy(u, act, sp, E, B, p)
    return static, dyn, parts, cfg, [gpu_field(c)[0, 0, 0] for c in E], [gpu_field(c)[0, 0, 0] for c in B], want


def _both(static, dyn, parts, cfg, E, B):
    desc = make_desc(static, dyn)
    xp, up, act = particle_planes(parts)
    sp = make_species(cfg, xp[0].shape[1])
    pp = pic_particles(xp, up, act)
    s = N.stream_handle()
    nbytes = int(N.lib().pic_energy_workspace_bytes(desc, sp))
    ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    with_ws = torch.full((3,), -1.0, dtype=torch.float64, device="cuda")
    no_alloc = torch.full((3,), -1.0, dtype=torch.float64, device="cuda")
    N.call("pic_energy", desc, sp, pp, N.p3(E), N.p3(B), N.ptr(with_ws), N.ptr(ws), ws.numel(), s)
    N.call("pic_energy_noalloc", desc, sp, pp, N.p3(E), N.p3(B), N.ptr(no_alloc), s)
    torch.cuda.synchronize()
    return with_ws.cpu().numpy(), no_alloc.cpu().numpy()


def test_noalloc_energy_equals_workspace_energy_on_two_calls():
    big = _case((40, 36, 32), 200_000, 3)          # 46,080 cells, 400,000 slots: all 1024 partials in use
    small = _case((12, 10, 8), 3_000, 5)
    for static, dyn, parts, cfg, E, B, want in (big, small):
        a, b = _both(static, dyn, parts, cfg, E, B)
        np.testing.assert_array_equal(a, b)
        for got, w in zip(b, want):
            assert abs(got - w) <= 1e-10 * abs(w)


def test_noalloc_energy_rejects_invalid_arguments():
    static, dyn, parts, cfg, E, B, _ = _case((12, 10, 8), 100, 7)
    desc = make_desc(static, dyn)
    xp, up, act = particle_planes(parts)
    sp = make_species(cfg, xp[0].shape[1])
    rc = N.lib().pic_energy_noalloc(desc, sp, pic_particles(xp, up, act), N.p3(E), N.p3(B), None,
                                    N.stream_handle())
    assert rc < 0
