"""GPU PML (boundary_conditions/PML.py, first_order_yee.py:54-59, 133-138)
against the FP64 oracle: one update_B / update_E with random memories and
profiles on all six walls, the standalone derivative stretch, the 1-D pulse
absorption run, and the allocation-free runner against the functional step."""

import numpy as np
import pytest
import torch

from oracle import pic_oracle as O
from pypic3d_amd import workloads
from pypic3d_amd.boundary_conditions.PML import (apply_tiled_pml_to_b_curl, initialize_tiled_pml_state,
                                                 load_pml_from_toml)
from pypic3d_amd.evolve import time_loop_electrodynamic
from pypic3d_amd.runner import ElectrodynamicRunner
from pypic3d_amd.solvers.first_order_yee import update_B, update_E
from tests.gpu_helpers import TOL, gpu_field, host, oracle_params, rel_err
from tests.helpers import random_field

pytestmark = pytest.mark.gpu

WALLS = [{"wall": "-x", "thickness": 3, "order": 3.0, "sigma_max": 8.0},
         {"wall": "+x", "thickness": 2, "order": 2.0},
         {"wall": "-y", "thickness": 2, "order": 3.0, "sigma_max": 5.0},
         {"wall": "+y", "thickness": 3, "order": 1.0, "sigma_max": 6.0},
         {"wall": "-z", "thickness": 2, "order": 3.0, "sigma_max": 7.0},
         {"wall": "+z", "thickness": 1, "order": 3.0, "sigma_max": 9.0}]


def _setup(N_=(10, 8, 6), seed=1, walls=WALLS, bc=(0, 0, 0)):
    static, dyn = workloads.make_parameters(N_, wind=(5.0, 4.0, 3.0), dt=0.05,
                                            bc=bc)
    p = oracle_params(static, dyn)
    _, _, _, _, prof = load_pml_from_toml(walls, static, dyn)
    state = initialize_tiled_pml_state(static, dyn, prof)
    rng = np.random.default_rng(seed)
    mem = [rng.standard_normal(p.N).astype(np.float32).astype(np.float64) * 0.1 for _ in range(12)]
    state = (tuple(torch.tensor(m, dtype=torch.float32, device="cuda").view(1, 1, 1, *p.N) for m in mem[:6]),
             tuple(torch.tensor(m, dtype=torch.float32, device="cuda").view(1, 1, 1, *p.N) for m in mem[6:]),
             state[2])
    oprof = [O.pml_ghosted(pr, p) for pr in O.pml_profiles(O.pml_layers(walls, p), p)]
    return static, dyn, p, state, mem, oprof


def test_profiles_equal_the_oracle_profiles():
    static, dyn, p, state, mem, oprof = _setup()
    for a in range(3):
        np.testing.assert_allclose(host(state[2][a]), oprof[a], rtol=1e-6, atol=0)


def test_one_update_B_and_update_E_match_the_oracle():
    static, dyn, p, state, mem, oprof = _setup()
    E = [random_field(p, 10 + c) for c in range(3)]
    B = [random_field(p, 20 + c) for c in range(3)]
    J = [random_field(p, 30 + c) for c in range(3)]
    Bg, st1 = update_B(tuple(map(gpu_field, E)), tuple(map(gpu_field, B)), static, dyn, state)
    Bo, mb = O.update_B(E, B, p, pml=(mem[6:], oprof))
    for c in range(3):
        assert rel_err(host(Bg[c]), Bo[c]) <= TOL
    for q in range(6):
        assert rel_err(host(st1[1][q]), mb[q]) <= TOL
        assert torch.equal(st1[0][q], state[0][q])            # the E memory is untouched
    assert not torch.equal(st1[1][0], state[1][0])          # inputs are not modified in place
    Eg, st2 = update_E(tuple(map(gpu_field, E)), tuple(map(gpu_field, B)), tuple(map(gpu_field, J)), static, dyn,
                       state)
    Eo, me = O.update_E(E, B, J, p, pml=(mem[:6], oprof))
    for c in range(3):
        assert rel_err(host(Eg[c]), Eo[c]) <= TOL
    for q in range(6):
        assert rel_err(host(st2[0][q]), me[q]) <= TOL


def test_standalone_stretch_matches_the_oracle():
    static, dyn, p, state, mem, oprof = _setup(seed=4)
    rng = np.random.default_rng(5)
    dv = [rng.standard_normal(p.N) for _ in range(6)]
    curl, st = apply_tiled_pml_to_b_curl(tuple(torch.tensor(v, dtype=torch.float32, device="cuda") for v in dv),
                                         static, dyn, state)
    want, mb = O._pml_curl(dv, mem[6:], oprof, p)
    for c in range(3):
        assert rel_err(curl[c].double().cpu().numpy(), want[c]) <= TOL
    for q in range(6):
        assert rel_err(host(st[1][q]), mb[q]) <= TOL


def test_pulse_is_absorbed_like_the_oracle():
    walls = [{"wall": "-x", "thickness": 8, "order": 3.0, "sigma_max": 60.0},
             {"wall": "+x", "thickness": 8, "order": 3.0, "sigma_max": 60.0}]
    static, dyn = workloads.make_parameters((40, 1, 1), wind=(1.0, 1.0, 1.0), d=(1.0 / 40, 1.0, 1.0),
                                            dt=0.5 / 40, bc=(1, 0, 0))
    p = oracle_params(static, dyn)
    _, _, _, _, prof = load_pml_from_toml(walls, static, dyn)
    state = initialize_tiled_pml_state(static, dyn, prof)
    g = p.g
    xv = -0.5 + (np.arange(40) + 0.5) / 40
    pulse = np.exp(-((xv + 0.30) / 0.04) ** 2)
    E = [np.zeros(p.shape) for _ in range(3)]
    B = [np.zeros(p.shape) for _ in range(3)]
    E[1][g:-g, g, g] = pulse
    B[2][g:-g, g, g] = pulse
    E = [O.refresh(O.refresh(c, p).astype(np.float32).astype(np.float64), p) for c in E]
    B = [O.refresh(O.refresh(c, p).astype(np.float32).astype(np.float64), p) for c in B]
    J = [np.zeros(p.shape) for _ in range(3)]
    oprof = [O.pml_ghosted(pr, p) for pr in O.pml_profiles(O.pml_layers(walls, p), p)]
    me = [np.zeros(p.N) for _ in range(6)]
    mb = [np.zeros(p.N) for _ in range(6)]
    Eg, Bg, Jg = tuple(map(gpu_field, E)), tuple(map(gpu_field, B)), tuple(map(gpu_field, J))
    for _ in range(120):
        Eg, state = update_E(Eg, Bg, Jg, static, dyn, state)
        Bg, state = update_B(Eg, Bg, static, dyn, state)
        E, me = O.update_E(E, B, J, p, pml=(me, oprof))
        B, mb = O.update_B(E, B, p, pml=(mb, oprof))

    def energy(E, B):
        A = slice(g, -g)
        return 0.5 * sum(np.sum(c[A, A, A] ** 2) for c in list(E) + list(B)) * p.d[0]

    e_gpu = energy([host(c) for c in Eg], [host(c) for c in Bg])
    e_ora = energy(E, B)
    e0 = 0.5 * 2 * np.sum(pulse ** 2) * p.d[0]
    assert e_ora < 0.1 * e0
    assert abs(e_gpu - e_ora) <= 1e-3 * e0


def test_runner_with_pml_matches_the_functional_step():
    static, dyn, p, state, mem, oprof = _setup(N_=(12, 10, 8), walls=WALLS[:4])
    specs = [workloads.SpeciesSpec(8, weight=0.01, vth=(0.1,) * 3)]
    parts, cfg = workloads.plasma(static, dyn, specs, seed=3)
    fields = workloads.empty_fields(static, dyn)
    E0 = tuple(gpu_field(random_field(p, 40 + c, scale=0.1)) for c in range(3))
    fields = (E0,) + fields[1:6] + (state,) + fields[7:]
    r = ElectrodynamicRunner(parts, cfg, fields, static, dyn)
    ref_parts, ref_fields = parts, fields
    for _ in range(3):
        r.step()
        ref_parts, ref_fields = time_loop_electrodynamic(ref_parts, cfg, ref_fields, static, dyn)
    r_parts, r_fields = r.state()
    for k in (0, 1):
        for c in range(3):
            assert rel_err(host(r_fields[k][c]), host(ref_fields[k][c])) <= 1e-5
    for which in (0, 1):
        for q in range(6):
            assert rel_err(host(r_fields[6][which][q]), host(ref_fields[6][which][q])) <= 1e-5


def test_axis_lines_give_the_same_bits_as_field_shaped_profiles():
    """PML.py's profiles vary along their own axis only, so the kernels read
    one line per axis (PicPml.sigma_line); the results equal, bit for bit,
    those of the field-shaped profiles, and a profile that varies along
    another axis is detected and keeps the field-shaped path."""
    import ctypes
    from pypic3d_amd import _native as N
    from pypic3d_amd._tensors import make_desc
    from pypic3d_amd.solvers.first_order_yee import pml_struct, sigma_lines
    static, dyn, p, state, mem, oprof = _setup(seed=7)
    desc = make_desc(static, dyn)
    prof = state[2]
    assert sigma_lines(desc, prof) is not None
    E = [gpu_field(random_field(p, 40 + c)).reshape(p.shape) for c in range(3)]
    B = [gpu_field(random_field(p, 50 + c)).reshape(p.shape) for c in range(3)]

    def run(with_lines, profiles):
        m = [t.reshape(p.N).clone() for t in state[1]]
        out = [torch.empty_like(c) for c in B]
        s = pml_struct(m, profiles, desc=desc if with_lines else None)
        assert (s.sigma_line[0] is not None) == with_lines
        N.call("pic_yee_update_b_pml", desc, N.p3(E), N.p3(B), N.p3(out), ctypes.byref(s), N.stream_handle())
        torch.cuda.synchronize()
        return out, m

    o1, m1 = run(True, prof)
    o0, m0 = run(False, prof)
    for a, b in zip(o1 + m1, o0 + m0):
        assert torch.equal(a, b)
    bent = tuple(t.clone() for t in prof)
    g = int(desc.g)
    at = (g + 1, g + 2, g + 1)
    keep = bent[0].view(p.shape)[at].clone()
    bent[0].view(p.shape)[at] = keep + 1.0                  # the x profile now varies along y
    assert sigma_lines(desc, bent) is None
    bent[0].view(p.shape)[at] = keep                        # restored in place: checked again
    assert sigma_lines(desc, bent) is not None
