"""Host side of the PML (PML.py:17-248): profiles built by
load_pml_from_toml equal the oracle's, the reference's validation messages,
and the single-tile state layout (no GPU needed)."""

import numpy as np
import pytest

from oracle import pic_oracle as O
from pypic3d_amd import workloads
from pypic3d_amd.boundary_conditions import PML
from tests.gpu_helpers import oracle_params


def _params(N_=(10, 8, 6)):
    static, dyn = workloads.make_parameters(N_, wind=(5.0, 4.0, 3.0), dt=0.05)
    return static, dyn, oracle_params(static, dyn)


def test_profiles_match_the_oracle_and_the_flags_follow_the_walls():
    static, dyn, p = _params()
    walls = [{"wall": "-x", "thickness": 3, "order": 3.0, "sigma_max": 8.0}, {"wall": "+y", "thickness": 2},
             {"wall": "+z", "thickness": 1, "target_reflection": 1e-6}]
    active, px, py, pz, prof = PML.load_pml_from_toml(walls, static, dyn, device="cpu")
    assert (active, px, py, pz) == (True, True, True, True)
    want = O.pml_profiles(O.pml_layers(walls, p), p)
    for a in range(3):
        assert tuple(prof[a].shape) == (12, 10, 8)
        np.testing.assert_allclose(prof[a].double().numpy(), want[a], rtol=1e-7, atol=0)
    active, px, py, pz, prof = PML.load_pml_from_toml(None, static, dyn, device="cpu")
    assert (active, px, py, pz) == (False, False, False, False)
    assert not any(bool(t.any()) for t in prof)


def test_validation_messages():
    static, dyn, p = _params()
    with pytest.raises(ValueError, match="Invalid PML wall"):
        PML.load_pml_from_toml({"wall": "up", "thickness": 1}, static, dyn, device="cpu")
    with pytest.raises(ValueError, match="Duplicate PML wall"):
        PML.load_pml_from_toml([{"wall": "-x", "thickness": 1}, {"wall": "-x", "thickness": 2}], static, dyn,
                               device="cpu")
    with pytest.raises(ValueError, match="exceeds active cells"):
        PML.load_pml_from_toml({"wall": "+z", "thickness": 7}, static, dyn, device="cpu")
    with pytest.raises(ValueError, match="must be positive"):
        PML.load_pml_from_toml({"wall": "+z", "thickness": -1}, static, dyn, device="cpu")


def test_tiled_state_is_the_single_tile_layout():
    static, dyn, p = _params()
    _, _, _, _, prof = PML.load_pml_from_toml({"wall": "-x", "thickness": 2, "sigma_max": 3.0}, static, dyn,
                                              device="cpu")
    e_mem, b_mem, tiled = PML.initialize_tiled_pml_state(static, dyn, prof)
    assert len(e_mem) == len(b_mem) == 6 and len(tiled) == 3
    assert all(tuple(m.shape) == (1, 1, 1, 10, 8, 6) for m in e_mem + b_mem)
    assert all(tuple(t.shape) == (1, 1, 1) + p.shape for t in tiled)
    np.testing.assert_array_equal(tiled[0][0, 0, 0].double().numpy(), O.pml_ghosted(prof[0].double().numpy(), p))
    e_mem, b_mem, same = PML.initialize_pml_state(dyn, prof)
    assert all(tuple(m.shape) == (10, 8, 6) for m in e_mem + b_mem) and same is prof
