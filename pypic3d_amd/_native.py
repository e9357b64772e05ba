"""ctypes binding of libpypic3d_hip.so (C ABI declared in include/pypic3d_amd.h).

The library is loaded lazily, after ``torch`` (so it binds to the HIP runtime
torch already loaded; both carry the SONAME libamdhip64.so.7).  There is no
fallback: if the library is missing or no GPU is present, calls raise.
"""

from __future__ import annotations

import ctypes
import os

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpypic3d_hip.so")
MAX_SPECIES = 16
PUSHERS = {"boris": 0, "higuera_cary": 1}
FUSED_VARIANTS = 6
ABI_VERSION = 12

c_int = ctypes.c_int
c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p
P3 = c_vp * 3


class PicDesc(ctypes.Structure):
    _fields_ = [
        ("n", c_i32 * 3), ("g", c_i32), ("dims", c_i32 * 3), ("shape_factor", c_i32),
        ("relativistic", c_i32), ("bc", c_i32 * 3), ("pbc", c_i32 * 3), ("variant", c_i32),
        ("d", c_dbl * 3), ("wind", c_dbl * 3), ("org_c", c_dbl * 3), ("sp_c", c_dbl * 3),
        ("org_v", c_dbl * 3), ("sp_v", c_dbl * 3),
        ("dt", c_dbl), ("C", c_dbl), ("eps", c_dbl), ("mu", c_dbl), ("alpha", c_dbl),
        ("pusher", c_i32), ("pad1", c_i32),
    ]


class PicSpecies(ctypes.Structure):
    _fields_ = [
        ("count", c_i32), ("slots", c_i32),
        ("charge", c_dbl * MAX_SPECIES), ("mass", c_dbl * MAX_SPECIES), ("weight", c_dbl * MAX_SPECIES),
        ("update_x", (ctypes.c_uint8 * 3) * MAX_SPECIES), ("update_u", (ctypes.c_uint8 * 3) * MAX_SPECIES),
    ]


class PicParticles(ctypes.Structure):
    _fields_ = [("x", c_vp * 3), ("u", c_vp * 3), ("active", c_vp)]


class PicPml(ctypes.Structure):
    _fields_ = [("mem", c_vp * 6), ("sigma", c_vp * 3), ("sigma_line", c_vp * 3)]


DP = ctypes.POINTER(PicDesc)
PMLP = ctypes.POINTER(PicPml)
SP = ctypes.POINTER(PicSpecies)
PP = ctypes.POINTER(PicParticles)

# name -> (restype, argtypes); must match include/pypic3d_amd.h
SIGNATURES = {
    "pic_abi_version": (c_int, []),
    "pic_fill_f32": (c_int, [c_vp, c_i64, ctypes.c_float, c_vp]),
    "pic_halo_refresh": (c_int, [DP, c_vp, c_int, c_vp]),
    "pic_halo_fold": (c_int, [DP, c_vp, c_int, c_vp]),
    "pic_zero_boundary": (c_int, [DP, c_vp, c_int, c_vp]),
    "pic_constant_boundary": (c_int, [DP, c_vp, c_int, c_vp]),
    "pic_filter_digital": (c_int, [DP, c_vp, c_vp, c_dbl, c_vp]),
    "pic_filter_bilinear": (c_int, [DP, c_vp, c_vp, c_vp]),
    "pic_yee_update_b": (c_int, [DP, P3, P3, P3, c_vp]),
    "pic_yee_update_e": (c_int, [DP, P3, P3, P3, P3, c_vp]),
    "pic_push": (c_int, [DP, SP, PP, P3, P3, c_vp, c_vp, P3, c_vp]),
    "pic_advance_positions": (c_int, [DP, SP, PP, c_dbl, P3, c_vp]),
    "pic_particle_boundaries": (c_int, [DP, SP, PP, c_vp]),
    "pic_deposit_esirkepov": (c_int, [DP, SP, PP, P3, c_vp]),
    "pic_deposit_direct": (c_int, [DP, SP, PP, P3, c_vp]),
    "pic_deposit_rho": (c_int, [DP, SP, PP, c_vp, c_vp]),
    "pic_deposit_direct_tiled": (c_int, [DP, SP, PP, P3, c_i32, c_vp, c_vp]),
    "pic_deposit_rho_tiled": (c_int, [DP, SP, PP, c_vp, c_i32, c_vp, c_vp]),
    "pic_fused_push_esirkepov": (c_int, [DP, SP, PP, P3, P3, c_vp, c_vp, PP, P3, c_vp]),
    "pic_fused_push_esirkepov_tiled": (c_int, [DP, SP, PP, P3, P3, c_vp, c_vp, PP, P3, c_i32, c_vp, c_i32, c_i32,
                                               c_vp, c_vp]),
    "pic_fused_push_esirkepov_tiled_gather": (c_int, [DP, SP, PP, c_vp, P3, P3, c_vp, c_vp, PP, P3, c_i32, c_vp,
                                                      c_i32, c_i32, c_vp, c_vp]),
    "pic_fused_push_direct_tiled_gather": (c_int, [DP, SP, PP, c_vp, P3, P3, c_vp, c_vp, PP, P3, c_i32, c_vp,
                                                   c_i32, c_vp, c_vp]),
    "pic_fused_push_direct_tiled": (c_int, [DP, SP, PP, P3, P3, c_vp, c_vp, PP, P3, c_i32, c_vp, c_i32, c_vp,
                                            c_vp]),
    "pic_tile_exact_grid": (c_int, [DP]),
    "pic_cell_keys": (c_int, [DP, SP, PP, c_vp, c_vp]),
    "pic_tile_count": (c_i64, [DP, c_i32]),
    "pic_tile_keys": (c_int, [DP, SP, PP, c_i32, c_vp, c_vp]),
    "pic_tile_ranges": (c_int, [DP, SP, c_i32, c_vp, c_vp, c_vp]),
    "pic_tile_rebin_workspace_bytes": (c_i64, [DP, SP, c_i32]),
    "pic_tile_rebin": (c_int, [DP, SP, PP, c_i32, c_dbl, PP, c_vp, c_vp, c_vp, c_i64, c_vp]),
    "pic_sort_workspace_bytes": (c_i64, [c_i64]),
    "pic_sort_permutation": (c_int, [c_vp, c_vp, c_vp, c_i64, c_i32, c_vp, c_i64, c_vp]),
    "pic_permute_particles": (c_int, [PP, PP, c_vp, c_i64, c_vp]),
    "pic_cg_workspace_bytes": (c_i64, [DP]),
    "pic_poisson_cg": (c_int, [DP, c_vp, c_vp, c_dbl, c_dbl, c_i32, c_vp, c_i64,
                               ctypes.POINTER(c_i32), ctypes.POINTER(c_dbl), c_vp]),
    "pic_yee_update_b_pml": (c_int, [DP, P3, P3, P3, PMLP, c_vp]),
    "pic_yee_update_e_pml": (c_int, [DP, P3, P3, P3, P3, PMLP, c_vp]),
    "pic_pml_stretch": (c_int, [DP, c_vp * 6, PMLP, c_vp]),
    "pic_efield_from_phi": (c_int, [DP, c_vp, P3, c_vp]),
    "pic_energy_workspace_bytes": (c_i64, [DP, SP]),
    "pic_energy": (c_int, [DP, SP, PP, P3, P3, c_vp, c_vp, c_i64, c_vp]),
    "pic_energy_noalloc": (c_int, [DP, SP, PP, P3, P3, c_vp, c_vp]),
    "pic_total_momentum": (c_int, [SP, PP, c_vp, c_vp, c_i64, c_vp]),
    "pic_total_charge": (c_int, [DP, c_vp, c_vp, c_vp, c_i64, c_vp]),
    "pic_halo_refresh_f64": (c_int, [DP, c_vp, c_int, c_vp]),
    "pic_halo_fold_f64": (c_int, [DP, c_vp, c_int, c_vp]),
    "pic_zero_boundary_f64": (c_int, [DP, c_vp, c_int, c_vp]),
    "pic_constant_boundary_f64": (c_int, [DP, c_vp, c_int, c_vp]),
    "pic_filter_digital_f64": (c_int, [DP, c_vp, c_vp, c_dbl, c_vp]),
    "pic_filter_bilinear_f64": (c_int, [DP, c_vp, c_vp, c_vp]),
    "pic_yee_update_b_f64": (c_int, [DP, P3, P3, P3, c_vp]),
    "pic_yee_update_e_f64": (c_int, [DP, P3, P3, P3, P3, c_vp]),
    "pic_push_f64": (c_int, [DP, SP, PP, P3, P3, c_vp, c_vp, P3, c_vp]),
    "pic_advance_positions_f64": (c_int, [DP, SP, PP, c_dbl, P3, c_vp]),
    "pic_particle_boundaries_f64": (c_int, [DP, SP, PP, c_vp]),
    "pic_deposit_esirkepov_f64": (c_int, [DP, SP, PP, P3, c_vp]),
    "pic_deposit_direct_f64": (c_int, [DP, SP, PP, P3, c_vp]),
    "pic_deposit_rho_f64": (c_int, [DP, SP, PP, c_vp, c_vp]),
    "pic_energy_f64": (c_int, [DP, SP, PP, P3, P3, c_vp, c_vp, c_i64, c_vp]),
    "pic_wrap_periodic": (c_int, [c_vp, c_i64, c_dbl, c_vp]),
    "pic_wrap_periodic_f64": (c_int, [c_vp, c_i64, c_dbl, c_vp]),
    "pic_batch_check": (c_int, [c_vp, c_i32, c_i64, c_i64]),
    "pic_batch_fused_push_esirkepov": (c_int, [c_vp, c_vp, c_i32, c_i32, c_i32, PP, P3, P3, PP, P3, c_vp]),
    "pic_batch_field_step": (c_int, [c_vp, c_i32, P3, P3, P3, P3, P3, c_vp]),
    "pic_fill_f64": (c_int, [c_vp, c_i64, c_dbl, c_vp]),
    "pic_periodic_expand_f64": (c_int, [DP, c_vp, c_vp, c_vp]),
    "pic_periodic_fold_add_f64": (c_int, [DP, c_vp, c_vp, c_vp]),
    "pic_adjoint_yee": (c_int, [DP, P3, P3, P3, c_vp * 6, c_vp]),
    "pic_adjoint_particles": (c_int, [DP, SP, PP, P3, P3, c_vp, c_vp, P3, P3, P3, P3, P3, c_i32, c_vp, c_vp,
                                      c_vp]),
}

_lib = None


class NativeLibraryMissing(RuntimeError):
    pass


def load(path: str = LIB_PATH):
    """Load (once) and type the C ABI.  Raises if the library is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise NativeLibraryMissing(
            f"{path} is missing: build it with `python -m pypic3d_amd._build` "
            "(there is no CPU fallback for the PIC hot path)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.pic_abi_version() != ABI_VERSION:
        raise NativeLibraryMissing(f"ABI version mismatch: library {lib.pic_abi_version()} != {ABI_VERSION}")
    _lib = lib
    return lib


def lib():
    return _lib if _lib is not None else load()


def check(rc: int, what: str):
    if rc == 0:
        return
    if rc < 0:
        raise ValueError(f"{what}: invalid argument or shape (code {rc})")
    raise RuntimeError(f"{what}: HIP error {rc}")


def call(name: str, *args):
    check(getattr(lib(), name)(*args), name)


def op(name: str, f64: bool) -> str:
    """Entry point name for FP32 or FP64 (``_f64`` suffix) tensors."""
    return name + "_f64" if f64 else name


def fill(t, value, stream):
    """Fill a CUDA tensor (float32 or float64) on the stream with a kernel."""
    f64 = t.dtype == torch.float64
    call("pic_fill_f64" if f64 else "pic_fill_f32", ptr(t), t.numel(), float(value), stream)


def stream_handle():
    return c_vp(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    return None if t is None else c_vp(t.data_ptr())


def p3(ts):
    return P3(*[t.data_ptr() for t in ts])
