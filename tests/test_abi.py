"""CPU checks of the C-ABI boundary: the library loads, exports every symbol
include/pypic3d_amd.h declares, and the ctypes structs match the C layout.
No compute call is made (there is no GPU here)."""

import ctypes
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pypic3d_amd.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pic_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import torch  # noqa: F401  (load order: torch's HIP runtime first)
    from pypic3d_amd import _native as N
    lib = N.load()
    names = _declared()
    assert len(names) >= 25
    for name in names:
        assert hasattr(lib, name), name
        assert name in N.SIGNATURES, f"{name} has no ctypes signature"
    assert set(N.SIGNATURES) == set(names)
    assert lib.pic_abi_version() == N.ABI_VERSION


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_struct_layout_matches_c(tmp_path):
    from pypic3d_amd import _native as N
    src = tmp_path / "sz.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "pypic3d_amd.h"\n'
                   'int main(){printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(PicDesc), sizeof(PicSpecies),'
                   ' sizeof(PicParticles), offsetof(PicDesc, d), offsetof(PicDesc, dt), offsetof(PicSpecies, update_u),'
                   ' sizeof(PicBatchSim), offsetof(PicBatchSim, species), offsetof(PicBatchSim, field_offset));}')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    from pypic3d_amd.batch import PicBatchSim
    want = [ctypes.sizeof(N.PicDesc), ctypes.sizeof(N.PicSpecies), ctypes.sizeof(N.PicParticles),
            N.PicDesc.d.offset, N.PicDesc.dt.offset, N.PicSpecies.update_u.offset,
            ctypes.sizeof(PicBatchSim), PicBatchSim.species.offset, PicBatchSim.field_offset.offset]
    assert [int(v) for v in out] == want


def test_descriptor_grid_lines_follow_reference_formula():
    from oracle import pic_oracle as O
    from pypic3d_amd import workloads
    from pypic3d_amd._tensors import make_desc
    static, dyn = workloads.make_parameters((8, 6, 1), wind=(4.0, 3.0, 1.0), dt=0.05)
    d = make_desc(static, dyn)
    p = O.Params(N=(8, 6, 1), wind=(4.0, 3.0, 1.0), g=2, dt=0.05)
    for a in range(3):
        assert (d.org_c[a], d.sp_c[a]) == O.axis_origin_spacing(p, a, "center")
        assert (d.org_v[a], d.sp_v[a]) == O.axis_origin_spacing(p, a, "vertex")
        assert d.dims[a] == p.shape[a]


def test_product_path_has_no_cpu_fallback(tmp_path):
    from pypic3d_amd import _native as N
    with pytest.raises(N.NativeLibraryMissing):
        N._lib_backup = N._lib
        try:
            N._lib = None
            N.load(str(tmp_path / "missing.so"))
        finally:
            N._lib = N._lib_backup


def test_product_does_not_import_oracle():
    pkg = os.path.join(ROOT, "pypic3d_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp")):
                text = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in text and "from oracle" not in text, f
