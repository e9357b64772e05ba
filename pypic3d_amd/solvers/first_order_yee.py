"""First-order Yee updates (PyPIC3D/solvers/first_order_yee.py:12-162).

``update_B`` advances B by half a step (dt/2) from curl E (forward
differences); ``update_E`` advances E by dt from C^2 curl B - J/eps (backward
differences), then applies the digital filter and conducting walls.  The
digital filter is the identity at alpha = 1 (centre weight 1, faces 0), so it
is skipped there.  With a ``pml_state`` = (e_memory, b_memory, profiles)
(boundary_conditions/PML.py) the six curl derivatives are coordinate-stretched
inside the kernels (first_order_yee.py:54-59, 133-138) and the updated
memory is returned in a new state.
"""

from __future__ import annotations

import ctypes
import weakref

import torch

from .. import _native as N
from .._tensors import as_field3, check_fields, is_f64, like_input, make_desc


def _check_pml(pml_state):
    """pml_state must be None or the 3-tuple (e_memory[6], b_memory[6],
    profiles[3]) of PML.initialize_(tiled_)pml_state."""
    if pml_state is None:
        return
    if not (isinstance(pml_state, (tuple, list)) and len(pml_state) == 3 and len(pml_state[0]) == 6
            and len(pml_state[1]) == 6 and len(pml_state[2]) == 3):
        raise ValueError("pml_state must be (e_memory[6], b_memory[6], profiles[3])")


def pml_memory3(desc, mem):
    """The six memory arrays as contiguous (Nx, Ny, Nz) float32 tensors."""
    out = []
    n = tuple(int(v) for v in desc.n)
    for m in mem:
        t = m.reshape(n)
        if t.dtype != torch.float32 or not t.is_contiguous() or not t.is_cuda:
            raise ValueError("PML memory must be contiguous float32 CUDA tensors of the interior shape")
        out.append(t)
    return out


_SIGMA_LINES = {}


def sigma_lines(desc, sigma3):
    """Profile a along its own axis (dims[a] float32 values, contiguous) for
    each of the three field-shaped profiles, when every profile is constant
    along the other two axes over the interior (PML.py's are); else None.
    Checked once per profile tensor and version (an in-place change of a
    profile, through any view, bumps its version and triggers a new check)."""
    key = tuple(id(t) for t in sigma3)
    hit = _SIGMA_LINES.get(key)
    if hit is not None:
        refs, versions, lines = hit
        if all(r() is t and t._version == v for r, t, v in zip(refs, sigma3, versions)):
            return lines
    g = int(desc.g)
    dims = tuple(int(v) for v in desc.dims)
    n = tuple(int(v) for v in desc.n)
    lines = []
    for a, t in enumerate(sigma3):
        f = t.reshape(dims)
        at = [g, g, g]
        at[a] = slice(None)
        line = f[tuple(at)].contiguous()
        shape = [1, 1, 1]
        shape[a] = n[a]
        want = line[g:g + n[a]].reshape(shape).expand(n)
        if not torch.equal(f[g:g + n[0], g:g + n[1], g:g + n[2]], want):
            lines = None
            break
        lines.append(line)
    for k in [k for k, (refs, _, _) in _SIGMA_LINES.items() if any(r() is None for r in refs)]:
        del _SIGMA_LINES[k]             # entries of profiles that no longer exist
    _SIGMA_LINES[key] = (tuple(weakref.ref(t) for t in sigma3), tuple(t._version for t in sigma3), lines)
    return lines


def pml_struct(mem6, sigma3, desc=None):
    """PicPml over six (Nx, Ny, Nz) memory tensors and three field-shaped
    profiles (with their axis lines when desc is given and the profiles are
    separable).  The returned struct keeps the line tensors alive."""
    s = N.PicPml()
    for q in range(6):
        s.mem[q] = mem6[q].data_ptr()
    for a in range(3):
        s.sigma[a] = sigma3[a].data_ptr()
    lines = sigma_lines(desc, sigma3) if desc is not None else None
    if lines is not None:
        for a in range(3):
            s.sigma_line[a] = lines[a].data_ptr()
        s._keep = lines
    return s


def yee_b(desc, E3, B3, out3, do_filter, stream, scratch=None, pml=None):
    """B half step on 3-D tensors into out3 (out3 must not alias E3).  pml:
    None or (memory tensors[6], profile tensors[3]), memory updated in place.
    float64 fields run the FP64 path (without PML)."""
    f64 = is_f64(E3[0])
    if f64 and pml is not None:
        raise ValueError("the PML runs in FP32 (float32 fields)")
    if pml is None:
        N.call(N.op("pic_yee_update_b", f64), desc, N.p3(E3), N.p3(B3), N.p3(out3), stream)
    else:
        N.call("pic_yee_update_b_pml", desc, N.p3(E3), N.p3(B3), N.p3(out3),
               ctypes.byref(pml_struct(*pml, desc=desc)), stream)
    if do_filter and desc.alpha != 1.0:
        tmp = scratch if scratch is not None else [c.clone() for c in out3]
        for c, t in zip(out3, tmp):
            N.call(N.op("pic_filter_digital", f64), desc, N.ptr(c), N.ptr(t), float(desc.alpha), stream)
            N.call(N.op("pic_halo_refresh", f64), desc, N.ptr(t), 0, stream)
            c.copy_(t)
    return out3


def yee_e(desc, E3, B3, J3, out3, stream, scratch=None, pml=None):
    """E full step on 3-D tensors into out3 (out3 must not alias B3 or J3).
    pml as for yee_b."""
    f64 = is_f64(E3[0])
    if f64 and pml is not None:
        raise ValueError("the PML runs in FP32 (float32 fields)")
    if pml is None:
        N.call(N.op("pic_yee_update_e", f64), desc, N.p3(E3), N.p3(B3), N.p3(J3), N.p3(out3), stream)
    else:
        N.call("pic_yee_update_e_pml", desc, N.p3(E3), N.p3(B3), N.p3(J3), N.p3(out3),
               ctypes.byref(pml_struct(*pml, desc=desc)), stream)
    if desc.alpha != 1.0:
        tmp = scratch if scratch is not None else [c.clone() for c in out3]
        for c, t in zip(out3, tmp):
            N.call(N.op("pic_filter_digital", f64), desc, N.ptr(c), N.ptr(t), float(desc.alpha), stream)
            c.copy_(t)
    walls = {0: (1, 2), 1: (0, 2), 2: (0, 1)}
    changed = desc.alpha != 1.0
    for axis in range(3):
        if desc.bc[axis] == 1:
            changed = True
            for comp in walls[axis]:
                N.call(N.op("pic_zero_boundary", f64), desc, N.ptr(out3[comp]), axis, stream)
    # the update kernels refresh the guard cells themselves; only a filter or the
    # conducting walls change the interior after that
    if changed:
        for c in out3:
            N.call(N.op("pic_halo_refresh", f64), desc, N.ptr(c), 0, stream)
    return out3


def _pml_args(desc, pml_state, which):
    """(cloned memory[6] as (Nx, Ny, Nz) tensors, profiles[3] as field tensors,
    the cloned memories in the caller's shapes) for which = 0 (E) or 1 (B)."""
    mem_in = pml_state[which]
    clones = [m.clone() for m in mem_in]
    mem3 = pml_memory3(desc, clones)
    prof3 = [as_field3(pr, "PML profile") for pr in pml_state[2]]
    check_fields(desc, *prof3, dtype=torch.float32)
    return mem3, prof3, tuple(clones)


def update_B(E_tiles, B_tiles, static_parameters, dynamic_parameters, pml_state=None, do_filter=False):
    """Returns (B, pml_state) (first_order_yee.py:96-162)."""
    _check_pml(pml_state)
    desc = make_desc(static_parameters, dynamic_parameters)
    E3 = [as_field3(c, "E") for c in E_tiles]
    B3 = [as_field3(c, "B") for c in B_tiles]
    check_fields(desc, *E3, *B3)
    if pml_state is not None and is_f64(E3[0]):
        raise ValueError("the PML runs in FP32 (float32 fields)")
    out = [b.new_empty(b.shape) for b in B3]
    pml = None
    if pml_state is not None:
        mem3, prof3, clones = _pml_args(desc, pml_state, 1)
        pml = (mem3, prof3)
        pml_state = (pml_state[0], clones, pml_state[2])
    yee_b(desc, E3, B3, out, do_filter, N.stream_handle(), pml=pml)
    return tuple(like_input(o, t) for o, t in zip(out, B_tiles)), pml_state


def update_E(E_tiles, B_tiles, J_tiles, static_parameters, dynamic_parameters, pml_state=None):
    """Returns (E, pml_state) (first_order_yee.py:12-93)."""
    _check_pml(pml_state)
    desc = make_desc(static_parameters, dynamic_parameters)
    E3 = [as_field3(c, "E") for c in E_tiles]
    B3 = [as_field3(c, "B") for c in B_tiles]
    J3 = [as_field3(c, "J") for c in J_tiles]
    check_fields(desc, *E3, *B3, *J3)
    if pml_state is not None and is_f64(E3[0]):
        raise ValueError("the PML runs in FP32 (float32 fields)")
    out = [e.new_empty(e.shape) for e in E3]
    pml = None
    if pml_state is not None:
        mem3, prof3, clones = _pml_args(desc, pml_state, 0)
        pml = (mem3, prof3)
        pml_state = (clones, pml_state[1], pml_state[2])
    yee_e(desc, E3, B3, J3, out, N.stream_handle(), pml=pml)
    return tuple(like_input(o, t) for o, t in zip(out, E_tiles)), pml_state
