"""First-order Yee updates (PyPIC3D/solvers/first_order_yee.py:12-162).

``update_B`` advances B by half a step (dt/2) from curl E (forward
differences); ``update_E`` advances E by dt from C^2 curl B - J/eps (backward
differences), then applies the digital filter and conducting walls.  The
digital filter is the identity at alpha = 1 (centre weight 1, faces 0), so it
is skipped there.  PML is not part of this library: a non-None ``pml_state``
raises NotImplementedError.
"""

from __future__ import annotations

from .. import _native as N
from .._tensors import as_field3, check_fields, like_input, make_desc


def _check_pml(pml_state):
    if pml_state is not None:
        raise NotImplementedError("PML curl stretching is not implemented on the GPU path (pml_state must be None)")


def yee_b(desc, E3, B3, out3, do_filter, stream, scratch=None):
    """B half step on 3-D tensors into out3 (out3 must not alias E3)."""
    N.call("pic_yee_update_b", desc, N.p3(E3), N.p3(B3), N.p3(out3), stream)
    if do_filter and desc.alpha != 1.0:
        tmp = scratch if scratch is not None else [c.clone() for c in out3]
        for c, t in zip(out3, tmp):
            N.call("pic_filter_digital", desc, N.ptr(c), N.ptr(t), float(desc.alpha), stream)
            N.call("pic_halo_refresh", desc, N.ptr(t), 0, stream)
            c.copy_(t)
    return out3


def yee_e(desc, E3, B3, J3, out3, stream, scratch=None):
    """E full step on 3-D tensors into out3 (out3 must not alias B3 or J3)."""
    N.call("pic_yee_update_e", desc, N.p3(E3), N.p3(B3), N.p3(J3), N.p3(out3), stream)
    if desc.alpha != 1.0:
        tmp = scratch if scratch is not None else [c.clone() for c in out3]
        for c, t in zip(out3, tmp):
            N.call("pic_filter_digital", desc, N.ptr(c), N.ptr(t), float(desc.alpha), stream)
            c.copy_(t)
    walls = {0: (1, 2), 1: (0, 2), 2: (0, 1)}
    for axis in range(3):
        if desc.bc[axis] == 1:
            for comp in walls[axis]:
                N.call("pic_zero_boundary", desc, N.ptr(out3[comp]), axis, stream)
    for c in out3:
        N.call("pic_halo_refresh", desc, N.ptr(c), 0, stream)
    return out3


def update_B(E_tiles, B_tiles, static_parameters, dynamic_parameters, pml_state=None, do_filter=False):
    """Returns (B, pml_state) (first_order_yee.py:96-162)."""
    _check_pml(pml_state)
    desc = make_desc(static_parameters, dynamic_parameters)
    E3 = [as_field3(c, "E") for c in E_tiles]
    B3 = [as_field3(c, "B") for c in B_tiles]
    check_fields(desc, *E3, *B3)
    out = [b.new_empty(b.shape) for b in B3]
    yee_b(desc, E3, B3, out, do_filter, N.stream_handle())
    return tuple(like_input(o, t) for o, t in zip(out, B_tiles)), pml_state


def update_E(E_tiles, B_tiles, J_tiles, static_parameters, dynamic_parameters, pml_state=None):
    """Returns (E, pml_state) (first_order_yee.py:12-93)."""
    _check_pml(pml_state)
    desc = make_desc(static_parameters, dynamic_parameters)
    E3 = [as_field3(c, "E") for c in E_tiles]
    B3 = [as_field3(c, "B") for c in B_tiles]
    J3 = [as_field3(c, "J") for c in J_tiles]
    check_fields(desc, *E3, *B3, *J3)
    out = [e.new_empty(e.shape) for e in E3]
    yee_e(desc, E3, B3, J3, out, N.stream_handle())
    return tuple(like_input(o, t) for o, t in zip(out, E_tiles)), pml_state
