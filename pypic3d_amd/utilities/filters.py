"""Digital and bilinear filters (PyPIC3D/utilities/filters.py:234-317).

The interior ``g:-g`` is replaced by a 3x3x3 VALID correlation over the
stencil window ``g-1 .. -g+1``; ghost cells are left unchanged.
"""

from __future__ import annotations

from .. import _native as N
from .._tensors import as_field3, is_f64, like_input, shape_desc


def digital_filter(phi, alpha, num_guard_cells=1):
    """Centre alpha, six faces (1-alpha)/6 (filters.py:259-291)."""
    f = as_field3(phi, "phi")
    out = f.clone()
    N.call(N.op("pic_filter_digital", is_f64(f)), shape_desc(f.shape, int(num_guard_cells)), N.ptr(f), N.ptr(out),
           float(alpha), N.stream_handle())
    return like_input(out, phi)


def bilinear_filter(phi, num_guard_cells=1):
    """(1,2,1)^3 / 64 (filters.py:234-256)."""
    f = as_field3(phi, "phi")
    out = f.clone()
    N.call(N.op("pic_filter_bilinear", is_f64(f)), shape_desc(f.shape, int(num_guard_cells)), N.ptr(f), N.ptr(out),
           N.stream_handle())
    return like_input(out, phi)


def digital_filter_vector(field, alpha, num_guard_cells=1):
    return tuple(digital_filter(c, alpha, num_guard_cells) for c in field)


def bilinear_filter_vector(field, num_guard_cells=1):
    return tuple(bilinear_filter(c, num_guard_cells) for c in field)
