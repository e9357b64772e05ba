"""Ghost refresh / fold and conducting walls for one tile per GPU
(PyPIC3D/boundary_conditions/ghost_cells.py:615-736).

Functional like the reference: inputs are never modified, new tensors are
returned.  ``bc_type`` 0 selects the field boundary codes, 1 the particle
codes (ghost_cells.py:33-39).
"""

from __future__ import annotations

import torch

from .. import _native as N
from .._tensors import as_field3, is_f64, like_input, make_desc

BC_TYPE_FIELD = 0
BC_TYPE_PARTICLE = 1


def _check_g(static_parameters, num_guard_cells):
    if int(num_guard_cells) != int(static_parameters.guard_cells):
        raise ValueError("num_guard_cells must equal static_parameters.guard_cells on the GPU path")


def _desc_for(field, static_parameters):
    """Descriptor without dynamic parameters: only shape and BC codes matter."""
    g = int(static_parameters.guard_cells)
    d = N.PicDesc()
    for a in range(3):
        d.dims[a] = field.shape[a]
        d.n[a] = field.shape[a] - 2 * g
        d.bc[a] = int(static_parameters.boundary_conditions[a])
        d.pbc[a] = int(static_parameters.particle_boundary_conditions[a])
        d.d[a] = 1.0
        d.sp_c[a] = d.sp_v[a] = 1.0
    d.g = g
    d.shape_factor = 1
    return d


def _apply(kind, field_tiles, static_parameters, bc_type=0, axis=None):
    f = as_field3(field_tiles, "field_tiles")
    out = f.clone()
    d = _desc_for(out, static_parameters)
    s = N.stream_handle()
    f64 = is_f64(out)
    if kind == "refresh":
        N.call(N.op("pic_halo_refresh", f64), d, N.ptr(out), int(bc_type), s)
    elif kind == "fold":
        N.call(N.op("pic_halo_fold", f64), d, N.ptr(out), int(bc_type), s)
    elif kind == "zero":
        N.call(N.op("pic_zero_boundary", f64), d, N.ptr(out), int(axis), s)
    elif kind == "constant":
        N.call(N.op("pic_constant_boundary", f64), d, N.ptr(out), int(axis), s)
    return like_input(out, field_tiles)


def update_tiled_ghost_cells(field_tiles, static_parameters, num_guard_cells=2, bc_type=BC_TYPE_FIELD):
    _check_g(static_parameters, num_guard_cells)
    return _apply("refresh", field_tiles, static_parameters, bc_type)


def update_tiled_vector_ghost_cells(field_tiles, static_parameters, num_guard_cells=2, bc_type=BC_TYPE_FIELD):
    return tuple(update_tiled_ghost_cells(c, static_parameters, num_guard_cells, bc_type) for c in field_tiles)


def fold_tiled_ghost_cells(field_tiles, static_parameters, num_guard_cells=2, bc_type=BC_TYPE_FIELD):
    _check_g(static_parameters, num_guard_cells)
    return _apply("fold", field_tiles, static_parameters, bc_type)


def fold_tiled_vector_ghost_cells(field_tiles, static_parameters, num_guard_cells=2, bc_type=BC_TYPE_FIELD):
    return tuple(fold_tiled_ghost_cells(c, static_parameters, num_guard_cells, bc_type) for c in field_tiles)


def apply_tiled_zero_boundary(field_tiles, static_parameters, axis, num_guard_cells=2):
    _check_g(static_parameters, num_guard_cells)
    return _apply("zero", field_tiles, static_parameters, axis=axis)


def apply_tiled_constant_boundary(field_tiles, static_parameters, axis, num_guard_cells=2):
    _check_g(static_parameters, num_guard_cells)
    return _apply("constant", field_tiles, static_parameters, axis=axis)


def make_field_mesh(tile_grid_shape):
    """One tile per GPU: the only supported topology is (1, 1, 1)."""
    if tuple(int(w) for w in tile_grid_shape) != (1, 1, 1):
        raise ValueError(
            "Tiled field communication requires one logical tile per device: "
            f"tile topology {tuple(tile_grid_shape)} needs {int(torch.tensor(tile_grid_shape).prod())} devices, "
            "but this build runs one tile per GPU.")
    return None
