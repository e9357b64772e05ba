"""Boundary codes, grid-axis builders and the periodic wrap
(PyPIC3D/boundary_conditions/grid_and_stencil.py).

The host-side pieces run as in the reference: the codes (:9-11), the axis
predicates used at set-up (:38-89) and the FP64 axis builders (:242-286).
``wrap_periodic_position`` (:14-35) runs on the GPU (``pic_wrap_periodic``).
The per-particle stencil primitives of the reference (anchor, offset, stencil
points, collapse of a reduced axis: :92-239) are device code inside the HIP
kernels (``anchor_of``, ``node_coord``, ``shape3`` and the axis structs of
pic_common.h and pic_particle_math.h), not separate calls.
"""

from __future__ import annotations

import numpy as np
import torch

from .. import _native as N

BC_PERIODIC = 0      # ghost cells mirror the opposite interior edge
BC_CONDUCTING = 1    # tangential E vanishes at the walls

PARTICLE_PERIODIC = 0
PARTICLE_REFLECTING = 1
PARTICLE_ABSORBING = 2


def wrap_periodic_position(x, wind):
    """mod(x + W/2, W) - W/2 for every element of a CUDA tensor, with a
    position at +W/2 kept at +W/2 (:14-35).  Returns a new tensor."""
    if not isinstance(x, torch.Tensor) or not x.is_cuda:
        raise ValueError("wrap_periodic_position takes a CUDA tensor")
    f64 = x.dtype == torch.float64
    out = x.to(torch.float64 if f64 else torch.float32).contiguous().clone()
    N.call(N.op("pic_wrap_periodic", f64), N.ptr(out), out.numel(), float(wind), N.stream_handle())
    return out


def axis_has_active_cells(axis_size, ghost_cells=False):
    """True when the axis holds more than one physical cell (:38-54)."""
    return int(axis_size) > (3 if ghost_cells else 1)


def inactive_axis_index(axis_size, ghost_cells=False):
    """Index of the single physical cell of a reduced axis (:57-73)."""
    return 1 if ghost_cells and int(axis_size) >= 3 else int(axis_size) // 2


def uniform_axis_spacing(axis):
    """axis[1] - axis[0], or 1.0 for a single-point axis (:76-89)."""
    axis = np.asarray(axis, dtype=np.float64).reshape(-1)
    return float(axis[1] - axis[0]) if axis.shape[0] > 1 else 1.0


def build_collocated_axis(minimum_physical, spacing, count):
    """Collocated line with one ghost node per side, count + 2 points from
    minimum - spacing to minimum + count * spacing (:242-265)."""
    return np.linspace(float(minimum_physical) - float(spacing),
                       float(minimum_physical) + int(count) * float(spacing), int(count) + 2)


def build_staggered_axis(minimum_physical, spacing, count):
    """Staggered (cell-centred) line, count + 2 points from
    minimum - spacing/2 to minimum + (count + 1/2) spacing (:268-286)."""
    return np.linspace(float(minimum_physical) - 0.5 * float(spacing),
                       float(minimum_physical) + (int(count) + 0.5) * float(spacing), int(count) + 2)
