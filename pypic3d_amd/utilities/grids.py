"""Yee grid lines (PyPIC3D/utilities/grids.py:14-165, grid_and_stencil.py:242-286).

Host-side FP64 NumPy arrays.  "center" is the collocated line with one ghost
node per side starting at -W/2 - d, "vertex" the staggered line starting at
-W/2 - d/2; the electrostatic solver uses collocated lines for both.  Tile
lines (one tile covering the domain) place local index l at
``global[0] + (l - (g-1)) * d``, so local l holds physical index l - g.
"""

from __future__ import annotations

import numpy as np

from ..boundary_conditions.grid_and_stencil import build_collocated_axis, build_staggered_axis


def _axis(minimum, d, n, half):
    return build_staggered_axis(minimum, d, n) if half else build_collocated_axis(minimum, d, n)


def build_collocated_grid(dynamic_parameters):
    """Collocated center and vertex lines (grids.py:14-39)."""
    dp = dynamic_parameters
    grid = tuple(_axis(-float(w) / 2, float(d), int(n), False)
                 for w, d, n in ((dp.x_wind, dp.dx, dp.Nx), (dp.y_wind, dp.dy, dp.Ny), (dp.z_wind, dp.dz, dp.Nz)))
    return grid, grid


def build_yee_grid(dynamic_parameters):
    """(center, vertex) lines of the electrodynamic solver (grids.py:42-77)."""
    dp = dynamic_parameters
    spec = ((dp.x_wind, dp.dx, dp.Nx), (dp.y_wind, dp.dy, dp.Ny), (dp.z_wind, dp.dz, dp.Nz))
    center = tuple(_axis(-float(w) / 2, float(d), int(n), False) for w, d, n in spec)
    vertex = tuple(_axis(-float(w) / 2, float(d), int(n), True) for w, d, n in spec)
    return center, vertex


def _tile_line(global_line, d, n, g):
    offsets = np.arange(n + 2 * g, dtype=np.float64)
    return (global_line[0] + (offsets + 0.0 * n - (g - 1)) * d).reshape(1, 1, 1, n + 2 * g)


def build_tiled_yee_grids(static_parameters, dynamic_parameters):
    """Tile lines of the single tile, shaped (1, 1, 1, n+2g) (grids.py:80-165)."""
    g = int(static_parameters.guard_cells)
    dp = dynamic_parameters
    ds = (float(dp.dx), float(dp.dy), float(dp.dz))
    ns = (int(dp.Nx), int(dp.Ny), int(dp.Nz))
    center = tuple(_tile_line(dp.grids.center[a], ds[a], ns[a], g) for a in range(3))
    vertex = tuple(_tile_line(dp.grids.vertex[a], ds[a], ns[a], g) for a in range(3))
    return center, vertex


def tiled_origin_spacing(wind, d, n, g, kind, electrostatic=False):
    """(line[0], line[1] - line[0]) of a tile line, computed like the reference."""
    half = kind == "vertex" and not electrostatic
    line = _tile_line(_axis(-float(wind) / 2, float(d), int(n), half), float(d), int(n), int(g)).reshape(-1)
    return float(line[0]), float(line[1] - line[0])
