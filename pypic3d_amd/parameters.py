"""Parameter containers with the reference's field names (PyPIC3D/parameters.py:7-53).

``StaticParameters`` selects branches and layout; ``DynamicParameters`` holds
scalars and the grid lines.  Values may be Python numbers, NumPy scalars or
CPU tensors.  ``field_mesh`` is kept for signature compatibility: this
library runs one tile per GPU, so it is ``None`` (or ignored).
"""

from __future__ import annotations

from typing import NamedTuple


class GridParameters(NamedTuple):
    vertex: tuple
    center: tuple
    tiled_vertex_grid: tuple
    tiled_center_grid: tuple


class StaticParameters(NamedTuple):
    name: str
    output_dir: str
    Nt: int
    verbose: bool
    GPUs: bool
    benchmark: bool
    solver: str
    electrostatic: bool
    relativistic: bool
    particle_pusher: str
    current_deposition: str
    current_filter: str
    shape_factor: int
    guard_cells: int
    tile_shape: tuple
    particle_tile_capacity_factor: float
    pml_active: bool
    boundary_conditions: tuple
    particle_boundary_conditions: tuple
    field_mesh: object


class DynamicParameters(NamedTuple):
    dt: object
    dx: object
    dy: object
    dz: object
    Nx: object
    Ny: object
    Nz: object
    x_wind: object
    y_wind: object
    z_wind: object
    C: object
    eps: object
    mu: object
    kb: object
    alpha: object
    grids: GridParameters


def _axis_tuple(values):
    if isinstance(values, (tuple, list)):
        return tuple(int(v) for v in values)
    return int(values["x"]), int(values["y"]), int(values["z"])


def _tile_shape(cfg):
    if "tile_shape" in cfg:
        return tuple(int(w) for w in cfg["tile_shape"])
    return int(cfg["particle_tile_nx"]), int(cfg["particle_tile_ny"]), int(cfg["particle_tile_nz"])


def build_static_parameters(static_config):
    """Same keys and defaults as parameters.py:89-123."""
    return StaticParameters(
        name=static_config.get("name", "Default Simulation"),
        output_dir=static_config.get("output_dir", "."),
        Nt=int(static_config.get("Nt", 0)),
        verbose=bool(static_config.get("verbose", False)),
        GPUs=bool(static_config.get("GPUs", False)),
        benchmark=bool(static_config.get("benchmark", False)),
        solver=static_config.get("solver", "electrodynamic_yee"),
        electrostatic=bool(static_config.get("electrostatic", False)),
        relativistic=bool(static_config.get("relativistic", True)),
        particle_pusher=static_config.get("particle_pusher", "boris"),
        current_deposition=static_config.get("current_deposition", "direct"),
        current_filter=static_config.get("current_filter", "none"),
        shape_factor=int(static_config["shape_factor"]),
        guard_cells=int(static_config["guard_cells"]),
        tile_shape=_tile_shape(static_config),
        particle_tile_capacity_factor=float(static_config.get("particle_tile_capacity_factor", 1.0)),
        pml_active=bool(static_config.get("pml_active", False)),
        boundary_conditions=_axis_tuple(static_config["boundary_conditions"]),
        particle_boundary_conditions=_axis_tuple(
            static_config.get("particle_boundary_conditions", {"x": 0, "y": 0, "z": 0})),
        field_mesh=static_config.get("field_mesh"),
    )


def build_dynamic_parameters(dynamic_config, extra_dynamic_config=None):
    """Same keys and defaults as parameters.py:126-161."""
    extra = extra_dynamic_config or {}
    grids = dynamic_config["grids"]
    if hasattr(grids, "_asdict"):
        grids = grids._asdict()
    grids = GridParameters(
        vertex=grids["vertex"], center=grids["center"],
        tiled_vertex_grid=grids["tiled_vertex_grid"], tiled_center_grid=grids["tiled_center_grid"],
    )
    get = lambda k, dflt: dynamic_config.get(k, extra.get(k, dflt))   # noqa: E731
    return DynamicParameters(
        dt=dynamic_config["dt"], dx=dynamic_config["dx"], dy=dynamic_config["dy"], dz=dynamic_config["dz"],
        Nx=dynamic_config["Nx"], Ny=dynamic_config["Ny"], Nz=dynamic_config["Nz"],
        x_wind=dynamic_config["x_wind"], y_wind=dynamic_config["y_wind"], z_wind=dynamic_config["z_wind"],
        C=get("C", 1.0), eps=get("eps", 1.0), mu=get("mu", 1.0), kb=get("kb", 1.0), alpha=get("alpha", 1.0),
        grids=grids,
    )
