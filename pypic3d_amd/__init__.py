"""pypic3d_amd -- MI355X (gfx950) HIP implementation of PyPIC3D's per-step PIC hot path.

Drop-in modules mirror the reference layout (PyPIC3D.<module> ->
pypic3d_amd.<module>): evolve, pusher.particle_push, deposition.{Esirkepov,
J_from_rhov, rho}, solvers.{first_order_yee, electrostatic_yee},
boundary_conditions.ghost_cells, particles.{particle_class,
particle_tile_communication}, utilities.{grids, filters}, utils, parameters.
All compute runs in libpypic3d_hip.so (C ABI: include/pypic3d_amd.h); there is
no CPU fallback.
"""

from .particles.particle_class import SpeciesConfig, TiledParticles  # noqa: F401
from .parameters import DynamicParameters, GridParameters, StaticParameters  # noqa: F401

__version__ = "0.1.0"
