"""Boundary-condition codes (PyPIC3D/boundary_conditions/grid_and_stencil.py:9-11).

The stencil helpers of the reference (anchors, offsets, stencil points,
collapse, periodic wrap) live on the device: see pic_common.h and
pic_particles.hip.
"""

BC_PERIODIC = 0      # ghost cells mirror the opposite interior edge
BC_CONDUCTING = 1    # tangential E vanishes at the walls

PARTICLE_PERIODIC = 0
PARTICLE_REFLECTING = 1
PARTICLE_ABSORBING = 2
