"""Host-side helpers of grid_and_stencil against the reference's known answers
(tests/code_tests/grid_and_stencil_test.py:35-107), CPU only."""

import numpy as np

from pypic3d_amd.boundary_conditions.grid_and_stencil import (axis_has_active_cells, build_collocated_axis,
                                                              build_staggered_axis, inactive_axis_index,
                                                              uniform_axis_spacing)


def test_axis_activity_and_inactive_index():
    assert axis_has_active_cells(6, ghost_cells=True)
    assert not axis_has_active_cells(3, ghost_cells=True)
    assert inactive_axis_index(3, ghost_cells=True) == 1
    assert inactive_axis_index(1, ghost_cells=False) == 0


def test_axis_builders_include_ghost_cells():
    col = build_collocated_axis(-1.0, 0.5, 4)
    stag = build_staggered_axis(-1.0, 0.5, 4)
    np.testing.assert_allclose(col, [-1.5, -1.0, -0.5, 0.0, 0.5, 1.0])
    np.testing.assert_allclose(stag, [-1.25, -0.75, -0.25, 0.25, 0.75, 1.25])
    assert uniform_axis_spacing(col) == 0.5
    assert uniform_axis_spacing([3.0]) == 1.0
