"""Frozen-cell evaluation (EvalPointResidualFrozenCell, residual_math.hpp:113-147)
as an independent check of the device's point rows: every point row whose
cell has four active nodes equals the frozen-cell evaluation at that cell,
scaled by the row's sqrt weight, to 1e-12 of the largest entry (node
weights and residual in both gradient modes, pose partials in the exact
mode, which is the frozen cell's gradient)."""
import pytest

from frozen_check import check_point_rows, seeded_problem

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["room_small", "office_small"])
@pytest.mark.parametrize("smooth", [0, 1])
def test_device_point_rows_match_frozen_cell(sdfslam, oracle, name, smooth):
    print(f"case frozen_cell {name} smooth={smooth}", flush=True)
    prob, g, m, p = seeded_problem(sdfslam, name, smooth)
    ctx = sdfslam.Context(prob, sdfslam.default_solver_options(), m, p)
    r, cols, vals = ctx.assemble()
    assert check_point_rows(oracle, prob, g, m, p, r, cols, vals, smooth) > 0.5 * prob.view.num_points
