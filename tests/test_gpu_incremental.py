"""Incremental runs (runner.cpp:226-285) on the device against the same driver
solving each increment with the CPU oracle, and reuse of one device context
across problems of different sizes."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,frames", [("office_small", 12), ("room_small", 8)])
def test_incremental_run_matches_oracle(sdfslam, oracle, name, frames):
    print(f"case incremental {name} {frames}", flush=True)

    def oracle_solve(view, opts, m, p, r):
        return oracle.lib.oracle_solve(view, opts, m, p, r)

    o = sdfslam.default_solver_options(max_iterations=5, step_tolerance=0.0)
    ref = sdfslam.run_incremental(name, 1, options=o, max_frames=frames, solve_fn=oracle_solve)
    dev = sdfslam.run_incremental(name, 1, options=o, max_frames=frames)
    assert len(dev["records"]) == len(ref["records"]) == frames - 1
    for a, b in zip(dev["records"], ref["records"]):
        assert a["frames"] == b["frames"] and a["iterations"] == b["iterations"]
        assert abs(a["final_cost"] - b["final_cost"]) <= 1e-9 * b["final_cost"]
    dp = dev["poses"] - ref["poses"]
    assert np.abs(dp[:, :2]).max() <= 1e-6
    assert np.abs(dp[:, 2]).max() <= 1e-7
    assert np.abs(dev["map"] - ref["map"]).max() <= 1e-6
    np.testing.assert_allclose(dev["total"]["cost_history"], ref["total"]["cost_history"], rtol=1e-9)


def test_context_update_matches_fresh_context(sdfslam):
    # One context reused for a larger, then a smaller problem solves each
    # exactly as a context created for it.
    print("case context update", flush=True)
    o = sdfslam.default_solver_options(max_iterations=6, step_tolerance=0.0)
    small = sdfslam.Problem.workload("corner", 1)
    large = sdfslam.Problem.workload("office_small", 1)
    ctx = sdfslam.Context(small, o)
    first = ctx.solve()
    for prob in (large, small):
        ctx.update(prob)
        got = ctx.solve()
        m, p = ctx.state()
        fresh = sdfslam.Context(prob, o)
        want = fresh.solve()
        fm, fp = fresh.state()
        assert np.array_equal(got["cost_history"], want["cost_history"])
        assert np.array_equal(m, fm) and np.array_equal(p, fp)
    assert np.array_equal(got["cost_history"], first["cost_history"])
