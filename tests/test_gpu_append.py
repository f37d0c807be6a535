"""Frames appended to a context's problem on the device
(sdfslam_hip_context_append_frames) against the problem rebuilt over all
frames on the host, as the reference's incremental runner rebuilds it before
each Solve (runner.cpp:245-258, problem.cpp:11-115): the device's point rows
equal the host builder's, and every increment's solve (cost history, map,
poses) is bitwise the same."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _compose(p, odo):
    c, s = np.cos(p[2]), np.sin(p[2])
    th = p[2] + odo[2]
    th = (th + np.pi) % (2 * np.pi) - np.pi
    return np.array([p[0] + c * odo[0] - s * odo[1], p[1] + s * odo[0] + c * odo[1], th])


def _options(S, ds, method, active):
    po = S.ProblemOptions()
    C.pointer(po)[0] = ds.problem_options
    po.normals_method = S.NORMAL_METHODS[method]
    po.active_region = 1 if active else 0
    po.active_margin = 2
    return po


# (workload, normal method, active region, frames built at once, appends:
# frames per append call before a solve; a tuple is several calls before one
# solve, the later ones superseding the earlier ones' deferred device work;
# a session's tenth frame is where appended poses' initial margins change).
CASES = [
    ("served3", "pca", False, 3, [1, 1, 1, 2, (1, 1), 1]),
    ("office_small", "hybrid", False, 4, [1, 2, 1, (1, 2)]),
    ("office_small", "pca", True, 4, [1, 1, 2, 1]),
]


@pytest.mark.parametrize("name,method,active,first,appends", CASES)
def test_appended_frames_match_rebuilt_problem(sdfslam, name, method, active, first, appends):
    S = sdfslam
    print(f"case append {name} {method} active={active}", flush=True)
    ds = S.Dataset(name, 2000 if name.startswith("served") else 1)
    po = _options(S, ds, method, active)
    so = S.default_solver_options(max_iterations=4, step_tolerance=0.0)
    g = ds.grid
    grid = (g.nx, g.ny, g.min_x, g.min_y, g.dx, g.dy)
    calls = [a if isinstance(a, tuple) else (a,) for a in appends]
    total = first + sum(sum(c) for c in calls)
    assert total <= ds.num_frames
    frames = [ds.frame(k) for k in range(total)]
    scans = [f[0] for f in frames]
    odo = np.array([f[2] for f in frames[1:]])
    poses = [frames[0][1]]
    for k in range(1, first):
        poses.append(_compose(poses[-1], odo[k - 1]))
    p0 = np.array(poses)
    prob = S.Problem.build(grid, p0, scans[:first], odo[:first - 1], options=po)
    m0, _ = prob.state()
    ctx_a = S.Context(prob, so, m0, p0.reshape(-1))
    ctx_b = S.Context(prob, so, m0, p0.reshape(-1))
    ra, rb = ctx_a.solve(so), ctx_b.solve(so)
    np.testing.assert_array_equal(ra["cost_history"], rb["cost_history"])
    m, p = ctx_a.state()
    fo = S.frame_options(po)
    have = first
    for call in calls:
        est = list(p.reshape(-1, 3))
        for part in call:
            for f in range(have, have + part):
                est.append(_compose(est[-1], odo[f - 1]))
            ctx_a.append_frames([scans[f] for f in range(have, have + part)], odo[have - 1:have + part - 1], fo,
                                m, np.array(est).reshape(-1))
            have += part
            # The same sequence of updates with the problem rebuilt (the
            # margins of appended poses depend on the frame count of the call).
            full = S.Problem.build(grid, np.array(est), scans[:have], odo[:have - 1], options=po, map_values=m)
            ctx_b.update(full, m, np.array(est).reshape(-1))
        # The device's rows are the host builder's, bit for bit.
        arr = full.arrays()
        f_, x_, y_, d_, w_ = ctx_a.point_rows()
        np.testing.assert_array_equal(f_, arr["point_frame"])
        np.testing.assert_array_equal(x_, arr["point_x"])
        np.testing.assert_array_equal(y_, arr["point_y"])
        np.testing.assert_array_equal(d_, arr["point_expected_sdf"])
        np.testing.assert_array_equal(w_, arr["point_sqrt_weight"])
        ia, ib = ctx_a.analysis(), ctx_b.analysis()
        for key in ("num_parameters", "num_residuals", "pattern_nnz_lower", "factor_nnz"):
            assert ia[key] == ib[key], key
        ra, rb = ctx_a.solve(so), ctx_b.solve(so)
        np.testing.assert_array_equal(ra["cost_history"], rb["cost_history"])
        m, p = ctx_a.state()
        mb, pb = ctx_b.state()
        np.testing.assert_array_equal(m, mb)
        np.testing.assert_array_equal(p, pb)
