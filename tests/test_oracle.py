"""Pins the CPU oracle (test infrastructure) before it is trusted as the checker.

Each check is independent of the oracle's own formulas: the reference's
known answers (tests/golden/reference_known_answers.json), finite
differences, closed forms on fields whose derivatives are known exactly,
and dense linear algebra (numpy/scipy) on the same matrices.
"""
import json
import math
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_known_answers.json")))


def grid(sdfslam, nx, ny, mn, mx):
    return sdfslam.Grid(nx, ny, mn[0], mn[1], (mx[0] - mn[0]) / (nx - 1), (mx[1] - mn[1]) / (ny - 1))


# ---------------------------------------------------------------- pose


def test_pose_known_answers(oracle):
    g = GOLD["pose_apply"]
    p = g["pose"]
    c, s = math.cos(p[2]), math.sin(p[2])
    out = oracle.compose(p, [g["point"][0], g["point"][1], 0.0])
    assert abs(out[0] - g["expected"][0]) < g["tol"] and abs(out[1] - g["expected"][1]) < g["tol"]
    assert abs(c * g["point"][0] - s * g["point"][1] + p[0] - out[0]) < 1e-15
    r = GOLD["pose_roundtrip"]
    rel = oracle.relative(r["a"], r["b"])
    back = oracle.compose(r["a"], rel)
    assert np.allclose(back[:2], r["b"][:2], atol=r["tol"])
    assert abs(oracle.wrap(back[2] - r["b"][2])) < r["tol"]
    for a, expected, kind in GOLD["wrap_angle"]["cases"]:
        w = oracle.wrap(a)
        if kind == "abs":
            assert abs(abs(w) - expected) < 1e-12
        else:
            assert abs(w - expected) < 1e-12


# ---------------------------------------------------------------- grid


def map_3x3(sdfslam):
    g = GOLD["grid_map_3x3"]
    gr = grid(sdfslam, g["nx"], g["ny"], g["min"], g["max"])
    vals = np.array([w + 10.0 * h for h in range(3) for w in range(3)])
    return gr, vals


def test_grid_known_answers(sdfslam, oracle):
    gr, vals = map_3x3(sdfslam)
    g = GOLD["grid_map_3x3"]
    assert abs(oracle.interpolate(gr, vals, *g["query"]) - g["expected"]) < g["tol"]
    ok, w, h, a, b = oracle.locate(gr, *GOLD["grid_line_cell"]["query"])
    assert ok and (w, h) == tuple(GOLD["grid_line_cell"]["cell"]) and abs(a) < 1e-12 and abs(b) < 1e-12
    ok, w, h, a, b = oracle.locate(gr, *GOLD["grid_upper_boundary"]["query"])
    assert ok and (w, h) == (1, 1) and abs(a - 1) < 1e-12 and abs(b - 1) < 1e-12
    for q in GOLD["grid_outside"]["queries"]:
        assert not oracle.locate(gr, *q)[0]


def test_grid_gradient_matches_fd(sdfslam, oracle):
    gr, vals = map_3x3(sdfslam)
    q = GOLD["grid_gradient_fd"]["query"]
    eps = GOLD["grid_gradient_fd"]["eps"]
    ok, _, _, _, dp = oracle.eval_point(gr, vals, [0.0, 0.0, 0.0], q[0], q[1], 0.0)
    assert ok
    fx = (oracle.interpolate(gr, vals, q[0] + eps, q[1]) - oracle.interpolate(gr, vals, q[0] - eps, q[1])) / (2 * eps)
    fy = (oracle.interpolate(gr, vals, q[0], q[1] + eps) - oracle.interpolate(gr, vals, q[0], q[1] - eps)) / (2 * eps)
    assert abs(dp[0] - fx) < 1e-5 and abs(dp[1] - fy) < 1e-5


# ---------------------------------------------------------------- residual Jacobians


def random_map(sdfslam, rng):
    j = GOLD["jacobian_fd"]["map"]
    gr = grid(sdfslam, j["nx"], j["ny"], j["min"], j["max"])
    return gr, rng.uniform(*j["value_range"], size=j["nx"] * j["ny"])


@pytest.mark.parametrize("seed", [42, 7, 43])
def test_point_jacobian_matches_fd(sdfslam, oracle, seed):
    eps = GOLD["jacobian_fd"]["eps"]
    tol = GOLD["jacobian_fd"]["tol"]
    rng = np.random.default_rng(seed)
    checked = 0
    for _ in range(20):
        gr, vals = random_map(sdfslam, rng)
        pose = np.array([rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5), 3 * rng.uniform(-0.5, 0.5)])
        px, py = rng.uniform(-1, 1, size=2)
        delta = rng.uniform(-1, 1)
        ok, r, ids, dn, dp = oracle.eval_point(gr, vals, pose, px, py, delta)
        if not ok:
            continue
        for k in range(4):
            v = vals.copy()
            v[ids[k]] += eps
            rp = oracle.eval_point(gr, v, pose, px, py, delta)[1]
            v[ids[k]] -= 2 * eps
            rm = oracle.eval_point(gr, v, pose, px, py, delta)[1]
            assert abs(dn[k] - (rp - rm) / (2 * eps)) < tol
        for k in range(3):
            pp = pose.copy()
            pp[k] += eps
            pm = pose.copy()
            pm[k] -= eps
            ep = oracle.eval_point(gr, vals, pp, px, py, delta)
            em = oracle.eval_point(gr, vals, pm, px, py, delta)
            if not (ep[0] and em[0]) or not np.array_equal(ep[2], ids) or not np.array_equal(em[2], ids):
                continue
            assert abs(dp[k] - (ep[1] - em[1]) / (2 * eps)) < tol
            checked += 1
    assert checked > 10


def test_smooth_gradients_exact_on_affine_field(sdfslam, oracle):
    # Both smooth-gradient modes return the exact gradient of an affine field,
    # including one-sided node differences at the border and the clamped
    # extrapolation of the meter stencil (residual_math.hpp:40-73).
    gr = grid(sdfslam, 7, 6, [-3.0, -2.0], [3.0, 3.0])
    a, b, c = 0.7, -1.3, 0.25
    xs = gr.min_x + gr.dx * np.tile(np.arange(7), 6)
    ys = gr.min_y + gr.dy * np.repeat(np.arange(6), 7)
    vals = a * xs + b * ys + c
    rng = np.random.default_rng(5)
    for _ in range(50):
        th = rng.uniform(-3, 3)
        px, py = rng.uniform(-2.8, 2.8), rng.uniform(-1.8, 2.8)
        pose = [0.0, 0.0, th]
        cth, sth = math.cos(th), math.sin(th)
        gxp, gyp = cth * px - sth * py, sth * px + cth * py
        if not (-3 <= gxp <= 3 and -2 <= gyp <= 3):
            continue
        for smooth, step in ((1, 0.0), (1, 0.4)):
            ok, r, ids, dn, dp = oracle.eval_point(gr, vals, pose, px, py, 0.0, smooth, step)
            assert ok
            assert abs(dp[0] - a) < 1e-12 and abs(dp[1] - b) < 1e-12
            dxdt = -sth * px - cth * py
            dydt = cth * px - sth * py
            assert abs(dp[2] - (a * dxdt + b * dydt)) < 1e-11


def test_node_smooth_gradient_exact_on_quadratic_interior(sdfslam, oracle):
    # Central node differences are exact for x^2 + y^2 and bilinear blending
    # reproduces the linear node gradients, so interior cells give (2x, 2y).
    gr = grid(sdfslam, 9, 9, [0.0, 0.0], [8.0, 8.0])
    xs = np.tile(np.arange(9.0), 9)
    ys = np.repeat(np.arange(9.0), 9)
    vals = xs ** 2 + ys ** 2
    rng = np.random.default_rng(2)
    for _ in range(40):
        x, y = rng.uniform(1.0, 7.0, size=2)
        ok, _, _, _, dp = oracle.eval_point(gr, vals, [0.0, 0.0, 0.0], x, y, 0.0, 1, 0.0)
        assert ok
        assert abs(dp[0] - 2 * x) < 1e-11 and abs(dp[1] - 2 * y) < 1e-11


def test_eikonal_and_odometry_jacobians_match_fd(sdfslam, oracle):
    eps = GOLD["jacobian_fd"]["eps"]
    rng = np.random.default_rng(43)
    gr, vals = random_map(sdfslam, rng)
    w, h = GOLD["eikonal_fd"]["node"]
    nx, ny = GOLD["eikonal_fd"]["normal"]
    r, ids, d = oracle.eval_eikonal(gr, vals, w, h, nx, ny)
    for k in range(3):
        v = vals.copy()
        v[ids[k]] += eps
        rp = oracle.eval_eikonal(gr, v, w, h, nx, ny)[0]
        v[ids[k]] -= 2 * eps
        rm = oracle.eval_eikonal(gr, v, w, h, nx, ny)[0]
        assert abs(d[k] - (rp - rm) / (2 * eps)) < 1e-5
    o = GOLD["odometry_fd"]
    pi, pj, m = np.array(o["pose_i"]), np.array(o["pose_j"]), np.array(o["meas"])
    r, ji, jj = oracle.eval_odom(pi, pj, m)
    for k in range(3):
        e = np.zeros(3)
        e[k] = eps
        fdi = (oracle.eval_odom(pi + e, pj, m)[0] - oracle.eval_odom(pi - e, pj, m)[0]) / (2 * eps)
        fdj = (oracle.eval_odom(pi, pj + e, m)[0] - oracle.eval_odom(pi, pj - e, m)[0]) / (2 * eps)
        assert np.abs(ji[:, k] - fdi).max() < 1e-5
        assert np.abs(jj[:, k] - fdj).max() < 1e-5


# ---------------------------------------------------------------- assembly and normal equations


def exact_gradient_problem(sdfslam):
    """The reference's e2e corner setup with the exact bilinear gradient."""
    base = sdfslam.Problem.workload("corner", 1)
    a = base.arrays()
    opts = sdfslam.default_problem_options(hallucination_points_per_scan_point=4,
                                           hallucination_step_size=0.2, normals_k_neighbors=8,
                                           normals_method=0, smooth_gradient=0)
    m, p = base.state()
    poses = p.reshape(-1, 3)
    frames = a["point_frame"]
    scans = []
    for f in range(len(poses)):
        sel = (frames == f) & (a["point_expected_sdf"] == 0.0)
        scans.append(np.stack([a["point_x"][sel], a["point_y"][sel]], axis=1))
    g = base.view.grid
    return sdfslam.Problem.build((g.nx, g.ny, g.min_x, g.min_y, g.dx, g.dy), poses, scans,
                                 odometry=np.array([[a["odom_dx"][0], a["odom_dy"][0], a["odom_dtheta"][0]]]),
                                 options=opts)


def test_assembled_jacobian_is_the_derivative_of_the_residuals(sdfslam, oracle):
    prob = exact_gradient_problem(sdfslam)
    m, p = prob.state()
    rng = np.random.default_rng(9)
    m = m + rng.normal(0, 0.3, size=m.shape)  # a map with gradients
    r0, jac = oracle.assemble(prob.view, m, p)
    J = jac.to_scipy()
    n = prob.num_parameters
    na = prob.view.num_active_nodes
    for trial in range(3):
        d = rng.normal(0, 1, size=n) * 1e-7
        m1, p1 = m.copy(), p.copy()
        m1[: len(m1)] += d[:na]  # all nodes are active in this problem
        p1[3:] += d[na:]
        r1, _ = oracle.assemble(prob.view, m1, p1)
        pred = J @ d
        # Rows whose point stays in its cell are differentiable.
        ok = np.abs(r1 - r0 - pred) <= 1e-6 * np.abs(pred).max() + 1e-12
        assert ok.mean() > 0.98


def test_normal_matrix_and_rhs_match_dense_algebra(sdfslam, oracle):
    prob = sdfslam.Problem.workload("room_small", 2)
    m, p = prob.state()
    rng = np.random.default_rng(1)
    m = m + rng.normal(0, 0.2, size=m.shape)
    r, jac = oracle.assemble(prob.view, m, p)
    J = jac.to_scipy().toarray()
    n = prob.num_parameters
    A = oracle.normal_matrix(jac).to_scipy().toarray()
    dense = J.T @ J
    scale = np.sqrt(np.outer(np.diag(dense), np.diag(dense))) + 1e-300
    assert np.all(np.abs(A - dense) <= 1e-13 * scale)
    rhs = oracle.neg_atb(jac, r, n)
    assert np.allclose(rhs, -J.T @ r, rtol=1e-12, atol=1e-12 * np.abs(J.T @ r).max())
    lam = 0.3
    Ad = oracle.normal_matrix(jac, lam, False).to_scipy().toarray()
    assert np.allclose(np.diag(Ad), np.diag(dense) + lam, rtol=1e-14, atol=1e-14)
    Am = oracle.normal_matrix(jac, lam, True).to_scipy().toarray()
    dd = np.diag(dense)
    assert np.allclose(np.diag(Am), dd + lam * np.maximum(dd, 1e-12), rtol=1e-14, atol=1e-14)


def random_spd(n, width, seed):
    # The reference's RandomSpd recipe (tests/test_parallel_ldlt.cpp:32-54)
    # with numpy draws.
    rng = np.random.default_rng(seed)
    rows, cols, vals = [], [], []
    for i in range(n):
        rows.append(i); cols.append(i); vals.append(30.0 + abs(rng.uniform(-1, 1)))
        if i + 1 < n:
            v = rng.uniform(-1, 1)
            rows += [i, i + 1]; cols += [i + 1, i]; vals += [v, v]
        if i + width < n:
            v = rng.uniform(-1, 1)
            rows += [i, i + width]; cols += [i + width, i]; vals += [v, v]
    return sp.csc_matrix((vals, (rows, cols)), shape=(n, n))


@pytest.mark.parametrize("seed", GOLD["random_spd"]["seeds"])
def test_ldlt_solves_random_spd(oracle, seed):
    n = 3000 + 500 * seed
    A = random_spd(n, GOLD["random_spd"]["width"], seed)
    b = np.random.default_rng(seed + 100).uniform(-1, 1, size=n)
    mat = oracle.from_csc(A)
    for order in (None, np.arange(n)[::-1].astype(np.int32).copy()):
        st, x = oracle.ldlt_solve(mat, b, order)
        assert st == 0
        x_ref = spla.spsolve(A, b)
        assert np.linalg.norm(x - x_ref) <= 1e-12 * np.linalg.norm(x_ref)
        assert np.linalg.norm(A @ x - b) <= 1e-13 * np.linalg.norm(b)


def test_ldlt_reports_singular_matrix(oracle):
    g = GOLD["singular_diag"]
    d = np.full(g["n"], g["diag"])
    d[g["zero_at"]] = 0.0
    A = sp.diags(d).tocsc()
    st, _ = oracle.ldlt_solve(oracle.from_csc(A), np.ones(g["n"]))
    assert st == 2


# ---------------------------------------------------------------- LM control end to end


def _recovered(sdfslam, oracle, prob, options, thresholds):
    m0, p0 = prob.state()
    st, res, m, p = oracle.solve(sdfslam, prob.view, options, m0, p0)
    assert st == 0
    _, _, _, gt = prob.workload_info()
    pose1 = p.reshape(-1, 3)[1]
    init = p0.reshape(-1, 3)[1]
    assert res["final_cost"] < thresholds["cost_fraction"] * res["initial_cost"]
    err = np.hypot(*(pose1[:2] - gt[1, :2]))
    assert err < thresholds["translation"] and err < np.hypot(*(init[:2] - gt[1, :2]))
    assert abs(oracle.wrap(pose1[2] - gt[1, 2])) < thresholds["rotation"]
    return res, m, p


def test_oracle_lm_recovers_perturbed_pose(sdfslam, oracle):
    prob = sdfslam.Problem.workload("corner", 1)
    opts = sdfslam.default_solver_options(max_iterations=40)
    _recovered(sdfslam, oracle, prob, opts, GOLD["e2e_thresholds"])


def test_oracle_trust_region_is_monotone(sdfslam, oracle):
    prob = sdfslam.Problem.workload("corner", 1)
    opts = sdfslam.default_solver_options(max_iterations=40, trust_region=1)
    res, _, _ = _recovered(sdfslam, oracle, prob, opts, GOLD["e2e_thresholds"])
    h = res["cost_history"]
    assert np.all(h[1:] <= h[:-1] + 1e-9)


# ------------------------------------------------ frozen-cell point residual


@pytest.mark.parametrize("seed", [7, 11])
def test_frozen_cell_matches_regular_inside_cell(sdfslam, oracle, seed):
    # test_jacobians.cpp:78-102: inside its build-time cell the frozen-cell
    # evaluation equals the regular one to 1e-12.
    rng = np.random.default_rng(seed)
    checked = 0
    for _ in range(30):
        gr, vals = random_map(sdfslam, rng)
        pose = np.array([rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5), rng.uniform(-1.5, 1.5)])
        px, py, delta = rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(-1, 1)
        valid, r, ids, dn, dp = oracle.eval_point(gr, vals, pose, px, py, delta, smooth=0)
        if not valid:
            continue
        w, h = ids[0] % gr.nx, ids[0] // gr.nx
        fr, fdn, fdp = oracle.eval_points_frozen(gr, vals, pose, [0], [px], [py], [delta], [w], [h])
        assert abs(fr[0] - r) <= 1e-12
        assert np.abs(fdn[0] - dn).max() <= 1e-12 and np.abs(fdp[0] - dp).max() <= 1e-12
        checked += 1
    assert checked >= 10


def test_frozen_cell_pose_jacobian_matches_fd_off_cell(sdfslam, oracle):
    # A point that left its cell extrapolates the cell's bilinear patch, which
    # is smooth in the pose: central differences pin d_pose everywhere.
    rng = np.random.default_rng(3)
    eps = GOLD["jacobian_fd"]["eps"]
    tol = GOLD["jacobian_fd"]["tol"]
    for _ in range(10):
        gr, vals = random_map(sdfslam, rng)
        pose = np.array([rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5), rng.uniform(-1.5, 1.5)])
        px, py, delta = rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(-1, 1)
        w, h = int(rng.integers(0, gr.nx - 1)), int(rng.integers(0, gr.ny - 1))
        _, _, dp = oracle.eval_points_frozen(gr, vals, pose, [0], [px], [py], [delta], [w], [h])
        for k in range(3):
            e = np.zeros(3)
            e[k] = eps
            rp = oracle.eval_points_frozen(gr, vals, pose + e, [0], [px], [py], [delta], [w], [h])[0][0]
            rm = oracle.eval_points_frozen(gr, vals, pose - e, [0], [px], [py], [delta], [w], [h])[0][0]
            assert abs(dp[0, k] - (rp - rm) / (2 * eps)) <= tol


@pytest.mark.parametrize("name", ["room_small", "office_small"])
@pytest.mark.parametrize("smooth", [0, 1])
def test_oracle_point_rows_match_frozen_cell(sdfslam, oracle, name, smooth):
    # A second pin of the assembly: the oracle's point rows equal the
    # frozen-cell evaluation at each row's own cell.
    from frozen_check import check_point_rows, oracle_rows, seeded_problem
    prob, g, m, p = seeded_problem(sdfslam, name, smooth)
    r, cols, vals = oracle_rows(oracle, prob, m, p)
    assert check_point_rows(oracle, prob, g, m, p, r, cols, vals, smooth) > 0.5 * prob.view.num_points


# ------------------------------------------------ cost and LM control pins


@pytest.mark.parametrize("name", ["room_small", "office_small"])
def test_cost_is_the_squared_norm_of_the_assembled_residuals(sdfslam, oracle, name):
    # Problem::Cost (problem.cpp:132-156) with Huber off is r^T r of the
    # residual vector the assembly writes (lm_solver.cpp:30-82, 275-296):
    # the same rows, summed in another order.
    prob = sdfslam.Problem.workload(name, 2)
    m, p = prob.state()
    rng = np.random.default_rng(5)
    m = m + rng.normal(0, 0.1, size=m.shape)
    p = p + np.r_[np.zeros(3), rng.normal(0, 0.02, size=len(p) - 3)]
    assert prob.view.huber_delta == 0.0
    r, _ = oracle.assemble(prob.view, m, p)
    c = oracle.cost(prob.view, m, p)
    assert abs(c - float(np.dot(r, r))) <= 1e-12 * c


def huber_toy(sdfslam, delta):
    """Three point rows on the field D(x, y) = x (bilinear interpolation is
    exact on it), frame 0 at the origin: residuals 0.2, 0.5, 1.5 with unit
    weights; eikonal rows carry weight 0."""
    po = sdfslam.default_problem_options(hallucination_points_per_scan_point=0, weight_eikonal=0.0,
                                         normals_k_neighbors=2, huber_delta=delta, smooth_gradient=0)
    vals = np.array([float(w) for h in range(3) for w in range(3)])
    scan = np.array([[0.2, 0.7], [0.5, 1.3], [1.5, 0.4]])
    return sdfslam.Problem.build((3, 3, 0.0, 0.0, 1.0, 1.0), np.zeros((1, 3)), [scan], options=po,
                                 map_values=vals)


def test_huber_cost_and_irls_rows_closed_form(sdfslam, oracle):
    # huber_delta 0.4 (problem.hpp:115-121, problem.cpp:140-146):
    #   |0.2| <= 0.4: 0.2^2                = 0.04
    #   |0.5|  > 0.4: 0.4 * (2*0.5 - 0.4)  = 0.24
    #   |1.5|  > 0.4: 0.4 * (2*1.5 - 0.4)  = 1.04      cost 1.32
    # IRLS rows (lm_solver.cpp:39-47): w = delta/|r| outside, row scaled by
    # sqrt(w): 0.5 -> sqrt(0.2), 1.5 -> sqrt(0.6); node weights of (0.2, 0.7)
    # in cell (0, 0): (0.8*0.3, 0.2*0.3, 0.8*0.7, 0.2*0.7).
    prob = huber_toy(sdfslam, 0.4)
    m, p = prob.state()
    assert abs(oracle.cost(prob.view, m, p) - 1.32) <= 1e-14
    r, jac = oracle.assemble(prob.view, m, p)
    assert np.allclose(r[:3], [0.2, math.sqrt(0.2), math.sqrt(0.6)], rtol=0, atol=1e-15)
    J = jac.to_scipy().toarray()
    assert np.allclose(J[0, [0, 1, 3, 4]], [0.24, 0.06, 0.56, 0.14], rtol=0, atol=1e-15)
    # (0.5, 1.3) in cell (0, 1): weights (0.35, 0.35, 0.15, 0.15) times sqrt(0.8).
    assert np.allclose(J[1, [3, 4, 6, 7]], np.array([0.35, 0.35, 0.15, 0.15]) * math.sqrt(0.8),
                       rtol=0, atol=1e-15)
    # Huber off: the plain sum of squares 0.04 + 0.25 + 2.25.
    prob0 = huber_toy(sdfslam, 0.0)
    assert abs(oracle.cost(prob0.view, *prob0.state()) - 2.54) <= 1e-14


def _step_from(oracle, view, m, p, lam):
    st, step = oracle.linear_step(view, m, p, lam)
    assert st == 0
    return oracle.apply_step(view, step, m, p)


def test_fixed_factor_lambda_trace_with_reject(sdfslam, oracle):
    # room_small from its initial state with lambda_init 1e-2, factor 3 and
    # reject_worse_steps: iteration 1 lowers the cost, iterations 2-6 raise it
    # and are reverted, iteration 7 lowers it. By lm_solver.cpp:411-423 the
    # lambdas are 1e-2, then 1e-2/3 (success divides), then times 3 per
    # rejection; a rejection restores the backup bit for bit and repeats the
    # previous cost in the history.
    prob = sdfslam.Problem.workload("room_small", 1)
    v = prob.view
    m0, p0 = prob.state()
    lam0, f = 1e-2, 3.0

    def run(k, reject=1):
        o = sdfslam.default_solver_options(max_iterations=k, step_tolerance=0.0, lambda_init=lam0,
                                           lambda_factor=f, reject_worse_steps=reject)
        st, res, m, p = oracle.solve(sdfslam, v, o, m0, p0)
        assert st == 0 and res["iterations"] == k
        return res["cost_history"], m, p

    m1, p1 = _step_from(oracle, v, m0, p0, lam0)
    h, m, p = run(1)
    assert np.array_equal(m, m1) and np.array_equal(p, p1)
    c1 = oracle.cost(v, m1, p1)
    assert h[1] == c1 < h[0]
    lam = lam0 / f
    # Iteration 2's step (lambda0 / f) raises the cost; the solve reverts it.
    m2, p2 = _step_from(oracle, v, m1, p1, lam)
    assert oracle.cost(v, m2, p2) > c1
    h, m, p = run(2)
    assert np.array_equal(m, m1) and np.array_equal(p, p1) and h[2] == h[1]
    for _ in range(4):  # iterations 3-6, all rejected
        lam *= f
    h, m, p = run(6)
    assert np.array_equal(m, m1) and np.array_equal(p, p1) and np.all(h[2:] == h[1])
    lam *= f  # iteration 7 runs at lambda0 / f * f^5 and is accepted
    m7, p7 = _step_from(oracle, v, m1, p1, lam)
    h, m, p = run(7)
    assert np.array_equal(m, m7) and np.array_equal(p, p7)
    assert h[7] == oracle.cost(v, m7, p7) < c1
    # The lambdas a swapped update (multiply on success) would use instead.
    m7x, _ = _step_from(oracle, v, m1, p1, lam0 * f ** 6)
    assert not np.array_equal(m7x, m7)


def test_fixed_factor_lambda_trace_without_reject(sdfslam, oracle):
    # Same start without reject_worse_steps: the worse state of iteration 2
    # is kept and becomes the previous cost (lm_solver.cpp:416-421), lambda
    # goes 1e-2 -> 1e-2/3 -> (1e-2/3)*3 and iteration 3 starts from state 2.
    prob = sdfslam.Problem.workload("room_small", 1)
    v = prob.view
    m0, p0 = prob.state()
    lam0, f = 1e-2, 3.0
    o = sdfslam.default_solver_options(max_iterations=3, step_tolerance=0.0, lambda_init=lam0,
                                       lambda_factor=f, reject_worse_steps=0)
    st, res, m, p = oracle.solve(sdfslam, v, o, m0, p0)
    m1, p1 = _step_from(oracle, v, m0, p0, lam0)
    m2, p2 = _step_from(oracle, v, m1, p1, lam0 / f)
    m3, p3 = _step_from(oracle, v, m2, p2, lam0 / f * f)
    c = [oracle.cost(v, m0, p0), oracle.cost(v, m1, p1), oracle.cost(v, m2, p2), oracle.cost(v, m3, p3)]
    assert c[2] > c[1] and c[3] < c[2]
    assert np.array_equal(m, m3) and np.array_equal(p, p3)
    assert list(res["cost_history"]) == c


# ------------------------------------------------ jitter ensembles


def test_ensemble_summary_known_answer():
    import ensemble
    g = GOLD["stability_ensemble_summary"]
    s = ensemble.summarize(g["scores"], g["divergence_factor"])
    assert s["best"] == pytest.approx(g["best"]) and s["diverged"] == g["diverged"]
    assert s["n"] == g["n"] and s["spread"] == pytest.approx(g["spread"])
    j = GOLD["stability_ensemble_jitter"]
    assert ensemble.jittered_lambda(j["lambda_init"], 0, j["jitter"]) == j["lambda_init"]
    lam1 = ensemble.jittered_lambda(j["lambda_init"], 1, j["jitter"])
    assert lam1 == j["lambda_init"] * (1.0 + 1e-9) and lam1 != j["lambda_init"]


def test_mean_translation_error_closed_form():
    import ensemble
    gt = np.zeros((4, 3))
    est = np.array([[0, 0, 0], [3, 4, 0.1], [0, 1, 0], [1, 0, -2.0]])
    assert ensemble.mean_translation_error(est, gt) == pytest.approx((0 + 5 + 1 + 1) / 4)
    with pytest.raises(ValueError):
        ensemble.mean_translation_error(est[:1], gt[:1])


# ------------------------------------------------ ordering and threads


@pytest.mark.parametrize("name", ["room_small", "office_small"])
def test_amd_order_is_a_permutation_with_less_fill(sdfslam, oracle, name):
    # The approximate minimum degree order (the reference's default for
    # SimplicialLDLT) is a permutation and fills less than the geometric
    # order with the poses last; solving with either gives the same step.
    prob = sdfslam.Problem.workload(name, 1)
    m, p = prob.state()
    n = prob.num_parameters
    r, jac = oracle.assemble(prob.view, m, p)
    a = oracle.normal_matrix(jac, 0.5, False)
    amd = oracle.amd_order(a, n)
    assert np.array_equal(np.sort(amd), np.arange(n))
    rhs = oracle.neg_atb(jac, r, n)
    st, x_amd = oracle.ldlt_solve(a, rhs, amd)
    nnz_amd, flops_amd = oracle.last_factor_stats()
    st2, x_geo = oracle.ldlt_solve(a, rhs, oracle.order(prob.view, n))
    nnz_geo, flops_geo = oracle.last_factor_stats()
    assert st == st2 == 0
    assert nnz_amd < nnz_geo and flops_amd < flops_geo
    assert np.linalg.norm(x_amd - x_geo) <= 1e-10 * np.linalg.norm(x_geo)
    A = a.to_scipy()
    assert np.linalg.norm(A @ x_amd - rhs) <= 1e-12 * np.linalg.norm(rhs)


def test_amd_order_small_closed_forms(oracle):
    # An arrow matrix (vertex 0 joined to all others): minimum degree
    # eliminates the leaves first and the hub last, so L has no fill.
    n = 20
    rows = [0] * (n - 1) + list(range(1, n))
    cols = list(range(1, n)) + [0] * (n - 1)
    A = sp.csc_matrix((np.ones(2 * (n - 1)), (rows, cols)), shape=(n, n)) + sp.eye(n) * n
    mat = oracle.from_csc(A)
    o = oracle.amd_order(mat, n)
    assert o[-1] == 0 and sorted(o) == list(range(n))
    st, _ = oracle.ldlt_solve(mat, np.ones(n), o)
    assert st == 0 and oracle.last_factor_stats()[0] == n - 1


def test_threaded_oracle_is_bitwise_serial(sdfslam, oracle):
    # Assembly chunks, J^T J column ranges, -J^T r and the LDL^T over
    # elimination subtrees give the serial results bit for bit.
    prob = sdfslam.Problem.workload("office_small", 1)
    m0, p0 = prob.state()
    o = sdfslam.default_solver_options(max_iterations=3, step_tolerance=0.0)
    results = []
    for t in (1, 4, 7):
        oracle.set_threads(t)
        try:
            results.append(oracle.solve(sdfslam, prob.view, o, m0, p0))
            results[-1] += (oracle.linear_step(prob.view, m0, p0, 0.3)[1],)
        finally:
            oracle.set_threads(1)
    for st, res, m, p, x in results[1:]:
        assert st == 0
        assert np.array_equal(m, results[0][2]) and np.array_equal(p, results[0][3])
        assert np.array_equal(res["cost_history"], results[0][1]["cost_history"])
        assert np.array_equal(x, results[0][4])
