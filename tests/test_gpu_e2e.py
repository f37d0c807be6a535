"""End-to-end recovery through Solve(problem, {backend: hip}), with the
reference's e2e thresholds (tests/test_e2e.cpp:63-88), and full-size checks
that need no CPU factorization."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_known_answers.json")))["e2e_thresholds"]


def _check_recovered(oracle, prob, res, gt):
    m, p = prob.state()
    pose1 = p.reshape(-1, 3)[1]
    assert res["final_cost"] < GOLD["cost_fraction"] * res["initial_cost"]
    assert np.hypot(*(pose1[:2] - gt[1, :2])) < GOLD["translation"]
    assert abs(oracle.wrap(pose1[2] - gt[1, 2])) < GOLD["rotation"]


@pytest.mark.parametrize("variant", ["lm", "trust", "huber"])
def test_hip_backend_recovers_perturbed_pose(sdfslam, oracle, variant):
    print(f"case e2e {variant}", flush=True)
    prob = sdfslam.Problem.workload("corner", 1)
    _, _, _, gt = prob.workload_info()
    opts = sdfslam.default_solver_options(max_iterations=40)
    if variant == "trust":
        opts.trust_region = 1
    if variant == "huber":
        base = prob.arrays()
        m, p = prob.state()
        scans = []
        for f in range(2):
            sel = (base["point_frame"] == f) & (base["point_expected_sdf"] == 0.0)
            scans.append(np.stack([base["point_x"][sel], base["point_y"][sel]], 1))
        g = prob.view.grid
        po = sdfslam.default_problem_options(hallucination_points_per_scan_point=4,
                                             hallucination_step_size=0.2, normals_k_neighbors=8,
                                             normals_method=0, huber_delta=0.5)
        odo = np.array([[base["odom_dx"][0], base["odom_dy"][0], base["odom_dtheta"][0]]])
        prob = sdfslam.Problem.build((g.nx, g.ny, g.min_x, g.min_y, g.dx, g.dy), p.reshape(-1, 3),
                                     scans, odometry=odo, options=po)
    res = prob.solve(opts, backend="hip")
    _check_recovered(oracle, prob, res, gt)
    if variant == "trust":
        h = res["cost_history"]
        assert np.all(h[1:] <= h[:-1] + 1e-9)


def test_full_size_step_solves_the_damped_system(sdfslam):
    # 42,727 unknowns: the device step satisfies the device's own damped
    # normal equations (an independent check of factorization and solves).
    print("case full-size step residual", flush=True)
    prob = sdfslam.Problem.workload("office200", 1)
    ctx = sdfslam.Context(prob, sdfslam.default_solver_options())
    ctx.iterate(3)
    lam = 2.0
    cp, ri, va, _, rhs = ctx.normal_equations(lam)
    import scipy.sparse as sp
    n = prob.num_parameters
    L = sp.csc_matrix((va, ri, cp), shape=(n, n))
    A = L + sp.tril(L, -1).T
    x = ctx.linear_step(lam)
    assert np.linalg.norm(A @ x - rhs) <= 1e-10 * np.linalg.norm(rhs)
    assert np.array_equal(ctx.linear_step(lam), x)


def test_full_size_normal_equations_match_oracle(sdfslam, oracle):
    print("case full-size normal equations", flush=True)
    prob = sdfslam.Problem.workload("office200", 1)
    m, p = prob.state()
    rng = np.random.default_rng(3)
    m = m + rng.normal(0, 0.2, size=m.shape)
    ctx = sdfslam.Context(prob, sdfslam.default_solver_options(), m, p)
    lam = 5.0
    cp, ri, va, _, rhs = ctx.normal_equations(lam)
    import scipy.sparse as sp
    n = prob.num_parameters
    r_or, jac = oracle.assemble(prob.view, m, p)
    a_or = oracle.normal_matrix(jac, lam, False).to_scipy()
    dev = sp.csc_matrix((va, ri, cp), shape=(n, n))
    diff = (dev - sp.tril(a_or)).tocoo()
    d = a_or.diagonal()
    assert np.all(np.abs(diff.data) <= 1e-10 * np.sqrt(d[diff.row] * d[diff.col]))
    rhs_or = oracle.neg_atb(jac, r_or, n)
    assert np.all(np.abs(rhs - rhs_or) <= 1e-10 * np.sqrt(d) * np.linalg.norm(r_or))


@pytest.mark.parametrize("name", ["office_small", "office200"])
def test_reference_factor_counts_match_the_oracle_under_the_device_order(sdfslam, oracle, name):
    # The unpadded factor size the roofline is priced on: the oracle's
    # symbolic LDL^T of the reference pattern under the device's elimination
    # order gives the same nonzeros and the same sum of c^2 + 3c; the
    # device's own (padded) factor is at least that large.
    print(f"case reference factor counts {name}", flush=True)
    prob = sdfslam.Problem.workload(name, 1)
    m, p = prob.state()
    ctx = sdfslam.Context(prob, sdfslam.default_solver_options(), m, p)
    nnz, flops = ctx.reference_factor_counts()
    order = ctx.elimination_order()
    n = prob.num_parameters
    assert np.array_equal(np.sort(order), np.arange(n))
    r, jac = oracle.assemble(prob.view, m, p)
    a = oracle.normal_matrix(jac, 1.0, False)
    st, _ = oracle.ldlt_solve(a, oracle.neg_atb(jac, r, n), order)
    assert st == 0
    o_nnz, o_flops = oracle.last_factor_stats()
    assert (nnz, flops) == (o_nnz, o_flops)
    info = ctx.analysis()
    assert info["factor_flops"] >= flops
