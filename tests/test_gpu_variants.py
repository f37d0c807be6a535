"""Device path vs the oracle on the problem options and residual terms the
seeded workloads leave at their defaults: Huber IRLS, the meter-step smooth
gradient with clamped interpolation, the exact bilinear gradient, the active
region (constant nodes) and relations between frames far apart.

Tolerances as in test_gpu_parity.py: residuals and J 1e-12 of the largest
entry, J^T J 1e-10 of sqrt(A_ii A_jj) over the exact reference pattern,
-J^T r 1e-10 of sqrt(A_ii)*||r||, steps 1e-10, cost 1e-12, trajectories
1e-6 m / 1e-7 rad / map 1e-6.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from variants import VARIANTS, variant_problem

pytestmark = pytest.mark.gpu

CASES = [(name, v) for name in ("office_small", "room_small") for v in VARIANTS]


@pytest.fixture(scope="module")
def problems(sdfslam):
    cache = {}

    def get(name, variant):
        if (name, variant) not in cache:
            cache[(name, variant)] = variant_problem(sdfslam, name, 1, variant)
        return cache[(name, variant)]
    return get


def _state_after(sdfslam, oracle, prob, iterations):
    m, p = prob.state()
    o = sdfslam.default_solver_options(max_iterations=iterations, step_tolerance=0.0)
    st, _, m2, p2 = oracle.solve(sdfslam, prob.view, o, m, p)
    assert st == 0
    return m2, p2


def _rows_to_coo(cols, vals, n):
    rows = np.repeat(np.arange(cols.shape[0]), 7).reshape(cols.shape)
    mask = cols >= 0
    return sp.coo_matrix((vals[mask], (rows[mask], cols[mask])), shape=(cols.shape[0], n))


@pytest.mark.parametrize("name,variant", CASES)
def test_variant_assembly_and_cost_match_oracle(sdfslam, oracle, problems, name, variant):
    print(f"case variant assembly {name} {variant}", flush=True)
    prob = problems(name, variant)
    m, p = _state_after(sdfslam, oracle, prob, 2)
    ctx = sdfslam.Context(prob, sdfslam.default_solver_options(), m, p)
    r_dev, cols, vals = ctx.assemble()
    r_or, jac = oracle.assemble(prob.view, m, p)
    assert np.abs(r_dev - r_or).max() <= 1e-12 * max(1.0, np.abs(r_or).max())
    d = jac.to_scipy().tocoo()
    dd = _rows_to_coo(cols, vals, prob.num_parameters)
    assert dd.nnz == d.nnz
    ko, kd = np.lexsort((d.col, d.row)), np.lexsort((dd.col, dd.row))
    assert np.array_equal(d.row[ko], dd.row[kd]) and np.array_equal(d.col[ko], dd.col[kd])
    assert np.abs(d.data[ko] - dd.data[kd]).max() <= 1e-12 * max(1.0, np.abs(d.data).max())
    if variant == "active":
        # Constant nodes have no column: every node id of a point row that maps
        # to NodeColumn -1 is dropped, as problem.cpp:39-70 / lm_solver.cpp:48-52.
        assert prob.view.num_active_nodes < prob.view.grid.nx * prob.view.grid.ny
    c_or = oracle.cost(prob.view, m, p)
    assert abs(ctx.cost() - c_or) <= 1e-12 * c_or


@pytest.mark.parametrize("name,variant", CASES)
def test_variant_normal_equations_and_step_match_oracle(sdfslam, oracle, problems, name, variant):
    print(f"case variant normal/step {name} {variant}", flush=True)
    prob = problems(name, variant)
    m, p = _state_after(sdfslam, oracle, prob, 2)
    ctx = sdfslam.Context(prob, sdfslam.default_solver_options(), m, p)
    lam = 0.7
    cp, ri, va, structural, rhs = ctx.normal_equations(lam)
    n = prob.num_parameters
    r_or, jac = oracle.assemble(prob.view, m, p)
    a_or = oracle.normal_matrix(jac, lam, False).to_scipy()
    lower = sp.tril(a_or).tocoo()
    dev_cols = np.repeat(np.arange(n), np.diff(cp))
    key_dev = ri[structural].astype(np.int64) * n + dev_cols[structural]
    key_or = lower.row.astype(np.int64) * n + lower.col
    od, oo = np.argsort(key_dev), np.argsort(key_or)
    assert np.array_equal(key_dev[od], key_or[oo])
    assert np.all(va[~structural] == 0.0)
    diag = a_or.diagonal()
    bound = 1e-10 * np.sqrt(diag[lower.row[oo]] * diag[lower.col[oo]])
    assert np.all(np.abs(va[structural][od] - lower.data[oo]) <= bound)
    rhs_or = oracle.neg_atb(jac, r_or, n)
    assert np.all(np.abs(rhs - rhs_or) <= 1e-10 * np.sqrt(diag) * np.linalg.norm(r_or))
    for lam in (1.0, 0.05):
        x_dev = ctx.linear_step(lam)
        st, x_or = oracle.linear_step(prob.view, m, p, lam)
        assert st == 0
        assert np.linalg.norm(x_dev - x_or) <= 1e-10 * np.linalg.norm(x_or)


@pytest.mark.parametrize("name,variant", CASES)
def test_variant_trajectory_matches_oracle(sdfslam, oracle, problems, name, variant):
    print(f"case variant trajectory {name} {variant}", flush=True)
    prob = problems(name, variant)
    m0, p0 = prob.state()
    o = sdfslam.default_solver_options(max_iterations=8, step_tolerance=0.0)
    st, res_or, m_or, p_or = oracle.solve(sdfslam, prob.view, o, m0, p0)
    assert st == 0
    res, m_dev, p_dev = sdfslam.solve(prob, o, m0, p0)
    assert res["iterations"] == res_or["iterations"] == 8
    dp = (p_dev - p_or).reshape(-1, 3)
    assert np.abs(dp[:, :2]).max() <= 1e-6 and np.abs(dp[:, 2]).max() <= 1e-7
    assert np.abs(m_dev - m_or).max() <= 1e-6
    np.testing.assert_allclose(res["cost_history"], res_or["cost_history"], rtol=1e-9)
    if variant == "active":
        # Constant nodes keep their values bit for bit.
        const = np.asarray(prob.arrays()["node_to_active"]) < 0
        assert const.any() and np.array_equal(m_dev[const], m0[const])
