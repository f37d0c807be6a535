"""Shared check: assembled point rows against the frozen-cell evaluation
(EvalPointResidualFrozenCell, residual_math.hpp:113-147) at each row's own
cell, read from the row's four node columns."""
import numpy as np


def seeded_problem(sdfslam, name, smooth, frames=10):
    """The first `frames` frames of a workload with the given gradient mode,
    and a map with gradient in every cell (the workload map plus noise)."""
    ds = sdfslam.Dataset(name, 1)
    fr = [ds.frame(k) for k in range(min(ds.num_frames, frames))]
    g = ds.grid
    opts = sdfslam.ProblemOptions.from_buffer_copy(ds.problem_options)
    opts.smooth_gradient = smooth
    poses = np.array([f[1] for f in fr])
    odo = np.array([f[2] for f in fr[1:]])
    prob = sdfslam.Problem.build((g.nx, g.ny, g.min_x, g.min_y, g.dx, g.dy), poses, [f[0] for f in fr],
                                 odometry=odo, options=opts)
    m, p = prob.state()
    m = m + np.random.default_rng(1).normal(0.0, 0.05, size=m.shape)
    return prob, g, m, p


def check_point_rows(oracle, prob, g, m, p, residuals, cols, vals, smooth):
    """residuals (rows,), cols/vals (rows, 7) as sdfslam_hip_assemble returns
    them. Node weights and residuals are compared in both gradient modes, pose
    partials in the exact mode; 1e-12 of the largest entry. Returns the
    number of rows checked."""
    a = prob.arrays()
    npts = len(a["point_x"])
    act = a["node_to_active"]
    num_active = int((act >= 0).sum())
    active_to_node = np.full(num_active, -1)
    active_to_node[act[act >= 0]] = np.nonzero(act >= 0)[0]
    rows, cw, ch = [], [], []
    for k in range(npts):
        node_cols = [c for c in cols[k] if 0 <= c < num_active]
        if len(node_cols) != 4:  # out of the domain, or a constant node in the cell
            continue
        base = active_to_node[node_cols].min()
        rows.append(k)
        cw.append(base % g.nx)
        ch.append(base // g.nx)
    rows = np.array(rows)
    fr, fdn, fdp = oracle.eval_points_frozen(g, m, p, a["point_frame"][rows], a["point_x"][rows],
                                             a["point_y"][rows], a["point_expected_sdf"][rows], cw, ch)
    sw = a["point_sqrt_weight"][rows]
    scale = max(1.0, np.abs(vals).max())
    assert np.abs(residuals[rows] - sw * fr).max() <= 1e-12 * max(1.0, np.abs(residuals).max())
    for j, k in enumerate(rows):
        b = cw[j] + ch[j] * g.nx
        node_ids = [b, b + 1, b + g.nx, b + g.nx + 1]
        for c, v in zip(cols[k], vals[k]):
            if 0 <= c < num_active:
                assert abs(v - sw[j] * fdn[j, node_ids.index(active_to_node[c])]) <= 1e-12 * scale
            elif c >= num_active and not smooth:
                assert abs(v - sw[j] * fdp[j, (c - num_active) % 3]) <= 1e-12 * scale
    return len(rows)


def oracle_rows(oracle, prob, m, p):
    """The oracle's assembly in the device's row-slot form (7 slots a row)."""
    r, jac = oracle.assemble(prob.view, m, p)
    j = jac.to_scipy().tocsr()
    cols = np.full((j.shape[0], 7), -1, dtype=np.int64)
    vals = np.zeros((j.shape[0], 7))
    for k in range(j.shape[0]):
        lo, hi = j.indptr[k], j.indptr[k + 1]
        cols[k, :hi - lo] = j.indices[lo:hi]
        vals[k, :hi - lo] = j.data[lo:hi]
    return r, cols, vals
