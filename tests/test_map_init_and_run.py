"""Map seeding from scans and whole runs with output files, on the host.

The oracle's InitializeFromScans (src/core/map_init.cpp:31-69) is pinned by
the reference's known answers (tests/test_map_init.cpp) and by a vectorised
numpy nearest-point search. RunWorkload (Run, src/pipeline/runner.cpp:196-303)
is driven with the oracle as its solve function, and its files are checked
against the reference's formats (runner.cpp:73-113, 172-192) and against the
states the same solves return.
"""
import json
import os

import numpy as np
import pytest


def grid_of(sdfslam, nx, ny, lo, hi):
    # GridMap(nx, ny, min, max): spacing (max - min) / (n - 1) (grid_map.hpp:18-19).
    return sdfslam.Grid(nx, ny, lo[0], lo[1], (hi[0] - lo[0]) / (nx - 1), (hi[1] - lo[1]) / (ny - 1))


def wall_at_x2():
    # test_map_init.cpp:12-18: a wall along x = 2 seen from the origin.
    return np.array([[2.0, 0.1 * i] for i in range(-20, 21)])


def test_oracle_map_init_reference_known_answers(sdfslam, oracle):
    g = grid_of(sdfslam, 11, 11, (0.0, -2.0), (4.0, 2.0))
    z = np.zeros(121)

    def val(m, w, h):
        return m[h * 11 + w]

    m = oracle.init_from_scans(g, z, [wall_at_x2()], [[0, 0, 0]], True)
    assert abs(val(m, 5, 5)) <= 1e-9  # on the surface (test_map_init.cpp:20-25)
    assert abs(val(m, 3, 5) - 0.8) <= 1e-9  # sensor side (27-32)
    assert abs(val(m, 7, 5) + 0.8) <= 1e-9  # behind (34-39)
    m = oracle.init_from_scans(g, z, [wall_at_x2()], [[1.0, 0, 0]], True)
    assert abs(val(m, 7, 5) - 0.2) <= 1e-9 and abs(val(m, 9, 5) + 0.6) <= 1e-9  # pose placement (41-47)
    m = oracle.init_from_scans(g, z, [wall_at_x2(), wall_at_x2()], [[0, 0, 0], [1.0, 0, 0]], True)
    assert abs(abs(val(m, 7, 5)) - 0.2) <= 1e-9  # nearest across scans (49-55)
    m = oracle.init_from_scans(g, z, [wall_at_x2()], [[0, 0, 0]], False)
    assert abs(val(m, 3, 5) - 0.8) <= 1e-9 and abs(val(m, 7, 5) - 0.8) <= 1e-9  # unsigned (57-62)
    # Unsigned has no sign seam (64-83).
    g21 = grid_of(sdfslam, 21, 21, (0.0, -2.0), (4.0, 2.0))
    scans, poses = [wall_at_x2(), wall_at_x2()], [[0, 0, 0], [2.0, 0, 0]]
    s = oracle.init_from_scans(g21, np.zeros(441), scans, poses, True).reshape(21, 21)
    u = oracle.init_from_scans(g21, np.zeros(441), scans, poses, False).reshape(21, 21)
    assert np.abs(np.diff(s, axis=1)).max() > np.abs(np.diff(u, axis=1)).max()
    # Empty scans leave the map untouched (85-89).
    g5 = grid_of(sdfslam, 5, 5, (0.0, 0.0), (1.0, 1.0))
    m = oracle.init_from_scans(g5, np.full(25, 0.25), [], [], True)
    assert np.all(m == 0.25)


@pytest.mark.parametrize("signed", [True, False])
def test_oracle_map_init_matches_numpy_search(sdfslam, oracle, signed):
    rng = np.random.default_rng(5)
    g = grid_of(sdfslam, 23, 17, (-3.0, -2.0), (4.0, 3.0))
    scans = [rng.uniform(-2.5, 2.5, size=(int(rng.integers(5, 40)), 2)) for _ in range(4)]
    poses = np.column_stack([rng.uniform(-1, 1, 4), rng.uniform(-1, 1, 4), rng.uniform(-3, 3, 4)])
    got = oracle.init_from_scans(g, np.zeros(23 * 17), scans, poses, signed)
    pts, org = [], []
    for sc, (x, y, t) in zip(scans, poses):
        c, s = np.cos(t), np.sin(t)
        pts.append(np.column_stack([c * sc[:, 0] - s * sc[:, 1] + x, s * sc[:, 0] + c * sc[:, 1] + y]))
        org.append(np.tile([x, y], (len(sc), 1)))
    pts, org = np.concatenate(pts), np.concatenate(org)
    w, h = np.meshgrid(np.arange(g.nx), np.arange(g.ny))
    q = np.column_stack([g.min_x + w.ravel() * g.dx, g.min_y + h.ravel() * g.dy])
    d2 = ((q[:, None, :] - pts[None, :, :]) ** 2).sum(axis=2)
    k = d2.argmin(axis=1)
    want = np.sqrt(d2[np.arange(len(q)), k])
    if signed:
        behind = np.linalg.norm(q - org[k], axis=1) > np.linalg.norm(pts[k] - org[k], axis=1)
        want = np.where(behind, -want, want)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)


def read_poses(path):
    lines = open(path).read().splitlines()
    assert lines[0] == "x,y,theta"
    return np.array([[float(v) for v in line.split(",")] for line in lines[1:]])


def read_map(path):
    lines = open(path).read().splitlines()
    head = lines[0].split()
    nx, ny = int(head[0]), int(head[1])
    vals = np.array([[float(v) for v in line.split(" ")] for line in lines[1:]])
    assert vals.shape == (ny, nx)
    return nx, ny, [float(v) for v in head[2:]], vals.ravel()


def test_batch_run_writes_reference_outputs(sdfslam, oracle, tmp_path):
    def oracle_solve(view, opts, m, p, r):
        return oracle.lib.oracle_solve(view, opts, m, p, r)

    o = sdfslam.default_solver_options(max_iterations=4, step_tolerance=0.0)
    out = str(tmp_path / "corner")
    total = sdfslam.run("corner", 1, output_dir=out, options=o, solve_fn=oracle_solve)
    for f in ["poses_odometry.csv", "poses_estimated.csv", "map.txt", "scan_points_global.csv",
              "poses_ground_truth.csv", "metrics.json"]:
        assert os.path.exists(os.path.join(out, f)), f
    # The estimate and map in the files are the solved state, to the last bit
    # (17 significant digits round-trip a double).
    prob = sdfslam.Problem.workload("corner", 1)
    m0, p0 = prob.state()
    st, res, m1, p1 = oracle.solve(sdfslam, prob.view, o, m0, p0)
    assert st == 0
    est = read_poses(os.path.join(out, "poses_estimated.csv"))
    assert np.array_equal(est.ravel(), np.asarray(p1).ravel())
    nx, ny, head, vals = read_map(os.path.join(out, "map.txt"))
    g = prob.view.grid
    assert (nx, ny) == (g.nx, g.ny) and head == [g.min_x, g.min_y, g.dx, g.dy]
    assert np.array_equal(vals, np.asarray(m1))
    metrics = json.load(open(os.path.join(out, "metrics.json")))
    assert metrics["iterations"] == res["iterations"] == total["iterations"]
    assert metrics["final_cost"] == res["final_cost"] and metrics["initial_cost"] == res["initial_cost"]
    gt = read_poses(os.path.join(out, "poses_ground_truth.csv"))
    wrap = lambda a: np.arctan2(np.sin(a), np.cos(a))
    assert np.isclose(metrics["mean_translation_error"],
                      np.mean(np.linalg.norm(est[:, :2] - gt[:, :2], axis=1)), rtol=1e-12)
    assert np.isclose(metrics["mean_rotation_error"], np.mean(np.abs(wrap(est[:, 2] - gt[:, 2]))),
                      rtol=1e-12, atol=1e-15)
    pts = open(os.path.join(out, "scan_points_global.csv")).read().splitlines()
    assert pts[0] == "x,y" and len(pts) > 1


def test_incremental_run_writes_snapshots(sdfslam, oracle, tmp_path):
    def oracle_solve(view, opts, m, p, r):
        return oracle.lib.oracle_solve(view, opts, m, p, r)

    o = sdfslam.default_solver_options(max_iterations=2, step_tolerance=0.0)
    out = str(tmp_path / "office")
    sdfslam.run("office_small", 1, output_dir=out, incremental=True, max_frames=6, snapshot_every=2,
                options=o, solve_fn=oracle_solve)
    snaps = sorted(os.listdir(os.path.join(out, "snapshots")))
    # Increments end at frames 2..6; every second one and the last (runner.cpp:271-284).
    assert snaps == ["map_000003.txt", "map_000005.txt", "map_000006.txt",
                     "poses_000003.csv", "poses_000005.csv", "poses_000006.csv"]
    ref = sdfslam.run_incremental("office_small", 1, options=o, max_frames=6, solve_fn=oracle_solve)
    est = read_poses(os.path.join(out, "poses_estimated.csv"))
    assert np.array_equal(est, ref["poses"])
    assert np.array_equal(read_poses(os.path.join(out, "snapshots", "poses_000006.csv")), ref["poses"])
    assert np.array_equal(read_map(os.path.join(out, "map.txt"))[3], ref["map"])
    assert len(read_poses(os.path.join(out, "poses_odometry.csv"))) == 6
