"""Map seeding on the device (sdfslam_hip_init_from_scans through the host
mirror) against the oracle's exhaustive search, bit for bit, and a whole
device run with map seeding and output files against the same run solved by
the oracle."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def grid_of(sdfslam, nx, ny, lo, hi):
    return sdfslam.Grid(nx, ny, lo[0], lo[1], (hi[0] - lo[0]) / (nx - 1), (hi[1] - lo[1]) / (ny - 1))


def frames_of(sdfslam, name, count):
    ds = sdfslam.Dataset(name, 1)
    scans, poses = [], []
    for k in range(min(count, ds.num_frames)):
        xy, pose, _ = ds.frame(k)
        scans.append(xy)
        poses.append(pose)
    return ds.grid, scans, np.array(poses)


@pytest.mark.parametrize("signed", [True, False])
def test_reference_cases_match_oracle(sdfslam, oracle, signed):
    print(f"case map_init reference signed={signed}", flush=True)
    wall = np.array([[2.0, 0.1 * i] for i in range(-20, 21)])
    for nx, scans, poses in [(11, [wall], [[0, 0, 0]]), (11, [wall], [[1.0, 0, 0]]),
                             (11, [wall, wall], [[0, 0, 0], [1.0, 0, 0]]),
                             (21, [wall, wall], [[0, 0, 0], [2.0, 0, 0]])]:
        g = grid_of(sdfslam, nx, nx, (0.0, -2.0), (4.0, 2.0))
        z = np.zeros(nx * nx)
        dev = sdfslam.init_from_scans(g, z, scans, poses, signed)
        ref = oracle.init_from_scans(g, z, scans, poses, signed)
        assert np.array_equal(dev, ref)


@pytest.mark.parametrize("name,count", [("room", 76), ("office150", 120)])
@pytest.mark.parametrize("signed", [True, False])
def test_workload_scans_match_oracle(sdfslam, oracle, name, count, signed):
    print(f"case map_init {name} {count} signed={signed}", flush=True)
    g, scans, poses = frames_of(sdfslam, name, count)
    z = np.full(g.nx * g.ny, 0.5)
    dev = sdfslam.init_from_scans(g, z, scans, poses, signed)
    ref = oracle.init_from_scans(g, z, scans, poses, signed)
    assert np.array_equal(dev, ref)
    assert (ref < 0).any() == signed


def test_empty_scans_leave_map_unchanged(sdfslam):
    print("case map_init empty", flush=True)
    g = grid_of(sdfslam, 5, 5, (0.0, 0.0), (1.0, 1.0))
    m = sdfslam.init_from_scans(g, np.full(25, 0.25), [], [], True)
    assert np.all(m == 0.25)
    m = sdfslam.init_from_scans(g, np.full(25, 0.25), (np.zeros((0, 2)), np.zeros((0, 2))), None, True,
                                device=False)
    assert np.all(m == 0.25)


@pytest.mark.parametrize("incremental", [False, True])
def test_seeded_run_matches_oracle_run(sdfslam, oracle, tmp_path, incremental):
    print(f"case run room_small incremental={incremental}", flush=True)

    def oracle_solve(view, opts, m, p, r):
        return oracle.lib.oracle_solve(view, opts, m, p, r)

    o = sdfslam.default_solver_options(max_iterations=6 if not incremental else 3, step_tolerance=0.0)
    kw = dict(incremental=incremental, max_frames=8, snapshot_every=3, init_from_scans=True,
              init_signed=True, options=o)
    dev_dir, ref_dir = str(tmp_path / "dev"), str(tmp_path / "ref")
    dev = sdfslam.run("room_small", 1, output_dir=dev_dir, **kw)
    ref = sdfslam.run("room_small", 1, output_dir=ref_dir, solve_fn=oracle_solve, **kw)
    assert dev["iterations"] == ref["iterations"]
    assert abs(dev["final_cost"] - ref["final_cost"]) <= 1e-9 * abs(ref["final_cost"])
    assert sorted(os.listdir(dev_dir)) == sorted(os.listdir(ref_dir))

    def table(d, f):
        rows = open(os.path.join(d, f)).read().splitlines()[1:]
        return np.array([[float(v) for v in r.replace(",", " ").split()] for r in rows])

    dp = table(dev_dir, "poses_estimated.csv") - table(ref_dir, "poses_estimated.csv")
    assert np.abs(dp[:, :2]).max() <= 1e-6 and np.abs(dp[:, 2]).max() <= 1e-7
    assert np.abs(table(dev_dir, "map.txt") - table(ref_dir, "map.txt")).max() <= 1e-6
    md, mr = (json.load(open(os.path.join(d, "metrics.json"))) for d in (dev_dir, ref_dir))
    assert abs(md["mean_translation_error"] - mr["mean_translation_error"]) <= 1e-6
