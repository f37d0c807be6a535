"""Writes the oracle's batch-LM trajectories that the GPU tests compare with.

Runs in the build container (CPU only) and calls only the oracle (the CPU
restatement of the reference's SolveLm, oracle/liboracle.so) on problems the
host library builds from the seeded synthetic workloads:

  room       10x10 m room, 3 box obstacles, 76 poses, 360 beams, 100x100
             grid (10,225 unknowns), 100 LM iterations from its initial state
  office200  50x50 m office floorplan, 910 poses, 180 beams, 200x200 grid
             (42,727 unknowns), 200 LM iterations from the warm start

For each it stores the cost history, the poses after each recorded
iteration count, the map after some of them, and a fingerprint of the
problem's specs so a test can tell that it rebuilt the same problem.

    python tests/golden/make_batch_trajectories.py room office200

The office200 run takes about an hour on one core (the oracle's LDL^T).
"""
import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "sdf-slam_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

import sdfslam as S  # noqa: E402
import oracle_ctypes as O  # noqa: E402

# workload: (iteration counts with recorded poses, counts with recorded maps)
PLANS = {
    "room": ([10, 25, 50, 100], [10, 25, 50, 100]),
    "office200": ([10, 25, 50, 100, 150, 200], [50, 200]),
}


def fingerprint(prob):
    """SHA-256 over the spec arrays and the initial state."""
    h = hashlib.sha256()
    a = prob.arrays()
    for k in sorted(a):
        h.update(k.encode())
        h.update(np.ascontiguousarray(a[k]).tobytes())
    m, p = prob.state()
    h.update(m.tobytes())
    h.update(p.tobytes())
    return h.hexdigest()


def path_of(name):
    return os.path.join(HERE, f"trajectory_{name}.npz")


def make(name, seed=1):
    O.configure(S)
    prob = S.Problem.workload(name, seed)
    m0, p0 = prob.state()
    opts, _, _, _ = prob.workload_info()
    pose_its, map_its = PLANS[name]
    its = sorted(set(pose_its) | set(map_its))
    t0 = time.time()
    st, res, m, p, rm, rp = O.solve_recorded(S, prob.view, opts, m0, p0, its)
    assert st == 0, st
    print(f"{name}: {res['iterations']} iterations in {time.time() - t0:.1f} s, "
          f"cost {res['initial_cost']:.9g} -> {res['final_cost']:.9g}", flush=True)
    reached = [k for k in its if k <= res["iterations"]]
    pose_keep = [i for i, k in enumerate(its) if k in pose_its and k in reached]
    map_keep = [i for i, k in enumerate(its) if k in map_its and k in reached]
    np.savez_compressed(
        path_of(name),
        workload=name, seed=seed, fingerprint=fingerprint(prob),
        unknowns=prob.num_parameters, options=np.array([opts.max_iterations, opts.lambda_init,
                                                         opts.lambda_factor, opts.step_tolerance]),
        iterations=res["iterations"], converged=res["converged"],
        cost_history=res["cost_history"],
        pose_iterations=np.array([its[i] for i in pose_keep], dtype=np.int32),
        poses=rp[pose_keep],
        map_iterations=np.array([its[i] for i in map_keep], dtype=np.int32),
        maps=rm[map_keep])
    print(f"wrote {path_of(name)} ({os.path.getsize(path_of(name)) // 1024} KiB)", flush=True)


if __name__ == "__main__":
    for n in sys.argv[1:] or list(PLANS):
        make(n)
