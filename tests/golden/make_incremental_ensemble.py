"""Writes the oracle's incremental runs of the office150 recipe that the GPU
tests compare with.

Runs in the build container (CPU only). Each increment is solved by the
oracle (the CPU restatement of the reference's SolveLm) through the host
library's incremental driver (runner.cpp:226-285 restated): frames are
appended one at a time, the problem is rebuilt over all frames so far and
refined with 5 LM iterations (lambda_init 5, factor 0.9, tolerance 1e-3).

  parity    the first PARITY_FRAMES frames, unjittered: per-increment
            iterations and final costs, the final poses and map
  ensemble  ENSEMBLE_RUNS runs over ENSEMBLE_FRAMES frames with lambda_init
            jittered by i * 1e-9 (tools/viz/stability_ensemble.py:33-40,
            63-64): each run's mean translation error against ground truth
            and its final poses

    python tests/golden/make_incremental_ensemble.py [processes]

About 15 minutes with 6 processes.
"""
import multiprocessing as mp
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "sdf-slam_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

WORKLOAD, SEED = "office150", 1
PARITY_FRAMES = 40
ENSEMBLE_FRAMES = 300
ENSEMBLE_RUNS = 8
JITTER = 1e-9
OUT = os.path.join(HERE, "incremental_office150.npz")


def _modules():
    import sdfslam as S
    import oracle_ctypes as O
    O.configure(S)
    return S, O


def _oracle_run(frames, lambda_init=None):
    S, O = _modules()

    def solve(view, opts, m, p, r):
        return O.lib.oracle_solve(view, opts, m, p, r)

    full = S.Problem.workload(WORKLOAD, SEED)
    o, _, per_increment, _ = full.workload_info()
    o.max_iterations = per_increment
    if lambda_init is not None:
        o.lambda_init = lambda_init
    t = time.time()
    out = S.run_incremental(WORKLOAD, SEED, options=o, max_frames=frames, solve_fn=solve)
    print(f"  {frames} frames, lambda_init {o.lambda_init!r}: {time.time() - t:.0f} s", flush=True)
    return out


def _ensemble_member(i):
    import ensemble
    S, _ = _modules()
    full = S.Problem.workload(WORKLOAD, SEED)
    o, _, _, gt = full.workload_info()
    lam = ensemble.jittered_lambda(o.lambda_init, i, JITTER)
    run = _oracle_run(ENSEMBLE_FRAMES, lam)
    err = ensemble.mean_translation_error(run["poses"], gt[:ENSEMBLE_FRAMES])
    return i, lam, err, run["poses"]


def main(processes):
    parity = _oracle_run(PARITY_FRAMES)
    with mp.get_context("spawn").Pool(processes) as pool:
        members = sorted(pool.map(_ensemble_member, range(ENSEMBLE_RUNS)))
    import ensemble
    errs = [m[2] for m in members]
    print("ensemble", ensemble.summarize(errs), flush=True)
    np.savez_compressed(
        OUT, workload=WORKLOAD, seed=SEED, jitter=JITTER,
        parity_frames=PARITY_FRAMES,
        parity_iterations=np.array([r["iterations"] for r in parity["records"]], dtype=np.int32),
        parity_final_cost=np.array([r["final_cost"] for r in parity["records"]]),
        parity_cost_history=parity["total"]["cost_history"],
        parity_poses=parity["poses"], parity_map=parity["map"],
        ensemble_frames=ENSEMBLE_FRAMES,
        ensemble_lambda_init=np.array([m[1] for m in members]),
        ensemble_error=np.array(errs),
        ensemble_poses=np.stack([m[3] for m in members]))
    print(f"wrote {OUT} ({os.path.getsize(OUT) // 1024} KiB)", flush=True)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 6)
