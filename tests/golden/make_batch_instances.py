"""Writes the oracle's 30-iteration trajectories of three instances of the
batched workload (64 uneven instances "batch<i>", seed 1000): the smallest
(batch24, 2,403 unknowns), a median one (batch7, 12,376) and the largest
(batch29, 25,699). Runs in the build container and calls only the oracle.

    python tests/golden/make_batch_instances.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "sdf-slam_amd"), os.path.join(ROOT, "tests"), HERE):
    if p not in sys.path:
        sys.path.insert(0, p)

import sdfslam as S  # noqa: E402
import oracle_ctypes as O  # noqa: E402
from make_batch_trajectories import fingerprint  # noqa: E402

SEED = 1000
INSTANCES = (24, 7, 29)
RECORD = (10, 30)
OUT = os.path.join(HERE, "batch_instances.npz")


def main():
    O.configure(S)
    out = {"seed": SEED, "instances": np.array(INSTANCES, dtype=np.int32),
           "record_iterations": np.array(RECORD, dtype=np.int32)}
    for i in INSTANCES:
        prob = S.Problem.workload(f"batch{i}", SEED)
        m0, p0 = prob.state()
        opts, _, _, _ = prob.workload_info()
        t = time.time()
        st, res, m, p, rm, rp = O.solve_recorded(S, prob.view, opts, m0, p0, RECORD)
        assert st == 0 and res["iterations"] == RECORD[-1]
        print(f"batch{i}: {prob.num_parameters} unknowns, {time.time() - t:.0f} s", flush=True)
        out[f"fingerprint_{i}"] = fingerprint(prob)
        out[f"cost_history_{i}"] = res["cost_history"]
        out[f"poses_{i}"] = rp
        out[f"map_{i}"] = rm[-1]
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({os.path.getsize(OUT) // 1024} KiB)", flush=True)


if __name__ == "__main__":
    main()
