#!/usr/bin/env python3
"""Benchmark of the joint-LM hot path on one MI355X.

A step is one LM iteration of the 42,727-unknown batch polish workload
(office200: 50x50 m synthetic office, 910 poses, 180 beams over 180 deg,
200x200 SDF grid, 4 hallucination samples per beam), run by the device
backend from its warm start: assembly, fixed-pattern J^T J and -J^T r,
supernodal LDL^T, triangular solves, step, cost and LM control.

Prints one JSON line with the metric, a roofline entry for the point-row
assembly kernel (HIP events on the backend's stream) and a CPU baseline (the
oracle's restated serial CPU path on a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sdf-slam_amd"))

METRIC = "LM iters/s at 42.7k unknowns; increments/s + p50/p99 latency, 32 sessions"
HBM_PEAK_GBS = 8000.0


def point_assembly_bytes(problem, info):
    """Algorithmic HBM bytes of one launch of the point-row assembly kernel.

    Per point row: the spec (frame int32 + x, y, expected SDF, sqrt weight:
    36 B), its pose's footprint bitmap word (4 B), and the outputs (residual
    8 B, 4 node weights 32 B, 3 pose partials 24 B, cell key 4 B): 108 B.
    Once per launch: the map values and the poses.
    """
    v = problem.view
    per_row = 36 + 4 + 8 + 32 + 24 + 4
    once = 8 * v.grid.nx * v.grid.ny + 24 * v.num_poses
    return per_row * v.num_points + once


def cpu_baseline(problem, iterations):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import sdfslam as S
    import oracle_ctypes as O
    O.configure(S)
    m, p = problem.state()
    opts, _, _, _ = problem.workload_info()
    opts.max_iterations = iterations
    opts.step_tolerance = 0.0
    t0 = time.perf_counter()
    st, res, _, _ = O.solve(S, problem.view, opts, m, p)
    dt = time.perf_counter() - t0
    if st != 0:
        return None
    return {"value": res["iterations"] / dt, "unit": "LM iters/s", "cores": 1, "kind": "port",
            "sample": f"{res['iterations']} LM iterations of the same workload from its warm start "
                      f"(restated serial CPU path: assembly, J^T J, simplicial LDL^T with "
                      f"nested-dissection order, cost), {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="office200")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-iterations", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the hot path is measured "
              "at N=1", file=sys.stderr)
        sys.exit(2)

    import sdfslam as S
    t_build = time.perf_counter()
    problem = S.Problem.workload(args.workload, args.seed)
    opts, _, _, _ = problem.workload_info()
    opts.max_iterations = args.steps + args.warmup + 16
    opts.step_tolerance = 0.0
    ctx = S.Context(problem, opts)
    t_build = time.perf_counter() - t_build
    print(f"[bench] built {args.workload}: n={problem.num_parameters} rows={problem.num_residuals} "
          f"in {t_build:.2f} s", file=sys.stderr, flush=True)
    ctx.iterate(args.warmup)
    info0 = ctx.analysis()
    t0 = time.perf_counter()
    ctx.iterate(args.steps)
    elapsed = time.perf_counter() - t0
    info1 = ctx.analysis()
    reanalyses = info1["pattern_epochs"] - info0["pattern_epochs"]
    reanalysis_s = info1["analysis_seconds"] - info0["analysis_seconds"]
    ms_per_step = 1e3 * elapsed / args.steps
    print(f"[bench] timed {args.steps} iterations: {ms_per_step:.3f} ms/iter", file=sys.stderr,
          flush=True)
    ctx.iterate(min(args.steps, 10), phase_times=True)
    phases = ctx.phase_times()
    info = ctx.analysis()
    bytes_launch = point_assembly_bytes(problem, info)
    kernel_ms = phases["point_assembly_kernel_ms"]
    achieved = bytes_launch / (kernel_ms * 1e-3) / 1e9
    line = {
        "metric": METRIC,
        "value": args.steps / elapsed,
        "unit": "LM iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: batch polish, seed {args.seed}, "
                        f"{info['num_parameters']} unknowns, {problem.view.num_points} point rows",
            "unknowns": info["num_parameters"],
            "residual_rows": info["num_residuals"],
            "factor_nnz": info["factor_nnz"],
            "factor_gflop": info["factor_flops"] / 1e9,
            "supernodes": info["num_supernodes"],
            "pattern_epochs": info["pattern_epochs"],
            "reanalyses_in_timed_steps": reanalyses,
            "reanalysis_ms_in_timed_steps": round(1e3 * reanalysis_s, 2),
            "phase_ms": {k: round(v, 4) for k, v in phases.items()},
        },
        "roofline": {
            "kernel": "KAssemblePoints (point-row residual + Jacobian assembly)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "bytes_per_launch": bytes_launch,
            "avg_launch_ms": kernel_ms,
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(problem, args.cpu_iterations)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
