#!/usr/bin/env python3
"""Benchmark of the joint-LM hot path on one MI355X.

A step is one LM iteration of the 42,727-unknown batch polish workload
(office200: 50x50 m synthetic office, 910 poses, 180 beams over 180 deg,
200x200 SDF grid, 4 hallucination samples per beam), run by the device
backend from its warm start: assembly, fixed-pattern J^T J and -J^T r,
supernodal LDL^T, triangular solves, step, cost and LM control.

Prints one JSON line with the metric, roofline entries for the dominant
kernel family (the trailing updates and update matrices of the supernodal
LDL^T on FP64 MFMA), for the 64-column panel step (FP64) and for the
point-row assembly kernel (HBM), all timed with HIP events on the backend's
stream (HBM traffic per launch from the committed counter summary), and a
CPU baseline (the oracle's restated CPU path, threaded as the reference's,
on a bounded sample of the same workload: all cores and one).
Three more legs report the other forms of the metric: an incremental run
(increments/s and per-increment latency), the served form (32 sessions with
Poisson 10 Hz frame arrivals: increments/s, p50/p99 latency, against the
plain one-session-at-a-time path and one session alone) and the batched form
(64 uneven instances: aggregate LM iterations/s).
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
FP64_PEAK_TFLOPS = 78.6


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


def calibration_copy(mib):
    """One device-to-device copy of `mib` MiB: under a counter run, its
    FETCH_SIZE / WRITE_SIZE readings give the scale from counter units to
    bytes (it reads and writes exactly that many bytes)."""
    import torch
    n = mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (profiles/office200_hbm_traffic.txt: lines `kernel fetch_bytes
    write_bytes`), or None when it has no line for the kernel."""
    path = os.path.join(ROOT, "profiles", "office200_hbm_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if line.startswith("#"):
            continue
        if len(parts) == 3 and parts[0] == kernel:
            return float(parts[1]) + float(parts[2])
    return None


def profile_avg_launch_us(kernels, counted=None):
    """Average kernel duration (us) of the kernels' launches in an LM iteration
    from the committed rocprofv3 summary (profiles/office200_kernel_stats.md,
    per-iteration table), or None when a kernel has no row: their time per
    iteration over the launches of `counted` (default: all of them; a step
    that is two kernels counts once). The HIP-event figures of this run
    bracket each launch and so also hold its dispatch."""
    path = os.path.join(ROOT, "profiles", "office200_kernel_stats.md")
    if not os.path.exists(path):
        return None
    text = open(path).read()
    if "## Per LM iteration" not in text:
        return None
    rows = {}
    for line in text.split("## Per LM iteration", 1)[1].splitlines():
        parts = [x.strip() for x in line.strip().strip("|").split("|")]
        if len(parts) == 5 and parts[0] in kernels:
            rows[parts[0]] = (float(parts[1]), float(parts[2]))
    if set(rows) != set(kernels):
        return None
    launches = sum(r[0] for k, r in rows.items() if counted is None or k in counted)
    return sum(r[1] for r in rows.values()) / launches if launches > 0 else None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(problem, timed):
    """The oracle's restated CPU path on the same workload: one LM iteration
    (threaded assembly, J^T J and -J^T r, approximate minimum degree order,
    simplicial LDL^T over elimination subtrees, step, cost) from the warm
    start, timed `timed` times after one warm-up, on all of this process's
    cores (at most 16) and on one; the median of each."""
    import statistics
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import sdfslam as S
    import oracle_ctypes as O
    O.configure(S)
    m, p = problem.state()
    opts, _, _, _ = problem.workload_info()
    opts.max_iterations = 1
    opts.step_tolerance = 0.0
    cores = min(16, len(os.sched_getaffinity(0)))

    def median_seconds(threads):
        O.set_threads(threads)
        times = []
        for k in range(1 + timed):
            t0 = time.perf_counter()
            st, res, _, _ = O.solve(S, problem.view, opts, m, p)
            dt = time.perf_counter() - t0
            if st != 0:
                return None, None
            if k > 0:
                times.append(dt)
        return statistics.median(times), [round(x, 3) for x in O.last_phase_seconds()]

    all_s, all_phases = median_seconds(cores)
    one_s, one_phases = median_seconds(1)
    O.set_threads(1)
    if all_s is None or one_s is None:
        return None
    lnz, flops = O.last_factor_stats()
    return {"value": 1.0 / all_s, "unit": "LM iters/s", "cores": cores, "kind": "port",
            "value_1_core": 1.0 / one_s, "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "factor_nnz_amd": lnz, "factor_gflop_amd": flops / 1e9,
            "phase_seconds": {"all_cores": all_phases, "one_core": one_phases,
                              "order": ["assembly", "normal equations + ordering", "LDL^T + solve",
                                        "step + cost"]},
            "sample": f"one LM iteration of the same workload from its warm start, median of {timed} "
                      f"after a warm-up, on {cores} threads and on 1 (restated reference CPU path: "
                      f"threaded assembly, J^T J, approximate minimum degree order, simplicial LDL^T "
                      f"over elimination subtrees, cost; not CHOLMOD, which is not in this image)"}


def incremental_leg(S, frames):
    """Incremental run (office150's first `frames` frames, 5 iterations per
    increment): increments/s and per-increment latency (host Problem
    rebuild + device solve)."""
    import numpy as np
    t0 = time.perf_counter()
    out = S.run_incremental("office150", 1, max_frames=frames)
    wall = time.perf_counter() - t0
    rec = out["records"]
    lat = np.array([r["build_seconds"] + r["solve_seconds"] for r in rec]) * 1e3
    return {"workload": f"office150 incremental, frames 1..{frames}, 5 LM iterations/increment",
            "increments": len(rec), "increments_per_s": len(rec) / wall,
            "latency_p50_ms": float(np.percentile(lat, 50)),
            "latency_p99_ms": float(np.percentile(lat, 99)),
            "build_mean_ms": float(np.mean([r["build_seconds"] for r in rec]) * 1e3),
            "solve_mean_ms": float(np.mean([r["solve_seconds"] for r in rec]) * 1e3),
            "lm_iterations": out["total"]["iterations"]}


def served_leg(S, duration, plain_duration, workers):
    """32 incremental sessions, Poisson frame arrivals at 10 Hz each; the
    served form (continuous batching: `workers` host threads prepare each
    increment on arrival, three solver threads batch what is prepared) against
    the plain path (one increment at a time, preparation and solve in turn)
    and one session alone; then the served form at 30 Hz per session for the
    throughput it sustains, and the 10 Hz load (and one session alone) on
    sessions advanced to frame 150 first (steady state)."""
    full = S.served_benchmark(sessions=32, rate_hz=10.0, duration_s=duration, workers=workers)
    plain = S.served_benchmark(sessions=32, rate_hz=10.0, duration_s=plain_duration, workers=1,
                               max_batch=1, pipelines=1, max_in_flight=1)
    alone = S.served_benchmark(sessions=1, rate_hz=10.0, duration_s=duration, workers=1,
                               max_batch=1, pipelines=1, max_in_flight=1)
    # Saturating load: 32 sessions at 30 Hz each (960 frames/s offered), the
    # increments/s the served form sustains.
    sat = S.served_benchmark(sessions=32, rate_hz=30.0, duration_s=duration, workers=workers)
    # Steady state: the same load on sessions already holding 150 frames
    # (each increment rebuilds a problem over every frame so far).
    steady = S.served_benchmark(sessions=32, rate_hz=10.0, duration_s=duration, workers=workers,
                                start_frame=150)
    steady_alone = S.served_benchmark(sessions=1, rate_hz=10.0, duration_s=duration, workers=1,
                                      max_batch=1, pipelines=1, max_in_flight=1, start_frame=150)
    r = lambda d: {k: (round(v, 3) if isinstance(v, float) else v) for k, v in d.items()}
    return {"sessions": 32, "rate_hz_per_session": 10.0, "duration_s": duration, "workers": workers,
            "increments_per_s": full["increments_per_second"],
            "latency_p50_ms": full["latency_p50_ms"], "latency_p99_ms": full["latency_p99_ms"],
            "throughput_vs_plain": full["increments_per_second"] / plain["increments_per_second"],
            "p99_vs_single_session": full["latency_p99_ms"] / alone["latency_p99_ms"],
            "saturated_increments_per_s": sat["increments_per_second"],
            "saturated_vs_plain": sat["increments_per_second"] / plain["increments_per_second"],
            "served": r(full), "plain_one_at_a_time": r(plain), "single_session": r(alone),
            "saturated_30hz": r(sat),
            "steady_state_from_frame_150": {
                "increments_per_s": steady["increments_per_second"],
                "latency_p50_ms": steady["latency_p50_ms"], "latency_p99_ms": steady["latency_p99_ms"],
                "p99_vs_single_session": steady["latency_p99_ms"] / steady_alone["latency_p99_ms"],
                "served": r(steady), "single_session": r(steady_alone)}}


def batched_leg(S):
    """64 uneven batch instances, 30 LM iterations each: aggregate LM
    iterations/s with batched solves of similar-sized instances on concurrent
    streams, the shape (instances per batched solve, solves in flight) picked
    by the autotuner from trials of the same iteration count (shorter trials
    over-weight the contexts' setup), against one solve at a time."""
    best, trials = S.batch_autotune(instances=64, max_batches=(4, 8, 16, 64), workers=(1, 4, 8, 12),
                                    trial_iterations=30)
    max_batch, workers = best[0], best[1]
    full = S.batch_benchmark(instances=64, iterations=30, max_batch=max_batch, workers=workers)
    # Every stage of an iteration one launch per level over all 64 instances.
    one = S.batch_benchmark(instances=64, iterations=30, max_batch=64, workers=1)
    serial = S.batch_benchmark(instances=64, iterations=30, max_batch=1, workers=1)
    return {"instances": 64, "iterations": 30, "max_batch": max_batch, "workers": workers,
            "lm_iters_per_s": full["lm_iterations_per_second"],
            "one_batch_of_64_lm_iters_per_s": one["lm_iterations_per_second"],
            "serial_lm_iters_per_s": serial["lm_iterations_per_second"],
            "speedup_vs_serial": full["lm_iterations_per_second"] / serial["lm_iterations_per_second"],
            "unknowns_total": full["unknowns"],
            "autotune_trials": [{"max_batch": t[0], "workers": t[1], "lm_iters_per_s": round(t[2], 1)}
                                for t in trials]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="office200")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-iterations", type=int, default=3,
                    help="timed oracle iterations per thread count (after one warm-up)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--served-duration", type=float, default=5.0)
    ap.add_argument("--plain-duration", type=float, default=2.0)
    ap.add_argument("--workers", type=int, default=12)
    ap.add_argument("--incremental-frames", type=int, default=150)
    ap.add_argument("--profile-run", action="store_true",
                    help="for kernel profiles: the iterations run as direct launches with per-launch "
                         "event timing only (no graph replays, no other forms, no CPU baseline)")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the incremental, served and batched legs")
    ap.add_argument("--calibration-copy-mb", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (the reference "
                         "kernel for scaling HBM counter readings)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the hot path is measured "
              "at N=1", file=sys.stderr)
        sys.exit(2)

    import sdfslam as S
    if args.calibration_copy_mb > 0:
        calibration_copy(args.calibration_copy_mb)
    t_build = time.perf_counter()
    problem = S.Problem.workload(args.workload, args.seed)
    opts, _, _, _ = problem.workload_info()
    opts.max_iterations = args.steps + args.warmup + 16
    opts.step_tolerance = 0.0
    ctx = S.Context(problem, opts)
    t_build = time.perf_counter() - t_build
    print(f"[bench] built {args.workload}: n={problem.num_parameters} rows={problem.num_residuals} "
          f"in {t_build:.2f} s", file=sys.stderr, flush=True)
    if args.profile_run:
        args.no_forms = args.no_cpu_baseline = True
    ctx.iterate(args.warmup, phase_times=args.profile_run)
    info0 = ctx.analysis()
    t0 = time.perf_counter()
    ctx.iterate(args.steps, phase_times=args.profile_run)
    elapsed = time.perf_counter() - t0
    info1 = ctx.analysis()
    reanalyses = info1["pattern_epochs"] - info0["pattern_epochs"]
    speculative = info1["speculative_epochs"] - info0["speculative_epochs"]
    reanalysis_s = info1["analysis_seconds"] - info0["analysis_seconds"]
    ms_per_step = 1e3 * elapsed / args.steps
    print(f"[bench] timed {args.steps} iterations: {ms_per_step:.3f} ms/iter", file=sys.stderr,
          flush=True)
    ctx.iterate(min(args.steps, 10), phase_times=True)
    phases = ctx.phase_times()
    info = ctx.analysis()
    # Flops of the factorization without this backend's padding (explicit
    # zeros of the superset pattern, dense supernode blocks): the reference
    # pattern's symbolic LDL^T under the same order. The roofline entries of
    # the factorization are priced on these useful flops.
    ref_nnz, ref_flops = ctx.reference_factor_counts()
    useful = ref_flops / info["factor_flops"] if info["factor_flops"] > 0 else 1.0
    bytes_launch = point_assembly_bytes(problem, info)
    kernel_ms = phases["point_assembly_kernel_ms"]
    achieved = bytes_launch / (kernel_ms * 1e-3) / 1e9
    # Dominant kernel family: the trailing updates (fused with the next
    # step's diagonal block and panel rows) and update matrices of the
    # supernodal LDL^T (KGemmAhead / KGemmLower / KGemmUpdate, FP64 MFMA).
    gemm_flops_padded = info["gemm_flops"] / max(info["gemm_launches"], 1)
    gemm_flops = gemm_flops_padded * useful
    gemm_ms = phases["gemm_kernel_ms"]
    gemm_tflops = gemm_flops / (gemm_ms * 1e-3) / 1e12 if gemm_ms > 0 else 0.0
    gemm_tflops_padded = gemm_flops_padded / (gemm_ms * 1e-3) / 1e12 if gemm_ms > 0 else 0.0
    # The first 64-column panel step of each level (the fronts' first
    # diagonal blocks, and the rows below them times W = L_kk^-T D^-1 on
    # FP64 MFMA).
    step_flops = info["panel_step_flops"] / max(info["panel_step_launches"], 1) * useful
    step_ms = phases["panel_step_kernel_ms"]
    step_tflops = step_flops / (step_ms * 1e-3) / 1e12 if step_ms > 0 else 0.0
    # Triangular solves: every factor entry (the supernodal panels as
    # stored) is read once forward and once backward, the kept 64x64
    # diagonal-block inverses once each way.
    solve_ms = phases["solve_ms"]
    solve_bytes = 2 * 8 * info["factor_nnz"] + 2 * 8 * 64 * 64 * info["diag_blocks"]
    solve_gbs = solve_bytes / (solve_ms * 1e-3) / 1e9 if solve_ms > 0 else 0.0
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
            "reference_factor_nnz": ref_nnz,
            "reference_factor_gflop": ref_flops / 1e9,
            "useful_flop_fraction": useful,
            "supernodes": info["num_supernodes"],
            "pattern_epochs": info["pattern_epochs"],
            "reanalyses_in_timed_steps": reanalyses,
            "speculative_reanalyses_in_timed_steps": speculative,
            "reanalysis_ms_in_timed_steps": round(1e3 * reanalysis_s, 2),
            "phase_ms": {k: round(v, 4) for k, v in phases.items()},
        },
        "roofline": {
            "kernel": "KGemmAhead + KGemmLower + KGemmUpdate: trailing updates of the supernodal "
                      "LDL^T on FP64 MFMA, each fused with the next step's diagonal block (factored "
                      "and inverted by every tile-column-0 workgroup) and panel rows (times W), and "
                      "update matrices (flat lists of 64x64 tiles)",
            "bound": "mfma",
            "achieved": gemm_tflops,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": gemm_tflops / FP64_PEAK_TFLOPS,
            "traffic": measured_traffic("KGemm"),
            "flops_per_launch": gemm_flops,
            "flops_per_launch_padded": gemm_flops_padded,
            "achieved_padded": gemm_tflops_padded,
            "launches_per_iteration": info["gemm_launches"],
            "avg_launch_ms": gemm_ms,
            "profile_avg_launch_us": profile_avg_launch_us(["KGemmAhead", "KGemmUpdate", "KGemmLower"]),
            "note": "flops are the useful ones (the padded count of the supernodal superset pattern "
                    "times reference_factor_gflop / factor_gflop; the next step's diagonal block and "
                    "TRSM counted once); latency-bound: the top fronts' steps form a dependent chain, "
                    "each launch as long as its tile-column-0 workgroups' update, 64x64 LDL^T + inverse "
                    "and TRSM",
        },
        "roofline_panel_step": {
            "kernel": "KPanelStep (+KDiagBlock): first 64-column panel step of each level of the "
                      "supernodal LDL^T, L21 = A21 W as a 64x64x64 FP64 MFMA product per 64 rows",
            "bound": "mfma",
            "achieved": step_tflops,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": step_tflops / FP64_PEAK_TFLOPS,
            "traffic": measured_traffic("KPanelStep"),
            "flops_per_launch": step_flops,
            "launches_per_iteration": info["panel_step_launches"],
            "avg_launch_ms": step_ms,
            "profile_avg_launch_us": profile_avg_launch_us(["KPanelStep", "KDiagBlock"], ["KPanelStep"]),
            "note": "latency-bound: one step per level, the fronts' first diagonal blocks then "
                    "their rows; flops are the TRSM's (rows x nb^2) and the block's 2/3 nb^3, not the "
                    "product's, times the useful fraction",
        },
        "roofline_solves": {
            "kernel": "KForward + KBackward (+ Rest): level-scheduled supernodal triangular solves",
            "bound": "hbm",
            "achieved": solve_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": solve_gbs / HBM_PEAK_GBS,
            "bytes_per_solve": solve_bytes,
            "solve_ms": solve_ms,
            "note": "bytes: the stored factor read once each way and the diagonal-block inverses; "
                    "time: both solves of an iteration, back to back; latency-bound: the top fronts' "
                    "blocks are solved one after another by one workgroup each",
        },
        "roofline_assembly": {
            "kernel": "KAssemblePoints (point-row residual + Jacobian assembly)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic("KAssemblePoints"),
            "bytes_per_launch": bytes_launch,
            "avg_launch_ms": kernel_ms,
            "profile_avg_launch_us": profile_avg_launch_us(["KAssemblePoints"]),
        },
        "cpu_baseline": None,
    }
    if not args.no_forms:
        del ctx
        print("[bench] incremental leg", file=sys.stderr, flush=True)
        line["incremental"] = incremental_leg(S, args.incremental_frames)
        print("[bench] served leg", file=sys.stderr, flush=True)
        line["served"] = served_leg(S, args.served_duration, args.plain_duration, args.workers)
        print("[bench] batched leg", file=sys.stderr, flush=True)
        line["batched"] = batched_leg(S)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(problem, args.cpu_iterations)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
