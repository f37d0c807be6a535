"""Every benchmark workload through the device LM path, against the oracle.

  room (10x10 m room, 10,225 unknowns, 100 iterations)      oracle trajectory fixture
  office150 incremental (910-pose office, 5 it/increment)   oracle-driven runner fixture
                                                            (first 40 frames) and a jitter
                                                            ensemble against CPU replicates
  office200 batch polish (42,727 unknowns, 200 iterations)  live LM step vs the oracle;
                                                            oracle trajectory fixture
  64 uneven batch instances                                 one batched solve of all 64 equals
                                                            each alone; three instances vs
                                                            oracle trajectory fixtures
  32 served sessions, Poisson 10 Hz                         sessions in the mix equal the same
                                                            sessions run alone, bit for bit

The fixtures are written in the build container by the scripts beside
them in tests/golden/ (oracle only). Tolerances: steps 1e-10 relative;
trajectories 1e-6 m / 1e-7 rad, map 1e-6, cost histories 1e-9 relative;
the 300-frame ensemble runs by their error distribution, each run's error
to 1e-7 relative and its poses within 1 mm / 1e-4 rad (chaotic growth).
"""
import os
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _fixture(name):
    path = os.path.join(GOLDEN, name)
    if not os.path.exists(path):
        pytest.fail(f"missing fixture {path}")
    return np.load(path)


def _fingerprint(prob):
    import sys
    sys.path.insert(0, GOLDEN)
    from make_batch_trajectories import fingerprint
    return fingerprint(prob)


def _check_state(tag, p_dev, p_ref, m_dev=None, m_ref=None):
    dp = (np.asarray(p_dev) - np.asarray(p_ref)).reshape(-1, 3)
    et, er = np.abs(dp[:, :2]).max(), np.abs(dp[:, 2]).max()
    em = np.abs(m_dev - m_ref).max() if m_ref is not None else 0.0
    print(f"  {tag}: pose {et:.3e} m {er:.3e} rad, map {em:.3e}", flush=True)
    return et <= 1e-6 and er <= 1e-7 and em <= 1e-6


def test_office200_step_matches_oracle(sdfslam, oracle):
    # 42,727 unknowns: the device's damped Gauss-Newton step against the
    # oracle's (assembly, J^T J, simplicial LDL^T) at a state the device
    # reached after three iterations.
    print("case office200 live step", flush=True)
    prob = sdfslam.Problem.workload("office200", 1)
    opts, _, _, _ = prob.workload_info()
    ctx = sdfslam.Context(prob, opts)
    ctx.iterate(3)
    m, p = ctx.state()
    for lam in (5.0, 0.5):
        x_dev = ctx.linear_step(lam)
        st, x_or = oracle.linear_step(prob.view, m, p, lam)
        assert st == 0
        rel = np.linalg.norm(x_dev - x_or) / np.linalg.norm(x_or)
        print(f"  lambda {lam}: step rel err {rel:.3e}", flush=True)
        assert rel <= 1e-10


@pytest.mark.parametrize("name", ["room", "office200"])
def test_batch_trajectory_matches_oracle_fixture(sdfslam, name):
    print(f"case trajectory fixture {name}", flush=True)
    fx = _fixture(f"trajectory_{name}.npz")
    prob = sdfslam.Problem.workload(name, int(fx["seed"]))
    assert _fingerprint(prob) == str(fx["fingerprint"])
    m0, p0 = prob.state()
    opts, _, _, _ = prob.workload_info()
    maps = {int(k): fx["maps"][i] for i, k in enumerate(fx["map_iterations"])}
    ok = {}
    for i, k in enumerate(fx["pose_iterations"]):
        opts.max_iterations = int(k)
        res, m, p = sdfslam.solve(prob, opts, m0, p0)
        assert res["iterations"] == k
        ok[int(k)] = _check_state(f"{k} iterations", p, fx["poses"][i], m, maps.get(int(k)))
        hist = fx["cost_history"][: k + 1]
        np.testing.assert_allclose(res["cost_history"], hist, rtol=1e-9)
    assert all(ok.values()), ok


def test_office150_incremental_matches_oracle_fixture(sdfslam):
    # The oracle-driven incremental runner over the first 40 frames.
    print("case office150 incremental parity", flush=True)
    fx = _fixture("incremental_office150.npz")
    frames = int(fx["parity_frames"])
    run = sdfslam.run_incremental(str(fx["workload"]), int(fx["seed"]), max_frames=frames)
    recs = run["records"]
    assert len(recs) == frames - 1
    assert [r["iterations"] for r in recs] == list(fx["parity_iterations"])
    fc = np.array([r["final_cost"] for r in recs])
    np.testing.assert_allclose(fc, fx["parity_final_cost"], rtol=1e-9)
    np.testing.assert_allclose(run["total"]["cost_history"], fx["parity_cost_history"], rtol=1e-9)
    assert _check_state(f"{frames} frames", run["poses"], fx["parity_poses"], run["map"], fx["parity_map"])


def test_office150_jitter_ensemble_matches_cpu_replicates(sdfslam):
    # Chaotic incremental runs are judged by their error distribution over
    # lambda_init * (1 + i * 1e-9), i = 0..7 (tools/viz/stability_ensemble.py).
    print("case office150 jitter ensemble", flush=True)
    import ensemble
    fx = _fixture("incremental_office150.npz")
    frames = int(fx["ensemble_frames"])
    name, seed = str(fx["workload"]), int(fx["seed"])
    full = sdfslam.Problem.workload(name, seed)
    o, _, per_increment, gt = full.workload_info()
    o.max_iterations = per_increment
    errs, same_run = [], []
    for i, lam in enumerate(fx["ensemble_lambda_init"]):
        assert lam == ensemble.jittered_lambda(fx["ensemble_lambda_init"][0], i, float(fx["jitter"]))
        o.lambda_init = float(lam)
        run = sdfslam.run_incremental(name, seed, options=o, max_frames=frames)
        errs.append(ensemble.mean_translation_error(run["poses"], gt[:frames]))
        print(f"  run {i}: mean translation error {errs[-1]:.5f} m (cpu {fx['ensemble_error'][i]:.5f})",
              flush=True)
        dp = (run["poses"] - fx["ensemble_poses"][i]).reshape(-1, 3)
        growth = (np.abs(dp[:, :2]).max(), np.abs(dp[:, 2]).max(), np.hypot(dp[:, 0], dp[:, 1]).max())
        print(f"  run {i} final poses: {growth[0]:.3e} m {growth[1]:.3e} rad", flush=True)
        same_run.append(growth)
    dev = ensemble.summarize(errs)
    cpu = ensemble.summarize(fx["ensemble_error"])
    print(f"  device {dev}\n  cpu    {cpu}", flush=True)
    # The distributions agree (how chaotic runs are judged) ...
    assert dev["diverged"] == cpu["diverged"]
    assert 0.8 * cpu["median"] <= dev["median"] <= 1.25 * cpu["median"]
    assert dev["worst"] <= 1.5 * cpu["worst"]
    # ... and each run stays near its CPU replicate. The runs are mildly
    # chaotic: rounding differences (1e-14 m at 40 frames) grow over the
    # increments to between 1e-9 and 4e-4 m at 300 frames, depending on the
    # solver's summation order, so the bound is a physical one (1 mm,
    # 1e-4 rad). A run's mean error (a mean of pose distances) moves by at
    # most its largest pose displacement, so each error must agree with its
    # replicate's to within that displacement.
    for e, c, g in zip(errs, fx["ensemble_error"], same_run):
        assert abs(e - c) <= g[2] + 1e-12, (e, c, g)
    assert max(g[0] for g in same_run) <= 1e-3, same_run
    assert max(g[1] for g in same_run) <= 1e-4, same_run


def _batch_names():
    return [(f"batch{i}", 1000) for i in range(64)]


def test_batch_of_64_instances_matches_each_alone(sdfslam):
    print("case batched 64 instances", flush=True)
    probs = [sdfslam.Problem.workload(n, s) for n, s in _batch_names()]
    o, _, _, _ = probs[0].workload_info()  # every instance: 30 iterations, lambda 5, factor 0.9
    assert o.max_iterations == 30
    alone = []
    for prob in probs:
        ctx = sdfslam.Context(prob, o)
        res = ctx.solve(o)
        alone.append((res, ctx.state()))
        ctx.close()
    ctxs = [sdfslam.Context(prob, o) for prob in probs]
    info = {}
    results, status = sdfslam.batch_solve(ctxs, o, info)
    assert np.all(status == 0)
    print(f"  batched solve {info['run_seconds']:.3f} s", flush=True)
    for i, ((res, (m, p)), r, ctx) in enumerate(zip(alone, results, ctxs)):
        bm, bp = ctx.state()
        assert r["iterations"] == res["iterations"] == 30, i
        assert np.array_equal(r["cost_history"], res["cost_history"]), i
        assert np.array_equal(bm, m) and np.array_equal(bp, p), i


def test_batch_instances_match_oracle_fixture(sdfslam):
    print("case batch instances vs oracle", flush=True)
    fx = _fixture("batch_instances.npz")
    seed = int(fx["seed"])
    rec = [int(k) for k in fx["record_iterations"]]
    ok = True
    for i in fx["instances"]:
        prob = sdfslam.Problem.workload(f"batch{i}", seed)
        assert _fingerprint(prob) == str(fx[f"fingerprint_{i}"])
        o, _, _, _ = prob.workload_info()
        m0, p0 = prob.state()
        for q, k in enumerate(rec):
            o.max_iterations = k
            res, m, p = sdfslam.solve(prob, o, m0, p0)
            np.testing.assert_allclose(res["cost_history"], fx[f"cost_history_{i}"][: k + 1], rtol=1e-9)
            ok &= _check_state(f"batch{i} {k} iterations", p, fx[f"poses_{i}"][q],
                               m if q == len(rec) - 1 else None,
                               fx[f"map_{i}"] if q == len(rec) - 1 else None)
    assert ok


def test_served_32_sessions_match_sessions_alone(sdfslam):
    # 32 sessions with Poisson 10 Hz frame arrivals for 3 s through one
    # server; four of them, run alone, end in the same state bit for bit.
    print("case served 32 sessions", flush=True)
    sessions, rate, duration = 32, 10.0, 3.0
    server = sdfslam.Server()
    data = [sdfslam.Dataset(f"served{i}", 2000) for i in range(sessions)]
    sids = []
    for d in data:
        xy, pose, _ = d.frame(0)
        sids.append(server.open_session(d.grid, d.problem_options, d.step_options, pose, xy))
    rng = np.random.default_rng(2000)
    events = []
    for s in range(sessions):
        t = rng.exponential(1.0 / rate)
        k = 1
        while t < duration and k < data[s].num_frames:
            events.append((t, s, k))
            t += rng.exponential(1.0 / rate)
            k += 1
    events.sort()
    t0 = time.perf_counter()
    tickets = []
    for t, s, k in events:
        delay = t0 + t - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        xy, _, odo = data[s].frame(k)
        tickets.append((sids[s], server.submit(sids[s], xy, odo)))
    lat = [server.wait(sid, tk)["latency_seconds"] for sid, tk in tickets]
    print(f"  {len(tickets)} increments, p50 {1e3 * np.median(lat):.1f} ms, "
          f"p99 {1e3 * np.quantile(lat, 0.99):.1f} ms", flush=True)
    frames = {s: 1 + sum(1 for e in events if e[1] == s) for s in range(sessions)}
    for s in (0, 9, 21, 31):
        p_mix, m_mix = server.state(sids[s])
        assert len(p_mix) == frames[s]
        run = sdfslam.run_incremental(f"served{s}", 2000, max_frames=frames[s])
        assert np.array_equal(run["poses"], p_mix) and np.array_equal(run["map"], m_mix), s
    server.close()
