"""Served form (many incremental sessions on one GPU) and the batched form:
a session gives the same results alone and in a mix, the same as the
incremental runner, and the benchmarks run end to end."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _open(sdfslam, server, name, frames):
    d = sdfslam.Dataset(name, 2000)
    xy, pose, _ = d.frame(0)
    sid = server.open_session(d.grid, d.problem_options, d.step_options, pose, xy)
    return d, sid, [d.frame(k) for k in range(1, frames)]


def test_session_alone_matches_mix_and_runner(sdfslam):
    print("case served alone vs mix", flush=True)
    frames = 6
    alone = sdfslam.Server(workers=1, max_batch=1, pipelines=1)
    d0, s0, f0 = _open(sdfslam, alone, "served0", frames)
    for xy, _, odo in f0:
        alone.wait(s0, alone.submit(s0, xy, odo))
    p_alone, m_alone = alone.state(s0)
    alone.close()

    mix = sdfslam.Server(workers=4)
    opened = [_open(sdfslam, mix, name, frames) for name in ("served0", "served1", "served2")]
    tickets = []
    for k in range(frames - 1):
        for _, sid, fr in opened:
            xy, _, odo = fr[k]
            tickets.append((sid, mix.submit(sid, xy, odo)))
    recs = [mix.wait(sid, t) for sid, t in tickets]
    assert all(r["frames"] >= 2 and r["latency_seconds"] > 0 for r in recs)
    p_mix, m_mix = mix.state(opened[0][1])
    mix.close()
    assert np.array_equal(p_alone, p_mix) and np.array_equal(m_alone, m_mix)

    run = sdfslam.run_incremental("served0", 2000, max_frames=frames)
    assert np.array_equal(run["poses"], p_alone) and np.array_equal(run["map"], m_alone)


def test_benchmarks_run(sdfslam):
    print("case served/batch benchmarks", flush=True)
    b = sdfslam.batch_benchmark(instances=3, iterations=3, max_batch=2)
    assert b["lm_iterations"] == 9 and b["lm_iterations_per_second"] > 0
    r = sdfslam.served_benchmark(sessions=2, rate_hz=20.0, duration_s=0.4, workers=2)
    assert r["increments"] > 0 and r["latency_p50_ms"] <= r["latency_p99_ms"]
    assert r["lm_iterations"] > 0 and r["device_bytes"] > 0


@pytest.mark.parametrize("tol", [0.0, 1e-3])
def test_batch_solve_matches_each_alone(sdfslam, tol):
    # One launch per stage over uneven problems gives each problem exactly
    # the result it gets alone (also when some stop early on the step
    # tolerance while the others go on).
    print(f"case batch solve tol={tol}", flush=True)
    names = [("office_small", 1), ("corner", 1), ("room_small", 1), ("office_small", 2), ("batch3", 1000)]
    o = sdfslam.default_solver_options(max_iterations=8, step_tolerance=tol)
    alone = []
    alone_ctx = []
    for name, seed in names:
        ctx = sdfslam.Context(sdfslam.Problem.workload(name, seed), o)
        res = ctx.solve(o)
        alone.append((res, ctx.state()))
        alone_ctx.append(ctx)
    ctxs = [sdfslam.Context(sdfslam.Problem.workload(name, seed), o) for name, seed in names]
    info = {}
    results, status = sdfslam.batch_solve(ctxs, o, info)
    assert np.all(status == 0)
    assert info["graph_captures"] >= 1 and info["graph_launches"] >= 1
    assert info["run_seconds"] >= info["capture_seconds"] > 0.0
    for (res, (m, p)), r, ctx in zip(alone, results, ctxs):
        bm, bp = ctx.state()
        assert r["iterations"] == res["iterations"]
        assert np.array_equal(r["cost_history"], res["cost_history"])
        assert np.array_equal(bm, m) and np.array_equal(bp, p)
    # A second, smaller batch led by the same context (its batch buffers are
    # reused) continues from the current states exactly as solving alone.
    again = [(c.solve(o), c.state()) for c in alone_ctx[:3]]
    results, status = sdfslam.batch_solve(ctxs[:3], o)
    assert np.all(status == 0)
    for (res, (m, p)), r, ctx in zip(again, results, ctxs[:3]):
        bm, bp = ctx.state()
        assert r["iterations"] == res["iterations"]
        assert np.array_equal(r["cost_history"], res["cost_history"])
        assert np.array_equal(bm, m) and np.array_equal(bp, p)


def test_batch_autotune_picks_its_fastest_trial(sdfslam):
    print("case batch autotune", flush=True)
    best, trials = sdfslam.batch_autotune(instances=4, max_batches=(1, 2), workers=(1, 2), trial_iterations=2)
    assert [(t[0], t[1]) for t in trials] == [(1, 1), (1, 2), (2, 1), (2, 2)]
    assert all(t[2] > 0 for t in trials)
    assert best == max(trials, key=lambda t: t[2])


def test_sessions_with_different_step_options(sdfslam):
    # Two sessions with their own max_iterations and lambda_factor in one
    # server: each ends as its own incremental run with its options (they
    # are never batched together).
    print("case served per-session options", flush=True)
    frames = 6
    d = [sdfslam.Dataset(f"served{i}", 2000) for i in (3, 4)]
    opts = [d[0].step_options, d[1].step_options]
    opts[1].max_iterations = 3
    opts[1].lambda_factor = 2.0
    server = sdfslam.Server(workers=4)
    sids = []
    for di, o in zip(d, opts):
        xy, pose, _ = di.frame(0)
        sids.append(server.open_session(di.grid, di.problem_options, o, pose, xy))
    tickets = []
    for k in range(1, frames):
        for di, sid in zip(d, sids):
            xy, _, odo = di.frame(k)
            tickets.append((sid, server.submit(sid, xy, odo)))
    recs = {sid: [] for sid in sids}
    for sid, t in tickets:
        recs[sid].append(server.wait(sid, t))
    assert all(r["iterations"] <= 3 for r in recs[sids[1]])
    for i, (sid, o) in enumerate(zip(sids, opts)):
        p, m = server.state(sid)
        run = sdfslam.run_incremental(f"served{(3, 4)[i]}", 2000, options=o, max_frames=frames)
        assert np.array_equal(run["poses"], p) and np.array_equal(run["map"], m), i
        assert [r["iterations"] for r in recs[sid]] == [r["iterations"] for r in run["records"]]
    server.close()


def test_queued_session_is_admitted_after_another_closes(sdfslam):
    # A budget that holds one session: the second waits in the queue with
    # its frames accepted, and is admitted and served once the first closes.
    print("case served admission queue", flush=True)
    d = [sdfslam.Dataset(f"served{i}", 2000) for i in (0, 1, 2)]
    reserve = max((64 << 20) + 4096 * di.grid.nx * di.grid.ny for di in d)
    server = sdfslam.Server(workers=2, device_budget_bytes=reserve + (1 << 20))
    sids = []
    for di in d:
        xy, pose, _ = di.frame(0)
        sids.append(server.open_session(di.grid, di.problem_options, di.step_options, pose, xy))
    a, b, c = sids
    assert server.session_info(a)["admitted"]
    assert not server.session_info(b)["admitted"] and not server.session_info(c)["admitted"]
    tb = [server.submit(b, *(lambda f: (f[0], f[2]))(d[1].frame(k))) for k in (1, 2, 3)]
    tc = server.submit(c, *(lambda f: (f[0], f[2]))(d[2].frame(1)))
    ta = [server.submit(a, *(lambda f: (f[0], f[2]))(d[0].frame(k))) for k in (1, 2)]
    for t in ta:
        server.wait(a, t)
    info = server.session_info(b)
    assert not info["admitted"] and info["finished"] == 0 and info["submitted"] == 3
    # Closing the queued third session drops its frame.
    server.close_session(c)
    with pytest.raises(Exception, match="closed before it was admitted"):
        server.wait(c, tc)
    server.close_session(a)
    assert server.session_info(a)["closed"] and server.session_info(a)["device_bytes"] == 0
    assert server.session_info(b)["admitted"]
    for t in tb:
        server.wait(b, t)
    p, m = server.state(b)
    run = sdfslam.run_incremental("served1", 2000, max_frames=4)
    assert np.array_equal(run["poses"], p) and np.array_equal(run["map"], m)
    with pytest.raises(ValueError, match="closed"):
        server.submit(a, *(lambda f: (f[0], f[2]))(d[0].frame(3)))
    server.close()


def test_device_normals_match_host_normals(sdfslam):
    # Increments after a session's first take their eikonal normals from the
    # device inside the batched solve; every state equals the host-built
    # sessions' and the incremental runner's bit for bit.
    print("case served device normals", flush=True)
    frames = 7
    states = {}
    for device_normals in (False, True):
        server = sdfslam.Server(workers=4, device_normals=device_normals)
        opened = [_open(sdfslam, server, name, frames) for name in ("served5", "served6")]
        tickets = []
        for k in range(frames - 1):
            for _, sid, fr in opened:
                xy, _, odo = fr[k]
                tickets.append((sid, server.submit(sid, xy, odo)))
        for sid, t in tickets:
            server.wait(sid, t)
        states[device_normals] = [server.state(sid) for _, sid, _ in opened]
        server.close()
    for (p0, m0), (p1, m1) in zip(states[False], states[True]):
        assert np.array_equal(p0, p1) and np.array_equal(m0, m1)
    run = sdfslam.run_incremental("served5", 2000, max_frames=frames)
    assert np.array_equal(run["poses"], states[True][0][0]) and np.array_equal(run["map"], states[True][0][1])
