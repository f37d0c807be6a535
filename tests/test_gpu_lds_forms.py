"""The pose-block and panel extend-add kernels' LDS forms against their
global-memory forms (sdfslam_hip_set_lds_forms): the LDS forms serve the
seeded workloads (frames of at most 192 occupied cells, fronts of at most
2048 rows) and the global forms everything larger, so both are run here on
the same problems and every solve must agree bit for bit: cost history,
map and poses, on a single problem and on a batch of two."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solve(sdfslam, name, seed, iterations):
    prob = sdfslam.Problem.workload(name, seed)
    o, _, _, _ = prob.workload_info()
    o.max_iterations = iterations
    o.step_tolerance = 0.0
    ctx = sdfslam.Context(prob, o)
    r = ctx.solve(o)
    m, p = ctx.state()
    return r["cost_history"], m, p


@pytest.mark.parametrize("name,seed,iterations", [("office_small", 1, 12), ("room_small", 1, 12), ("office200", 1, 3)])
def test_lds_forms_match_global_forms(sdfslam, name, seed, iterations):
    print(f"case lds forms {name} seed {seed}", flush=True)
    try:
        sdfslam.set_lds_forms(True)
        a = _solve(sdfslam, name, seed, iterations)
        sdfslam.set_lds_forms(False)
        b = _solve(sdfslam, name, seed, iterations)
    finally:
        sdfslam.set_lds_forms(True)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_lds_forms_match_in_a_batch(sdfslam):
    print("case lds forms batch", flush=True)
    names = [("batch3", 1000), ("batch7", 1000)]

    def run():
        probs = [sdfslam.Problem.workload(n, s) for n, s in names]
        o, _, _, _ = probs[0].workload_info()
        o.max_iterations = 6
        o.step_tolerance = 0.0
        ctxs = [sdfslam.Context(p, o) for p in probs]
        res, st = sdfslam.batch_solve(ctxs, o)
        assert list(st) == [0] * len(ctxs)
        return [(r["cost_history"],) + tuple(c.state()) for r, c in zip(res, ctxs)]

    try:
        sdfslam.set_lds_forms(True)
        a = run()
        sdfslam.set_lds_forms(False)
        b = run()
    finally:
        sdfslam.set_lds_forms(True)
    for ra, rb in zip(a, b):
        for x, y in zip(ra, rb):
            np.testing.assert_array_equal(x, y)
