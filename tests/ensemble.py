"""Jitter ensembles: how chaotic incremental runs are judged.

A run of the incremental recipe is chaotically sensitive: a 1e-9 relative
change of lambda_init can move its result far. Such runs are compared as
distributions. N runs use lambda_init * (1 + i * jitter), i = 0..N-1
(tools/viz/stability_ensemble.py:33-40, 63-64), each is scored, and the
ensemble is summarised by best / median / worst / spread / diverged, where a
run diverged when its score exceeds divergence_factor times the best
(stability_ensemble.py:43-55).

The reference scores a run by its revisit consistency over real scans; the
synthetic workloads here have ground truth, so a run is scored by its mean
translation error against it (Evaluate, src/pipeline/runner.cpp:145-170).
"""
import numpy as np


def jittered_lambda(lambda_init, i, jitter=1e-9):
    """lambda_init of run i (stability_ensemble.py:33-40: value * (1 + i * jitter))."""
    return float(lambda_init) * (1.0 + i * jitter)


def summarize(scores, divergence_factor=5.0):
    """Ensemble statistics (stability_ensemble.py:43-55)."""
    values = np.asarray(scores, dtype=np.float64)
    best = values.min()
    return {
        "best": float(best),
        "median": float(np.median(values)),
        "worst": float(values.max()),
        "spread": float(values.max() - best),
        "diverged": int((values > divergence_factor * best).sum()),
        "n": int(len(values)),
    }


def mean_translation_error(estimated, ground_truth):
    """Evaluate's mean translation error (runner.cpp:145-170): the mean over
    poses of |t_est - t_gt|."""
    e = np.asarray(estimated, dtype=np.float64).reshape(-1, 3)
    g = np.asarray(ground_truth, dtype=np.float64).reshape(-1, 3)
    if e.shape != g.shape or len(e) < 2:
        raise ValueError("pose count mismatch between estimate and ground truth")
    return float(np.mean(np.hypot(e[:, 0] - g[:, 0], e[:, 1] - g[:, 1])))
