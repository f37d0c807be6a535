"""Problems with the ProblemOptions and residual terms that the seeded
workloads leave at their defaults, rebuilt from a workload's frames:

  huber        huber_delta 0.05 (problem.hpp:70, 115-121; IRLS rows,
               lm_solver.cpp:39-47; the Huber branch of Cost, problem.cpp:140-146)
  smooth_step  smooth_gradient_step = 0.6 cell (SmoothGradientMeters with
               clamped interpolation, residual_math.hpp:63-73; grid_map.hpp:75-83)
  exact        smooth_gradient off (the exact bilinear gradient, grid_map.hpp:98-106)
  active       active_region with margin 2 (problem.cpp:39-70: nodes outside the
               scans' dilated cells are constant, NodeColumn -1)
  relations    relation specs between frames far apart in the trajectory
               (problem.cpp:102-114; tests/test_e2e.cpp:103-116)
"""
import numpy as np

VARIANTS = ("huber", "smooth_step", "exact", "active", "relations")


def variant_problem(sdfslam, name, seed, variant):
    base = sdfslam.Problem.workload(name, seed)
    d = sdfslam.Dataset(name, seed)
    m0, p0 = base.state()
    poses = p0.reshape(-1, 3)
    frames = [d.frame(k) for k in range(d.num_frames)]
    scans = [f[0] for f in frames]
    odometry = np.array([f[2] for f in frames[1:]])
    po = d.problem_options
    g = d.grid
    relations = ()
    if variant == "huber":
        po.huber_delta = 0.05
    elif variant == "smooth_step":
        po.smooth_gradient = 1
        po.smooth_gradient_step = 0.6 * g.dx
    elif variant == "exact":
        po.smooth_gradient = 0
    elif variant == "active":
        po.active_region = 1
        po.active_margin = 2
    elif variant == "relations":
        # Loop-closure-like relations from frame i to frames far along, with
        # measurements from the dataset poses (which integrate noisy odometry).
        n = len(poses)
        pairs = [(0, n // 2), (1, n - 1), (n // 4, (3 * n) // 4), (n // 3, n - 2)]
        import oracle_ctypes
        relations = [(i, j, tuple(oracle_ctypes.relative(frames[i][1], frames[j][1])), 2.0)
                     for i, j in pairs if 0 <= i < j < n]
    else:
        raise ValueError(variant)
    prob = sdfslam.Problem.build((g.nx, g.ny, g.min_x, g.min_y, g.dx, g.dy), poses, scans,
                                 odometry=odometry, relations=relations, options=po, map_values=m0)
    return prob
