// Incremental driver of the host mirror: the reference's incremental run
// (src/pipeline/runner.cpp:226-285, dissertation algorithm 2). Each
// increment appends frames by composing odometry onto the last optimized
// pose, rebuilds the Problem over every frame so far, and refines it with a
// few solver iterations.
//
// The solve of each increment is pluggable: the device backend through a
// persistent session (one context whose device memory is reused from one
// increment to the next), or a caller-supplied function with the signature
// of sdfslam_hip_solve (how the tests run the same driver on the CPU path).
#pragma once

#include <functional>
#include <vector>

#include "problem.hpp"
#include "scene.hpp"
#include "solver.hpp"

namespace sdf_slam {

/// One device context reused across solves of growing problems.
class HipSession {
 public:
  HipSession() = default;
  HipSession(const HipSession&) = delete;
  HipSession& operator=(const HipSession&) = delete;
  ~HipSession();
  /// Solve(problem, options) on the device (same results and exceptions as
  /// SolveHip); the first call creates the context, later calls replace its
  /// problem and keep its memory.
  SolveResult Solve(Problem& problem, const SolverOptions& options);
  /// Device memory the context holds (0 before the first Solve).
  int64_t DeviceBytes() const;

  /// The two halves of Solve around a solve of the context, for callers that
  /// solve many sessions' contexts in one batch (sdfslam_hip_batch_solve):
  /// Load puts the problem and its state into the context (created on first
  /// use); Collect reads the solved state back into the problem and returns
  /// the result, rethrowing `status` as Solve would.
  sdfslam_hip_context* Load(const Problem& problem, const SolverOptions& options);
  SolveResult Collect(Problem& problem, const sdfslam_solve_result& result, int status,
                      const std::vector<double>& history);

 private:
  sdfslam_hip_context* ctx_{nullptr};
};

/// Per-increment record.
struct IncrementRecord {
  int frames{0};  // frames in the problem of this increment
  int iterations{0};
  double initial_cost{0.0};
  double final_cost{0.0};
  double build_seconds{0.0};  // Problem construction
  double solve_seconds{0.0};  // the increment's Solve, transfers included
};

struct IncrementalRun {
  SolveResult total;  // accumulated as runner.cpp:262-270 does
  std::vector<Pose2> estimated;
  GridMap map;
  std::vector<IncrementRecord> increments;
};

/// Solve function of one increment: (problem view, options, map values in/out,
/// poses in/out, result) -> status, as sdfslam_hip_solve.
using IncrementSolveFn = std::function<int(const sdfslam_problem_view*, const sdfslam_solver_options*,
                                           double*, double*, sdfslam_solve_result*)>;

/// Runs the workload incrementally over its first `max_frames` frames (all
/// when <= 0), `increment_size` frames per increment, each refined with
/// `solver` (whose max_iterations is the iterations per increment). With a
/// null `solve_fn` the device backend solves through one HipSession.
/// Node normals of a problem build on the device (sdfslam_hip_node_normals):
/// a NodeNormalsFn for Problem's constructor.
std::vector<NodeNormal> DeviceNodeNormals(const std::vector<Vec2>& points, const std::vector<Vec2>& origins,
                                          const GridMap& map, const NormalOptions& options);

IncrementalRun RunIncremental(const Workload& wl, const SolverOptions& solver, int increment_size,
                              int max_frames, const IncrementSolveFn& solve_fn = nullptr);

}  // namespace sdf_slam
