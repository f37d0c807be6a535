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
#include <string>
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
  /// problem and keep its memory. With `device_normals` the problem was built
  /// with DeferredNodeNormals and its eikonal normals are computed on the
  /// device with these options (the session must have a context already).
  SolveResult Solve(Problem& problem, const SolverOptions& options,
                    const NormalOptions* device_normals = nullptr);
  /// Whether the session has a device context (made by its first Solve/Load).
  bool HasContext() const { return ctx_ != nullptr; }
  /// Device memory the context holds (0 before the first Solve).
  int64_t DeviceBytes() const;
  /// Frees the device buffers the context outgrew (waits for the device).
  void Trim();
  /// Destroys the context (a later Solve or Load creates a new one).
  void Release();

  /// The two halves of Solve around a solve of the context, for callers that
  /// solve many sessions' contexts in one batch (sdfslam_hip_batch_solve):
  /// Load puts the problem and its state into the context (created on first
  /// use); Collect reads the solved state back into the problem and returns
  /// the result, rethrowing `status` as Solve would.
  sdfslam_hip_context* Load(const Problem& problem, const SolverOptions& options,
                            const NormalOptions* device_normals = nullptr);
  SolveResult Collect(Problem& problem, const sdfslam_solve_result& result, int status,
                      const std::vector<double>& history);
  /// Appends frames to the context's problem, built on the device
  /// (sdfslam_hip_context_append_frames): the next increment of a session
  /// whose context holds the problem over the earlier frames. `estimated`
  /// holds every pose (the new frames' initial poses last) and `map` the
  /// map the problem is built at.
  sdfslam_hip_context* Append(const std::vector<const Scan*>& scans, const std::vector<Pose2>& odometry,
                              const ProblemOptions& options, const GridMap& map,
                              const std::vector<Pose2>& estimated);
  /// Collect for an appended increment: the solved state into `map` and
  /// `estimated`.
  SolveResult CollectState(GridMap& map, std::vector<Pose2>& estimated, const sdfslam_solve_result& result,
                           int status, const std::vector<double>& history);
  /// Solves the context's current problem (after Append) and collects the
  /// state into `map` and `estimated`.
  SolveResult SolveLoaded(GridMap& map, std::vector<Pose2>& estimated, const SolverOptions& options);

 private:
  /// Creates the context or replaces its problem.
  void Place(const Problem& problem, const SolverOptions& options, const NormalOptions* device_normals);
  sdfslam_hip_context* ctx_{nullptr};
};

/// Whether a problem with these options can take its eikonal normals from
/// the device (DeferredNodeNormals): all-scans neighbourhood, pca or hybrid.
bool DeviceNormalsApply(const ProblemOptions& options);

/// Whether a session with these options can append its frames on the
/// device (HipSession::Append): device normals apply, no relations, and a
/// positive hallucination step when it hallucinates.
bool AppendApplies(const ProblemOptions& options, bool has_relations);

/// Per-increment record.
struct IncrementRecord {
  int frames{0};  // frames in the problem of this increment
  int iterations{0};
  double initial_cost{0.0};
  double final_cost{0.0};
  double build_seconds{0.0};  // Problem construction
  double solve_seconds{0.0};  // the increment's Solve, transfers included
  // Served increments: arrival to the start of the preparation, the
  // preparation, and prepared to the start of the batched solve.
  double queue_seconds{0.0};
  double prepare_seconds{0.0};
  double solve_wait_seconds{0.0};
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

/// Seeds every map node with the distance to the nearest scan point on the
/// device (InitializeFromScans, src/core/map_init.cpp:31-69): points are the
/// scans placed by their poses, each with its pose's translation as origin;
/// with `signed_distance` a node farther from the nearest point's origin
/// than the point is gets the negative distance. No scans: map unchanged.
void InitializeFromScans(GridMap& map, const std::vector<Scan>& scans, const std::vector<Pose2>& poses,
                         bool signed_distance);

/// Optional parts of an incremental run (runner.cpp:229-233, 271-284): map
/// seeding from the first scan before the first increment, and a callback
/// after each increment with the frames so far, the estimate and the map.
struct IncrementHooks {
  bool init_from_scans{false};
  bool init_signed{true};
  // Device runs extend the context's problem frame by frame on the device
  // (HipSession::Append) where that applies; false rebuilds every increment.
  bool append_frames{true};
  std::function<void(size_t frames, size_t increment, const std::vector<Pose2>& estimated, const GridMap& map)>
      after_increment;
};

IncrementalRun RunIncremental(const Workload& wl, const SolverOptions& solver, int increment_size,
                              int max_frames, const IncrementSolveFn& solve_fn = nullptr,
                              const IncrementHooks& hooks = IncrementHooks());

/// Trajectory errors against ground truth (runner.cpp:145-170): mean
/// translation and absolute wrapped rotation errors, and the same for the
/// relative motions of consecutive frames.
struct TrajectoryMetrics {
  double mean_translation_error{0.0};
  double mean_rotation_error{0.0};
  double mean_relative_translation_error{0.0};
  double mean_relative_rotation_error{0.0};
};
TrajectoryMetrics Evaluate(const std::vector<Pose2>& estimated, const std::vector<Pose2>& ground_truth);

/// Output files of a run, in the reference's formats (runner.cpp:73-113,
/// 172-192): poses as "x,y,theta" CSV and the map as a header line
/// "nx ny min_x min_y dx dy" plus one line of values per grid row, both at
/// 17 significant digits; global scan points as "x,y" CSV at 9 digits;
/// metrics.json with the solve totals and, given ground truth, the metrics.
void WritePoses(const std::string& path, const std::vector<Pose2>& poses);
void WriteMap(const std::string& path, const GridMap& map);
void WriteGlobalScanPoints(const std::string& path, const std::vector<Scan>& scans,
                           const std::vector<Pose2>& poses);
void WriteMetricsJson(const std::string& path, const SolveResult& result, const TrajectoryMetrics* metrics);

/// A whole run over a workload (Run, runner.cpp:196-303): batch (one solve
/// from the workload's initial poses) or incremental (`increment_size`
/// frames per increment, the solver's max_iterations per increment), map
/// seeding from the scans when asked, snapshots every `snapshot_every`
/// increments (and after the last), and the output files in `output_dir`
/// (none when empty): poses_odometry.csv, poses_estimated.csv, map.txt,
/// scan_points_global.csv, poses_ground_truth.csv and metrics.json.
struct RunOptions {
  bool incremental{false};
  int increment_size{1};
  int max_frames{0};  // incremental runs: frames used (all when <= 0)
  int snapshot_every{0};
  bool init_from_scans{false};
  bool init_signed{true};
  std::string output_dir;
};
IncrementalRun RunWorkload(const Workload& wl, const SolverOptions& solver, const RunOptions& options,
                           const IncrementSolveFn& solve_fn = nullptr);

}  // namespace sdf_slam
