// Incremental driver (runner.cpp:226-285) and the persistent device session.
#include "runner.hpp"

#include <algorithm>
#include <chrono>
#include <stdexcept>
#include <string>

namespace sdf_slam {

namespace {

double Seconds(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

/// Maps a status of the C ABI to the exception SolveLm throws.
void Rethrow(int status) {
  if (status == SDFSLAM_OK) return;
  if (status == SDFSLAM_ERR_INVALID_ARGUMENT) throw std::invalid_argument(sdfslam_hip_last_error());
  if (status == SDFSLAM_ERR_NOT_SPD) throw std::runtime_error("sparse Cholesky factorization failed");
  throw std::runtime_error(std::string(sdfslam_hip_status_string(status)) + ": " +
                           sdfslam_hip_last_error());
}

}  // namespace

HipSession::~HipSession() {
  if (ctx_ != nullptr) (void)sdfslam_hip_context_destroy(ctx_);
}

int64_t HipSession::DeviceBytes() const {
  if (ctx_ == nullptr) return 0;
  sdfslam_hip_analysis_info info{};
  if (sdfslam_hip_analysis(ctx_, &info) != SDFSLAM_OK) return 0;
  return info.device_bytes;
}

sdfslam_hip_context* HipSession::Load(const Problem& problem, const SolverOptions& options) {
  const sdfslam_problem_view view = problem.View();
  const sdfslam_solver_options copts = ToC(options);
  std::vector<double> map_values;
  std::vector<double> poses;
  problem.GetState(map_values, poses);
  if (ctx_ == nullptr) {
    Rethrow(sdfslam_hip_context_create(&view, &copts, map_values.data(), poses.data(), &ctx_));
  } else {
    Rethrow(sdfslam_hip_context_update(ctx_, &view, map_values.data(), poses.data()));
  }
  return ctx_;
}

SolveResult HipSession::Collect(Problem& problem, const sdfslam_solve_result& r, int status,
                                const std::vector<double>& history) {
  Rethrow(status);
  std::vector<double> map_values;
  std::vector<double> poses;
  problem.GetState(map_values, poses);
  Rethrow(sdfslam_hip_get_state(ctx_, map_values.data(), poses.data()));
  problem.SetState(map_values.data(), poses.data());
  SolveResult result;
  result.converged = r.converged != 0;
  result.iterations = r.iterations;
  result.initial_cost = r.initial_cost;
  result.final_cost = r.final_cost;
  result.duration_seconds = r.duration_seconds;
  result.cost_history.assign(history.begin(),
                             history.begin() + std::min<std::ptrdiff_t>(
                                                   static_cast<std::ptrdiff_t>(history.size()),
                                                   std::min(r.cost_history_length, r.cost_history_capacity)));
  return result;
}

std::vector<NodeNormal> DeviceNodeNormals(const std::vector<Vec2>& points, const std::vector<Vec2>& origins,
                                          const GridMap& map, const NormalOptions& options) {
  const size_t n = points.size();
  std::vector<double> xy(2 * n);
  std::vector<double> oxy(2 * n);
  for (size_t i = 0; i < n; ++i) {
    xy[2 * i] = points[i].x;
    xy[2 * i + 1] = points[i].y;
    oxy[2 * i] = origins[i].x;
    oxy[2 * i + 1] = origins[i].y;
  }
  const sdfslam_grid g{map.nx(), map.ny(), map.min_corner().x, map.min_corner().y, map.dx(), map.dy()};
  const auto nodes = static_cast<size_t>(map.num_nodes());
  std::vector<double> nxy(2 * nodes);
  std::vector<double> scale(nodes);
  const int method = options.method == NormalMethod::kPca      ? SDFSLAM_NORMALS_PCA
                     : options.method == NormalMethod::kHybrid ? SDFSLAM_NORMALS_HYBRID
                                                               : SDFSLAM_NORMALS_WEIGHTED;
  Rethrow(sdfslam_hip_node_normals(xy.data(), oxy.data(), static_cast<int64_t>(n), options.k_neighbors, &g,
                                   method, options.corner_threshold, nxy.data(), scale.data()));
  std::vector<NodeNormal> out(nodes);
  for (size_t v = 0; v < nodes; ++v) {
    out[v].normal = Vec2{nxy[2 * v], nxy[2 * v + 1]};
    out[v].eikonal_scale = scale[v];
  }
  return out;
}

SolveResult HipSession::Solve(Problem& problem, const SolverOptions& options) {
  const auto t0 = std::chrono::steady_clock::now();
  const sdfslam_problem_view view = problem.View();
  const sdfslam_solver_options copts = ToC(options);
  std::vector<double> map_values;
  std::vector<double> poses;
  problem.GetState(map_values, poses);
  if (ctx_ == nullptr) {
    Rethrow(sdfslam_hip_context_create(&view, &copts, map_values.data(), poses.data(), &ctx_));
  } else {
    Rethrow(sdfslam_hip_context_update(ctx_, &view, map_values.data(), poses.data()));
  }
  SolveResult result;
  result.cost_history.assign(static_cast<size_t>(std::max(0, options.max_iterations)) + 1, 0.0);
  sdfslam_solve_result cres{};
  cres.cost_history = result.cost_history.data();
  cres.cost_history_capacity = static_cast<int32_t>(result.cost_history.size());
  Rethrow(sdfslam_hip_context_solve(ctx_, &copts, &cres));
  Rethrow(sdfslam_hip_get_state(ctx_, map_values.data(), poses.data()));
  problem.SetState(map_values.data(), poses.data());
  result.converged = cres.converged != 0;
  result.iterations = cres.iterations;
  result.initial_cost = cres.initial_cost;
  result.final_cost = cres.final_cost;
  result.duration_seconds = Seconds(t0);
  result.cost_history.resize(static_cast<size_t>(
      std::min(cres.cost_history_length, cres.cost_history_capacity)));
  return result;
}

IncrementalRun RunIncremental(const Workload& wl, const SolverOptions& solver, int increment_size,
                              int max_frames, const IncrementSolveFn& solve_fn) {
  const size_t total_frames =
      max_frames > 0 ? std::min(wl.scans.size(), static_cast<size_t>(max_frames)) : wl.scans.size();
  if (total_frames < 2) {
    throw std::runtime_error("dataset needs at least two scans");
  }
  if (increment_size < 1) {
    throw std::invalid_argument("increment_size must be >= 1");
  }
  HipSession session;
  IncrementalRun run;
  run.map = wl.map;
  run.estimated = {wl.dataset_poses[0]};
  run.total.initial_cost = -1.0;
  size_t next_frame = 1;
  while (next_frame < total_frames) {
    const size_t last = std::min(total_frames, next_frame + static_cast<size_t>(increment_size));
    for (size_t i = next_frame; i < last; ++i) {
      run.estimated.push_back(run.estimated.back().Compose(wl.odometry[i - 1]));
    }
    next_frame = last;
    IncrementRecord rec;
    rec.frames = static_cast<int>(next_frame);
    auto t0 = std::chrono::steady_clock::now();
    const std::vector<Scan> scans_so_far(wl.scans.begin(),
                                         wl.scans.begin() + static_cast<std::ptrdiff_t>(next_frame));
    const std::vector<Pose2> odom_so_far(wl.odometry.begin(),
                                         wl.odometry.begin() + static_cast<std::ptrdiff_t>(next_frame - 1));
    // The device runs build their node normals on the device too.
    Problem problem(std::move(run.map), run.estimated, scans_so_far, odom_so_far, wl.problem,
                    wl.relations, solve_fn ? NodeNormalsFn() : NodeNormalsFn(DeviceNodeNormals));
    rec.build_seconds = Seconds(t0);
    t0 = std::chrono::steady_clock::now();
    SolveResult step;
    if (solve_fn) {
      const sdfslam_problem_view view = problem.View();
      const sdfslam_solver_options copts = ToC(solver);
      std::vector<double> map_values;
      std::vector<double> poses;
      problem.GetState(map_values, poses);
      step.cost_history.assign(static_cast<size_t>(std::max(0, solver.max_iterations)) + 1, 0.0);
      sdfslam_solve_result cres{};
      cres.cost_history = step.cost_history.data();
      cres.cost_history_capacity = static_cast<int32_t>(step.cost_history.size());
      Rethrow(solve_fn(&view, &copts, map_values.data(), poses.data(), &cres));
      problem.SetState(map_values.data(), poses.data());
      step.converged = cres.converged != 0;
      step.iterations = cres.iterations;
      step.initial_cost = cres.initial_cost;
      step.final_cost = cres.final_cost;
      step.cost_history.resize(static_cast<size_t>(
          std::min(cres.cost_history_length, cres.cost_history_capacity)));
    } else {
      step = session.Solve(problem, solver);
    }
    rec.solve_seconds = Seconds(t0);
    step.duration_seconds = rec.solve_seconds;
    run.estimated = problem.poses();
    run.map = problem.map();
    rec.iterations = step.iterations;
    rec.initial_cost = step.initial_cost;
    rec.final_cost = step.final_cost;
    run.increments.push_back(rec);
    // Totals as runner.cpp:262-270.
    if (run.total.initial_cost < 0.0) {
      run.total.initial_cost = step.initial_cost;
    }
    run.total.iterations += step.iterations;
    run.total.final_cost = step.final_cost;
    run.total.duration_seconds += step.duration_seconds;
    run.total.converged = step.converged;
    run.total.cost_history.insert(run.total.cost_history.end(), step.cost_history.begin(),
                                  step.cost_history.end());
  }
  return run;
}

}  // namespace sdf_slam
