// Incremental driver (runner.cpp:226-285) and the persistent device session.
#include "runner.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
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

void HipSession::Trim() {
  if (ctx_ != nullptr) Rethrow(sdfslam_hip_context_trim(ctx_));
}

void HipSession::Release() {
  if (ctx_ != nullptr) (void)sdfslam_hip_context_destroy(ctx_);
  ctx_ = nullptr;
}

bool DeviceNormalsApply(const ProblemOptions& options) {
  return !options.normals.per_scan && options.normals.method != NormalMethod::kWeighted;
}

void HipSession::Place(const Problem& problem, const SolverOptions& options, const NormalOptions* device_normals) {
  const sdfslam_problem_view view = problem.View();
  const sdfslam_solver_options copts = ToC(options);
  std::vector<double> map_values;
  std::vector<double> poses;
  problem.GetState(map_values, poses);
  if (ctx_ == nullptr) {
    if (device_normals != nullptr) throw std::logic_error("device normals need an existing context");
    Rethrow(sdfslam_hip_context_create(&view, &copts, map_values.data(), poses.data(), &ctx_));
  } else if (device_normals != nullptr) {
    // The frames' cos and sin come from ::sincos in the backend, as Pose2 computes them.
    sdfslam_normals_request req{};
    req.k_neighbors = device_normals->k_neighbors;
    req.method = device_normals->method == NormalMethod::kPca ? SDFSLAM_NORMALS_PCA : SDFSLAM_NORMALS_HYBRID;
    req.corner_threshold = device_normals->corner_threshold;
    req.frame_cos_sin = nullptr;
    Rethrow(sdfslam_hip_context_update_device_normals(ctx_, &view, map_values.data(), poses.data(), &req));
  } else {
    Rethrow(sdfslam_hip_context_update(ctx_, &view, map_values.data(), poses.data()));
  }
}

sdfslam_hip_context* HipSession::Load(const Problem& problem, const SolverOptions& options,
                                      const NormalOptions* device_normals) {
  Place(problem, options, device_normals);
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

bool AppendApplies(const ProblemOptions& options, bool has_relations) {
  return DeviceNormalsApply(options) && !has_relations &&
         (options.hallucination.points_per_scan_point <= 0 || options.hallucination.step_size > 0.0);
}

sdfslam_hip_context* HipSession::Append(const std::vector<const Scan*>& scans, const std::vector<Pose2>& odometry,
                                        const ProblemOptions& options, const GridMap& map,
                                        const std::vector<Pose2>& estimated) {
  if (ctx_ == nullptr) throw std::logic_error("appending frames needs an existing context");
  if (scans.size() != odometry.size() || scans.empty()) {
    throw std::invalid_argument("append needs one odometry measurement per scan");
  }
  std::vector<int64_t> counts;
  std::vector<double> xy;
  std::vector<double> odo;
  for (size_t f = 0; f < scans.size(); ++f) {
    counts.push_back(static_cast<int64_t>(scans[f]->points.size()));
    for (const Vec2& p : scans[f]->points) {
      xy.push_back(p.x);
      xy.push_back(p.y);
    }
    odo.push_back(odometry[f].x);
    odo.push_back(odometry[f].y);
    odo.push_back(odometry[f].theta);
  }
  sdfslam_frame_options fo{};
  fo.hallucination_points = options.hallucination.points_per_scan_point;
  fo.hallucination_step = options.hallucination.step_size;
  fo.hallucination_both_directions = options.hallucination.both_directions ? 1 : 0;
  fo.weight_scan = options.weights.scan;
  fo.weight_hallucination = options.weights.hallucination;
  fo.weight_eikonal = options.weights.eikonal;
  fo.weight_odometry = options.weights.odometry;
  fo.active_region = options.active_region ? 1 : 0;
  fo.active_margin = options.active_margin;
  fo.normals.k_neighbors = options.normals.k_neighbors;
  fo.normals.method = options.normals.method == NormalMethod::kPca ? SDFSLAM_NORMALS_PCA : SDFSLAM_NORMALS_HYBRID;
  fo.normals.corner_threshold = options.normals.corner_threshold;
  fo.normals.frame_cos_sin = nullptr;
  std::vector<double> poses(3 * estimated.size());
  for (size_t i = 0; i < estimated.size(); ++i) {
    poses[3 * i] = estimated[i].x;
    poses[3 * i + 1] = estimated[i].y;
    poses[3 * i + 2] = estimated[i].theta;
  }
  Rethrow(sdfslam_hip_context_append_frames(ctx_, static_cast<int32_t>(scans.size()), counts.data(), xy.data(),
                                            odo.data(), &fo, map.values().data(), poses.data()));
  return ctx_;
}

SolveResult HipSession::SolveLoaded(GridMap& map, std::vector<Pose2>& estimated, const SolverOptions& options) {
  if (ctx_ == nullptr) throw std::logic_error("no context to solve");
  const sdfslam_solver_options copts = ToC(options);
  std::vector<double> history(static_cast<size_t>(std::max(0, options.max_iterations)) + 1, 0.0);
  sdfslam_solve_result cres{};
  cres.cost_history = history.data();
  cres.cost_history_capacity = static_cast<int32_t>(history.size());
  const int st = sdfslam_hip_context_solve(ctx_, &copts, &cres);
  return CollectState(map, estimated, cres, st, history);
}

SolveResult HipSession::CollectState(GridMap& map, std::vector<Pose2>& estimated, const sdfslam_solve_result& r,
                                     int status, const std::vector<double>& history) {
  Rethrow(status);
  std::vector<double> poses(3 * estimated.size());
  Rethrow(sdfslam_hip_get_state(ctx_, map.values().data(), poses.data()));
  for (size_t i = 0; i < estimated.size(); ++i) estimated[i] = {poses[3 * i], poses[3 * i + 1], poses[3 * i + 2]};
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

SolveResult HipSession::Solve(Problem& problem, const SolverOptions& options, const NormalOptions* device_normals) {
  const auto t0 = std::chrono::steady_clock::now();
  const sdfslam_solver_options copts = ToC(options);
  std::vector<double> map_values;
  std::vector<double> poses;
  problem.GetState(map_values, poses);
  Place(problem, options, device_normals);
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

void InitializeFromScans(GridMap& map, const std::vector<Scan>& scans, const std::vector<Pose2>& poses,
                         bool signed_distance) {
  std::vector<double> xy;
  std::vector<double> oxy;
  for (size_t i = 0; i < scans.size() && i < poses.size(); ++i) {
    const Vec2 origin = poses[i].Translation();
    for (const Vec2& p : scans[i].points) {
      const Vec2 g = poses[i].Apply(p);
      xy.push_back(g.x);
      xy.push_back(g.y);
      oxy.push_back(origin.x);
      oxy.push_back(origin.y);
    }
  }
  if (xy.empty()) return;
  const sdfslam_grid g{map.nx(), map.ny(), map.min_corner().x, map.min_corner().y, map.dx(), map.dy()};
  Rethrow(sdfslam_hip_init_from_scans(xy.data(), oxy.data(), static_cast<int64_t>(xy.size() / 2), &g,
                                      signed_distance ? 1 : 0, map.values().data()));
}

namespace {

/// One solve through a function with the signature of sdfslam_hip_solve.
SolveResult SolveWith(const IncrementSolveFn& solve_fn, Problem& problem, const SolverOptions& solver) {
  SolveResult step;
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
  step.duration_seconds = cres.duration_seconds;
  step.cost_history.resize(static_cast<size_t>(std::min(cres.cost_history_length, cres.cost_history_capacity)));
  return step;
}

/// One increment's record and the run's totals (runner.cpp:262-270).
void Record(IncrementalRun& run, IncrementRecord& rec, const SolveResult& step) {
  rec.iterations = step.iterations;
  rec.initial_cost = step.initial_cost;
  rec.final_cost = step.final_cost;
  run.increments.push_back(rec);
  if (run.total.initial_cost < 0.0) {
    run.total.initial_cost = step.initial_cost;
  }
  run.total.iterations += step.iterations;
  run.total.final_cost = step.final_cost;
  run.total.duration_seconds += step.duration_seconds;
  run.total.converged = step.converged;
  run.total.cost_history.insert(run.total.cost_history.end(), step.cost_history.begin(), step.cost_history.end());
}

}  // namespace

IncrementalRun RunIncremental(const Workload& wl, const SolverOptions& solver, int increment_size,
                              int max_frames, const IncrementSolveFn& solve_fn, const IncrementHooks& hooks) {
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
  if (hooks.init_from_scans) {
    // runner.cpp:231-233: the first scan at the first dataset pose.
    InitializeFromScans(run.map, {wl.scans[0]}, {wl.dataset_poses[0]}, hooks.init_signed);
  }
  if (!solve_fn && hooks.append_frames && AppendApplies(wl.problem, !wl.relations.empty())) {
    // The device context is made over the first frame; every increment then
    // appends its frames to it (HipSession::Append).
    Problem first(run.map, run.estimated, {wl.scans[0]}, {}, wl.problem);
    (void)session.Load(first, solver);
  }
  size_t next_frame = 1;
  size_t increment = 0;
  while (next_frame < total_frames) {
    const size_t last = std::min(total_frames, next_frame + static_cast<size_t>(increment_size));
    for (size_t i = next_frame; i < last; ++i) {
      run.estimated.push_back(run.estimated.back().Compose(wl.odometry[i - 1]));
    }
    const size_t first_new = next_frame;
    next_frame = last;
    IncrementRecord rec;
    rec.frames = static_cast<int>(next_frame);
    auto t0 = std::chrono::steady_clock::now();
    if (!solve_fn && hooks.append_frames && session.HasContext() &&
        AppendApplies(wl.problem, !wl.relations.empty())) {
      // The context extends its own problem by the new frames, built on the
      // device (the same problem as the rebuild below).
      std::vector<const Scan*> new_scans;
      std::vector<Pose2> new_odometry;
      for (size_t i = first_new; i < last; ++i) {
        new_scans.push_back(&wl.scans[i]);
        new_odometry.push_back(wl.odometry[i - 1]);
      }
      session.Append(new_scans, new_odometry, wl.problem, run.map, run.estimated);
      rec.build_seconds = Seconds(t0);
      t0 = std::chrono::steady_clock::now();
      SolveResult step = session.SolveLoaded(run.map, run.estimated, solver);
      rec.solve_seconds = Seconds(t0);
      step.duration_seconds = rec.solve_seconds;
      Record(run, rec, step);
      ++increment;
      if (hooks.after_increment) hooks.after_increment(next_frame, increment, run.estimated, run.map);
      continue;
    }
    const std::vector<Scan> scans_so_far(wl.scans.begin(),
                                         wl.scans.begin() + static_cast<std::ptrdiff_t>(next_frame));
    const std::vector<Pose2> odom_so_far(wl.odometry.begin(),
                                         wl.odometry.begin() + static_cast<std::ptrdiff_t>(next_frame - 1));
    // The device runs build their node normals on the device too: the first
    // increment through a device call of its own, the later ones inside the
    // context update (computed at the start of the solve, no host round trip).
    const bool deferred = !solve_fn && session.HasContext() && DeviceNormalsApply(wl.problem);
    Problem problem(std::move(run.map), run.estimated, scans_so_far, odom_so_far, wl.problem, wl.relations,
                    deferred  ? NodeNormalsFn(DeferredNodeNormals{})
                    : solve_fn ? NodeNormalsFn()
                               : NodeNormalsFn(DeviceNodeNormals));
    rec.build_seconds = Seconds(t0);
    t0 = std::chrono::steady_clock::now();
    SolveResult step = solve_fn ? SolveWith(solve_fn, problem, solver)
                                : session.Solve(problem, solver, deferred ? &wl.problem.normals : nullptr);
    rec.solve_seconds = Seconds(t0);
    step.duration_seconds = rec.solve_seconds;
    run.estimated = problem.poses();
    run.map = problem.map();
    Record(run, rec, step);
    ++increment;
    if (hooks.after_increment) hooks.after_increment(next_frame, increment, run.estimated, run.map);
  }
  return run;
}

TrajectoryMetrics Evaluate(const std::vector<Pose2>& estimated, const std::vector<Pose2>& ground_truth) {
  if (estimated.size() != ground_truth.size() || estimated.size() < 2) {
    throw std::runtime_error("pose count mismatch between estimate and ground truth");
  }
  TrajectoryMetrics m;
  for (size_t i = 0; i < estimated.size(); ++i) {
    m.mean_translation_error += (estimated[i].Translation() - ground_truth[i].Translation()).norm();
    m.mean_rotation_error += std::abs(WrapAngle(estimated[i].theta - ground_truth[i].theta));
  }
  m.mean_translation_error /= static_cast<double>(estimated.size());
  m.mean_rotation_error /= static_cast<double>(estimated.size());
  const size_t pairs = estimated.size() - 1;
  for (size_t i = 0; i < pairs; ++i) {
    const Pose2 rel_est = estimated[i].RelativeTo(estimated[i + 1]);
    const Pose2 rel_gt = ground_truth[i].RelativeTo(ground_truth[i + 1]);
    m.mean_relative_translation_error += (rel_est.Translation() - rel_gt.Translation()).norm();
    m.mean_relative_rotation_error += std::abs(WrapAngle(rel_est.theta - rel_gt.theta));
  }
  m.mean_relative_translation_error /= static_cast<double>(pairs);
  m.mean_relative_rotation_error /= static_cast<double>(pairs);
  return m;
}

namespace {

std::ofstream OpenOutput(const std::string& path) {
  std::ofstream file(path);
  if (!file.is_open()) throw std::runtime_error("cannot write " + path);
  return file;
}

}  // namespace

void WritePoses(const std::string& path, const std::vector<Pose2>& poses) {
  std::ofstream file = OpenOutput(path);
  file << std::setprecision(17) << "x,y,theta\n";
  for (const Pose2& q : poses) file << q.x << "," << q.y << "," << q.theta << "\n";
}

void WriteMap(const std::string& path, const GridMap& map) {
  std::ofstream file = OpenOutput(path);
  file << std::setprecision(17) << map.nx() << " " << map.ny() << " " << map.min_corner().x << " "
       << map.min_corner().y << " " << map.dx() << " " << map.dy() << "\n";
  for (int h = 0; h < map.ny(); ++h) {
    for (int w = 0; w < map.nx(); ++w) {
      file << map.values()[static_cast<size_t>(map.NodeId(w, h))] << (w + 1 < map.nx() ? " " : "\n");
    }
  }
}

void WriteGlobalScanPoints(const std::string& path, const std::vector<Scan>& scans,
                           const std::vector<Pose2>& poses) {
  std::ofstream file = OpenOutput(path);
  file << std::setprecision(9) << "x,y\n";
  for (size_t i = 0; i < scans.size() && i < poses.size(); ++i) {
    for (const Vec2& p : scans[i].points) {
      const Vec2 g = poses[i].Apply(p);
      file << g.x << "," << g.y << "\n";
    }
  }
}

void WriteMetricsJson(const std::string& path, const SolveResult& result, const TrajectoryMetrics* metrics) {
  std::ofstream file = OpenOutput(path);
  file << std::setprecision(17) << "{\n";
  file << "  \"converged\": " << (result.converged ? "true" : "false") << ",\n";
  file << "  \"iterations\": " << result.iterations << ",\n";
  file << "  \"initial_cost\": " << result.initial_cost << ",\n";
  file << "  \"final_cost\": " << result.final_cost << ",\n";
  file << "  \"duration_seconds\": " << result.duration_seconds;
  if (metrics != nullptr) {
    file << ",\n";
    file << "  \"mean_translation_error\": " << metrics->mean_translation_error << ",\n";
    file << "  \"mean_rotation_error\": " << metrics->mean_rotation_error << ",\n";
    file << "  \"mean_relative_translation_error\": " << metrics->mean_relative_translation_error << ",\n";
    file << "  \"mean_relative_rotation_error\": " << metrics->mean_relative_rotation_error << "\n";
  } else {
    file << "\n";
  }
  file << "}\n";
}

IncrementalRun RunWorkload(const Workload& wl, const SolverOptions& solver, const RunOptions& options,
                           const IncrementSolveFn& solve_fn) {
  if (wl.scans.size() < 2) throw std::runtime_error("dataset needs at least two scans");
  namespace fs = std::filesystem;
  const bool write = !options.output_dir.empty();
  IncrementalRun run;
  size_t frames = wl.scans.size();
  if (!options.incremental) {
    // runner.cpp:209-225.
    std::vector<Pose2> initial = wl.initial.empty() ? wl.dataset_poses : wl.initial;
    GridMap map = wl.map;
    if (options.init_from_scans) InitializeFromScans(map, wl.scans, initial, options.init_signed);
    Problem problem(std::move(map), initial, wl.scans, wl.odometry, wl.problem, wl.relations);
    SolverOptions o = solver;
    o.backend = SolverBackend::kHip;
    run.total = solve_fn ? SolveWith(solve_fn, problem, o) : Solve(problem, o);
    run.estimated = problem.poses();
    run.map = problem.map();
  } else {
    // runner.cpp:226-285.
    IncrementHooks hooks;
    hooks.init_from_scans = options.init_from_scans;
    hooks.init_signed = options.init_signed;
    const fs::path snapshot_dir = fs::path(options.output_dir) / "snapshots";
    if (write && options.snapshot_every > 0) {
      fs::create_directories(snapshot_dir);
      const size_t last_frame =
          options.max_frames > 0 ? std::min(frames, static_cast<size_t>(options.max_frames)) : frames;
      hooks.after_increment = [&, last_frame](size_t next_frame, size_t increment, const std::vector<Pose2>& est,
                                              const GridMap& map) {
        if (increment % static_cast<size_t>(options.snapshot_every) == 0 || next_frame == last_frame) {
          // File names carry the number of frames in the snapshot.
          char tag[32];
          std::snprintf(tag, sizeof(tag), "%06zu", next_frame);
          WriteMap((snapshot_dir / ("map_" + std::string(tag) + ".txt")).string(), map);
          WritePoses((snapshot_dir / ("poses_" + std::string(tag) + ".csv")).string(), est);
        }
      };
    }
    run = RunIncremental(wl, solver, options.increment_size, options.max_frames, solve_fn, hooks);
    frames = run.estimated.size();
  }
  if (!write) return run;
  const std::vector<Scan> scans(wl.scans.begin(), wl.scans.begin() + static_cast<std::ptrdiff_t>(frames));
  const std::vector<Pose2> dataset(wl.dataset_poses.begin(),
                                   wl.dataset_poses.begin() + static_cast<std::ptrdiff_t>(frames));
  const fs::path out(options.output_dir);
  fs::create_directories(out);
  WritePoses((out / "poses_odometry.csv").string(), dataset);
  WritePoses((out / "poses_estimated.csv").string(), run.estimated);
  WriteMap((out / "map.txt").string(), run.map);
  WriteGlobalScanPoints((out / "scan_points_global.csv").string(), scans, run.estimated);
  // runner.cpp:114-135: scan points outside the map domain are unanchored.
  size_t outside = 0;
  size_t total = 0;
  GridMap::CellRef cell;
  for (size_t i = 0; i < scans.size(); ++i) {
    for (const Vec2& p : scans[i].points) {
      if (!run.map.Locate(run.estimated[i].Apply(p), cell)) ++outside;
      ++total;
    }
  }
  if (outside > 0) {
    std::cerr << "WARNING: " << outside << " of " << total << " scan points ("
              << 100.0 * static_cast<double>(outside) / static_cast<double>(total)
              << "%) fall outside the map domain and are unanchored (SDF 0, no gradient).\n";
  }
  if (!wl.ground_truth.empty()) {
    const std::vector<Pose2> gt(wl.ground_truth.begin(),
                                wl.ground_truth.begin() + static_cast<std::ptrdiff_t>(frames));
    WritePoses((out / "poses_ground_truth.csv").string(), gt);
    const TrajectoryMetrics metrics = Evaluate(run.estimated, gt);
    WriteMetricsJson((out / "metrics.json").string(), run.total, &metrics);
  } else {
    WriteMetricsJson((out / "metrics.json").string(), run.total, nullptr);
  }
  return run;
}

}  // namespace sdf_slam
