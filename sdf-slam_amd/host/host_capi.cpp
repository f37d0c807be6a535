// C API of the host library (include/sdfslam_host.h).
#include <algorithm>
#include <cstring>
#include <exception>
#include <fstream>
#include <memory>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>

#include "problem.hpp"
#include "runner.hpp"
#include "scene.hpp"
#include "server.hpp"
#include "sdfslam_host.h"
#include "solver.hpp"

using namespace sdf_slam;

struct sdfslam_host_problem {
  std::unique_ptr<Problem> problem;
  Workload workload;  // empty for problems built from arrays
  bool has_workload{false};
};

struct sdfslam_host_dataset {
  Workload workload;
};

struct sdfslam_server {
  std::unique_ptr<Server> server;
  std::vector<GridMap> maps;  // grid of each session, for state queries
};

namespace {

thread_local std::string g_error;

std::vector<Vec2> ScanFrom(const double* xy, int64_t n) {
  if (n < 0 || (n > 0 && xy == nullptr)) throw std::invalid_argument("bad scan arrays");
  std::vector<Vec2> pts(static_cast<size_t>(n));
  for (int64_t i = 0; i < n; ++i) pts[static_cast<size_t>(i)] = {xy[2 * i], xy[2 * i + 1]};
  return pts;
}

int Fail(int code, const std::string& message) {
  g_error = message;
  return code;
}

template <typename F>
int Guard(F&& fn) {
  try {
    return fn();
  } catch (const std::invalid_argument& e) {
    return Fail(SDFSLAM_ERR_INVALID_ARGUMENT, e.what());
  } catch (const std::logic_error& e) {
    return Fail(SDFSLAM_ERR_INVALID_ARGUMENT, e.what());
  } catch (const std::bad_alloc& e) {
    return Fail(SDFSLAM_ERR_OOM, e.what());
  } catch (const std::runtime_error& e) {
    const std::string what = e.what();
    if (what == "sparse Cholesky factorization failed") {
      return Fail(SDFSLAM_ERR_NOT_SPD, what);
    }
    return Fail(SDFSLAM_ERR_INTERNAL, what);
  } catch (const std::exception& e) {
    return Fail(SDFSLAM_ERR_INTERNAL, e.what());
  }
}

SolverOptions FromC(const sdfslam_solver_options& c) {
  SolverOptions o;
  o.max_iterations = c.max_iterations;
  o.step_tolerance = c.step_tolerance;
  o.lambda_init = c.lambda_init;
  o.lambda_factor = c.lambda_factor;
  o.reject_worse_steps = c.reject_worse_steps != 0;
  o.marquardt_scaling = c.marquardt_scaling != 0;
  o.trust_region = c.trust_region != 0;
  o.num_threads = c.num_threads;
  o.linear_solver = static_cast<LinearSolver>(c.linear_solver);
  return o;
}

void ToC(const SolveResult& r, sdfslam_solve_result* result) {
  result->converged = r.converged ? 1 : 0;
  result->iterations = r.iterations;
  result->initial_cost = r.initial_cost;
  result->final_cost = r.final_cost;
  result->duration_seconds = r.duration_seconds;
  result->cost_history_length = static_cast<int32_t>(r.cost_history.size());
  if (result->cost_history != nullptr) {
    const size_t n = std::min(r.cost_history.size(),
                              static_cast<size_t>(std::max(0, result->cost_history_capacity)));
    std::copy(r.cost_history.begin(), r.cost_history.begin() + static_cast<std::ptrdiff_t>(n),
              result->cost_history);
  }
}

void ToCProblem(const ProblemOptions& d, sdfslam_problem_options* out) {
  out->weight_scan = d.weights.scan;
  out->weight_hallucination = d.weights.hallucination;
  out->weight_eikonal = d.weights.eikonal;
  out->weight_odometry = d.weights.odometry;
  out->weight_relation = d.weights.relation;
  out->hallucination_points_per_scan_point = d.hallucination.points_per_scan_point;
  out->hallucination_step_size = d.hallucination.step_size;
  out->hallucination_both_directions = d.hallucination.both_directions ? 1 : 0;
  out->normals_method = static_cast<int32_t>(d.normals.method);
  out->normals_k_neighbors = d.normals.k_neighbors;
  out->normals_per_scan = d.normals.per_scan ? 1 : 0;
  out->normals_corner_threshold = d.normals.corner_threshold;
  out->huber_delta = d.huber_delta;
  out->smooth_gradient = d.smooth_gradient ? 1 : 0;
  out->smooth_gradient_step = d.smooth_gradient_step;
  out->active_region = d.active_region ? 1 : 0;
  out->active_margin = d.active_margin;
}

ProblemOptions FromC(const sdfslam_problem_options& o) {
  ProblemOptions p;
  p.weights.scan = o.weight_scan;
  p.weights.hallucination = o.weight_hallucination;
  p.weights.eikonal = o.weight_eikonal;
  p.weights.odometry = o.weight_odometry;
  p.weights.relation = o.weight_relation;
  p.hallucination.points_per_scan_point = o.hallucination_points_per_scan_point;
  p.hallucination.step_size = o.hallucination_step_size;
  p.hallucination.both_directions = o.hallucination_both_directions != 0;
  p.normals.method = static_cast<NormalMethod>(o.normals_method);
  p.normals.k_neighbors = o.normals_k_neighbors;
  p.normals.per_scan = o.normals_per_scan != 0;
  p.normals.corner_threshold = o.normals_corner_threshold;
  p.huber_delta = o.huber_delta;
  p.smooth_gradient = o.smooth_gradient != 0;
  p.smooth_gradient_step = o.smooth_gradient_step;
  p.active_region = o.active_region != 0;
  p.active_margin = o.active_margin;
  return p;
}

std::string Trim(const std::string& s) {
  const size_t a = s.find_first_not_of(" \t\r");
  if (a == std::string::npos) {
    return "";
  }
  const size_t b = s.find_last_not_of(" \t\r");
  return s.substr(a, b - a + 1);
}

bool ParseBool(const std::string& v) {
  if (v == "true" || v == "True" || v == "1") {
    return true;
  }
  if (v == "false" || v == "False" || v == "0") {
    return false;
  }
  throw std::runtime_error("bad boolean '" + v + "'");
}

}  // namespace

extern "C" {

const char* sdfslam_host_last_error(void) { return g_error.c_str(); }

void sdfslam_host_default_problem_options(sdfslam_problem_options* out) {
  ToCProblem(ProblemOptions{}, out);
}

void sdfslam_host_default_solver_options(sdfslam_solver_options* out) {
  *out = ToC(SolverOptions{});
}

int sdfslam_host_problem_create(const sdfslam_grid* grid, const double* map_values,
                                int32_t num_frames, const double* poses,
                                const int64_t* scan_offsets, const double* scan_xy,
                                const double* odometry, int32_t num_relations,
                                const int32_t* rel_frames, const double* rel_meas,
                                const double* rel_sqrt_weight,
                                const sdfslam_problem_options* options,
                                sdfslam_host_problem** out) {
  return Guard([&]() -> int {
    if (grid == nullptr || poses == nullptr || scan_offsets == nullptr || options == nullptr ||
        out == nullptr || num_frames < 1) {
      throw std::invalid_argument("null argument or no frames");
    }
    // GridMap takes min/max corners; dx = (max - min) / (n - 1).
    GridMap map(grid->nx, grid->ny, {grid->min_x, grid->min_y},
                {grid->min_x + grid->dx * (grid->nx - 1), grid->min_y + grid->dy * (grid->ny - 1)});
    if (map_values != nullptr) {
      std::copy(map_values, map_values + map.values().size(), map.values().begin());
    }
    std::vector<Pose2> pose_vec(static_cast<size_t>(num_frames));
    std::vector<Scan> scans(static_cast<size_t>(num_frames));
    for (int32_t f = 0; f < num_frames; ++f) {
      pose_vec[static_cast<size_t>(f)] = {poses[3 * f], poses[3 * f + 1], poses[3 * f + 2]};
      for (int64_t k = scan_offsets[f]; k < scan_offsets[f + 1]; ++k) {
        scans[static_cast<size_t>(f)].points.push_back({scan_xy[2 * k], scan_xy[2 * k + 1]});
      }
    }
    std::vector<Pose2> odo;
    if (odometry != nullptr) {
      for (int32_t f = 0; f + 1 < num_frames; ++f) {
        odo.push_back({odometry[3 * f], odometry[3 * f + 1], odometry[3 * f + 2]});
      }
    }
    std::vector<RelationMeasurement> rels;
    for (int32_t r = 0; r < num_relations; ++r) {
      RelationMeasurement m;
      m.frame_i = rel_frames[2 * r];
      m.frame_j = rel_frames[2 * r + 1];
      m.measurement = {rel_meas[3 * r], rel_meas[3 * r + 1], rel_meas[3 * r + 2]};
      m.sqrt_weight = rel_sqrt_weight != nullptr ? rel_sqrt_weight[r] : 1.0;
      rels.push_back(m);
    }
    auto holder = std::make_unique<sdfslam_host_problem>();
    holder->problem = std::make_unique<Problem>(std::move(map), std::move(pose_vec), scans, odo,
                                                FromC(*options), rels);
    *out = holder.release();
    return SDFSLAM_OK;
  });
}

int sdfslam_host_workload_create(const char* name, uint64_t seed, sdfslam_host_problem** out) {
  return Guard([&]() -> int {
    if (name == nullptr || out == nullptr) {
      throw std::invalid_argument("null argument");
    }
    auto holder = std::make_unique<sdfslam_host_problem>();
    holder->workload = MakeWorkload(name, seed);
    holder->has_workload = true;
    const Workload& wl = holder->workload;
    holder->problem = std::make_unique<Problem>(wl.map, wl.initial, wl.scans, wl.odometry,
                                                wl.problem, wl.relations);
    *out = holder.release();
    return SDFSLAM_OK;
  });
}

void sdfslam_host_problem_destroy(sdfslam_host_problem* problem) { delete problem; }

int sdfslam_host_node_normals(const double* points_xy, const double* origins_xy, int64_t num_points,
                              int32_t k_neighbors, const sdfslam_grid* grid, int32_t method,
                              double corner_threshold, double* node_normal_xy, double* node_scale) {
  return Guard([&]() -> int {
    if (grid == nullptr || node_normal_xy == nullptr || node_scale == nullptr || num_points < 0 ||
        (num_points > 0 && (points_xy == nullptr || origins_xy == nullptr))) {
      throw std::invalid_argument("null argument");
    }
    if (method < SDFSLAM_NORMALS_PCA || method > SDFSLAM_NORMALS_WEIGHTED) {
      throw std::invalid_argument("bad normal method");
    }
    std::vector<Vec2> points(static_cast<size_t>(num_points));
    std::vector<Vec2> origins(static_cast<size_t>(num_points));
    for (int64_t i = 0; i < num_points; ++i) {
      points[static_cast<size_t>(i)] = Vec2{points_xy[2 * i], points_xy[2 * i + 1]};
      origins[static_cast<size_t>(i)] = Vec2{origins_xy[2 * i], origins_xy[2 * i + 1]};
    }
    const GridMap map = GridMap::FromSpacing(grid->nx, grid->ny, Vec2{grid->min_x, grid->min_y}, grid->dx, grid->dy);
    NormalOptions options;
    options.method = method == SDFSLAM_NORMALS_PCA      ? NormalMethod::kPca
                     : method == SDFSLAM_NORMALS_HYBRID ? NormalMethod::kHybrid
                                                        : NormalMethod::kWeighted;
    options.k_neighbors = k_neighbors;
    options.corner_threshold = corner_threshold;
    const std::vector<NodeNormal> nn =
        AssignNodeNormals(map, ScanNormalsOf(std::move(points), origins, k_neighbors), options);
    for (size_t v = 0; v < nn.size(); ++v) {
      node_normal_xy[2 * v] = nn[v].normal.x;
      node_normal_xy[2 * v + 1] = nn[v].normal.y;
      node_scale[v] = nn[v].eikonal_scale;
    }
    return SDFSLAM_OK;
  });
}

int sdfslam_host_init_from_scans(const sdfslam_grid* grid, double* map_values, const double* scan_xy,
                                 const int64_t* scan_ptr, int32_t num_scans, const double* poses,
                                 int32_t num_poses, int32_t signed_distance) {
  return Guard([&]() -> int {
    if (grid == nullptr || map_values == nullptr || num_scans < 0 || num_poses < 0 ||
        (num_scans > 0 && (scan_ptr == nullptr || scan_xy == nullptr)) || (num_poses > 0 && poses == nullptr)) {
      throw std::invalid_argument("null argument");
    }
    GridMap map = GridMap::FromSpacing(grid->nx, grid->ny, Vec2{grid->min_x, grid->min_y}, grid->dx, grid->dy);
    const size_t nodes = static_cast<size_t>(map.num_nodes());
    std::copy(map_values, map_values + nodes, map.values().begin());
    std::vector<Scan> scans(static_cast<size_t>(num_scans));
    for (int32_t i = 0; i < num_scans; ++i) {
      if (scan_ptr[i + 1] < scan_ptr[i]) throw std::invalid_argument("scan_ptr is not ascending");
      for (int64_t k = scan_ptr[i]; k < scan_ptr[i + 1]; ++k) {
        scans[static_cast<size_t>(i)].points.push_back(Vec2{scan_xy[2 * k], scan_xy[2 * k + 1]});
      }
    }
    std::vector<Pose2> ps(static_cast<size_t>(num_poses));
    for (int32_t i = 0; i < num_poses; ++i) ps[static_cast<size_t>(i)] = {poses[3 * i], poses[3 * i + 1], poses[3 * i + 2]};
    InitializeFromScans(map, scans, ps, signed_distance != 0);
    std::copy(map.values().begin(), map.values().end(), map_values);
    return SDFSLAM_OK;
  });
}

int sdfslam_host_run(const char* workload, uint64_t seed, const sdfslam_run_options* run,
                     const sdfslam_solver_options* options, sdfslam_solve_fn solve_fn, void* user,
                     const char* output_dir, sdfslam_solve_result* total) {
  return Guard([&]() -> int {
    if (workload == nullptr || run == nullptr || total == nullptr) throw std::invalid_argument("null argument");
    const Workload wl = MakeWorkload(workload, seed);
    SolverOptions o = wl.solver;
    if (options != nullptr) {
      o = FromC(*options);
    } else if (run->incremental != 0) {
      o.max_iterations = wl.iterations_per_increment;
    }
    o.backend = SolverBackend::kHip;
    IncrementSolveFn fn;
    if (solve_fn != nullptr) {
      fn = [solve_fn, user](const sdfslam_problem_view* v, const sdfslam_solver_options* so, double* m,
                            double* p, sdfslam_solve_result* r) { return solve_fn(user, v, so, m, p, r); };
    }
    RunOptions ro;
    ro.incremental = run->incremental != 0;
    ro.increment_size = run->increment_size;
    ro.max_frames = run->max_frames;
    ro.snapshot_every = run->snapshot_every;
    ro.init_from_scans = run->init_from_scans != 0;
    ro.init_signed = run->init_signed != 0;
    if (output_dir != nullptr) ro.output_dir = output_dir;
    const IncrementalRun r = RunWorkload(wl, o, ro, fn);
    ToC(r.total, total);
    return SDFSLAM_OK;
  });
}

int sdfslam_host_problem_view(const sdfslam_host_problem* problem, sdfslam_problem_view* out) {
  if (problem == nullptr || out == nullptr) {
    return Fail(SDFSLAM_ERR_INVALID_ARGUMENT, "null argument");
  }
  *out = problem->problem->View();
  return SDFSLAM_OK;
}

int sdfslam_host_problem_get_state(const sdfslam_host_problem* problem, double* map_values,
                                   double* poses) {
  if (problem == nullptr) {
    return Fail(SDFSLAM_ERR_INVALID_ARGUMENT, "null argument");
  }
  std::vector<double> m;
  std::vector<double> p;
  problem->problem->GetState(m, p);
  if (map_values != nullptr) {
    std::copy(m.begin(), m.end(), map_values);
  }
  if (poses != nullptr) {
    std::copy(p.begin(), p.end(), poses);
  }
  return SDFSLAM_OK;
}

int sdfslam_host_problem_set_state(sdfslam_host_problem* problem, const double* map_values,
                                   const double* poses) {
  if (problem == nullptr) {
    return Fail(SDFSLAM_ERR_INVALID_ARGUMENT, "null argument");
  }
  problem->problem->SetState(map_values, poses);
  return SDFSLAM_OK;
}

int sdfslam_host_workload_info(const sdfslam_host_problem* problem,
                               sdfslam_solver_options* solver, int32_t* incremental,
                               int32_t* iterations_per_increment, double* ground_truth) {
  if (problem == nullptr || !problem->has_workload) {
    return Fail(SDFSLAM_ERR_INVALID_ARGUMENT, "not a synthetic workload");
  }
  const Workload& wl = problem->workload;
  if (solver != nullptr) {
    *solver = ToC(wl.solver);
  }
  if (incremental != nullptr) {
    *incremental = wl.incremental ? 1 : 0;
  }
  if (iterations_per_increment != nullptr) {
    *iterations_per_increment = wl.iterations_per_increment;
  }
  if (ground_truth != nullptr) {
    for (size_t i = 0; i < wl.ground_truth.size(); ++i) {
      ground_truth[3 * i] = wl.ground_truth[i].x;
      ground_truth[3 * i + 1] = wl.ground_truth[i].y;
      ground_truth[3 * i + 2] = wl.ground_truth[i].theta;
    }
  }
  return SDFSLAM_OK;
}

int sdfslam_host_solve(sdfslam_host_problem* problem, const char* backend,
                       const sdfslam_solver_options* options, sdfslam_solve_result* result) {
  return Guard([&]() -> int {
    if (problem == nullptr || options == nullptr || result == nullptr || backend == nullptr) {
      throw std::invalid_argument("null argument");
    }
    SolverOptions o = FromC(*options);
    const std::string b = backend;
    if (b == "hip") {
      o.backend = SolverBackend::kHip;
    } else if (b == "lm") {
      o.backend = SolverBackend::kLm;
    } else if (b == "ceres") {
      o.backend = SolverBackend::kCeres;
    } else {
      throw std::runtime_error("unknown solver.backend '" + b + "'");
    }
    const SolveResult r = Solve(*problem->problem, o);
    ToC(r, result);
    return SDFSLAM_OK;
  });
}

int sdfslam_host_run_incremental(const char* workload, uint64_t seed,
                                 const sdfslam_solver_options* options, int32_t increment_size,
                                 int32_t max_frames, sdfslam_solve_fn solve_fn, void* user,
                                 sdfslam_solve_result* total, sdfslam_increment_record* records,
                                 int32_t records_capacity, int32_t* num_records, double* poses_out,
                                 double* map_out) {
  return Guard([&]() -> int {
    if (workload == nullptr || total == nullptr) throw std::invalid_argument("null argument");
    const Workload wl = MakeWorkload(workload, seed);
    SolverOptions o = wl.solver;
    if (options != nullptr) {
      o = FromC(*options);
    } else {
      o.max_iterations = wl.iterations_per_increment;
    }
    o.backend = SolverBackend::kHip;
    IncrementSolveFn fn;
    if (solve_fn != nullptr) {
      fn = [solve_fn, user](const sdfslam_problem_view* v, const sdfslam_solver_options* so, double* m,
                            double* p, sdfslam_solve_result* r) { return solve_fn(user, v, so, m, p, r); };
    }
    const IncrementalRun run = RunIncremental(wl, o, increment_size, max_frames, fn);
    ToC(run.total, total);
    const auto count = static_cast<int32_t>(run.increments.size());
    if (num_records != nullptr) *num_records = count;
    for (int32_t k = 0; records != nullptr && k < std::min(count, records_capacity); ++k) {
      const IncrementRecord& r = run.increments[static_cast<size_t>(k)];
      records[k] = {r.frames, r.iterations, r.initial_cost, r.final_cost, r.build_seconds, r.solve_seconds};
    }
    if (poses_out != nullptr) {
      for (size_t i = 0; i < run.estimated.size(); ++i) {
        poses_out[3 * i] = run.estimated[i].x;
        poses_out[3 * i + 1] = run.estimated[i].y;
        poses_out[3 * i + 2] = run.estimated[i].theta;
      }
    }
    if (map_out != nullptr) {
      std::copy(run.map.values().begin(), run.map.values().end(), map_out);
    }
    return SDFSLAM_OK;
  });
}

int sdfslam_host_parse_solver_yaml(const char* path, sdfslam_solver_options* out, char* backend,
                                   int32_t backend_capacity) {
  return Guard([&]() -> int {
    if (path == nullptr || out == nullptr || backend == nullptr || backend_capacity < 8) {
      throw std::invalid_argument("null argument or backend buffer below 8 bytes");
    }
    std::ifstream file(path);
    if (!file.is_open()) {
      throw std::runtime_error(std::string("cannot open config: ") + path);
    }
    SolverOptions o;
    std::string backend_name = "lm";
    // config.cpp:171-175 plus this backend's selection value.
    const std::set<std::string> known = {"backend",          "max_iterations",     "step_tolerance",
                                         "lambda_init",      "lambda_factor",      "reject_worse_steps",
                                         "marquardt_scaling", "trust_region",      "num_threads",
                                         "linear_solver"};
    std::string line;
    bool in_solver = false;
    while (std::getline(file, line)) {
      const size_t hash = line.find('#');
      if (hash != std::string::npos) {
        line = line.substr(0, hash);
      }
      if (Trim(line).empty()) {
        continue;
      }
      const bool indented = line[0] == ' ' || line[0] == '\t';
      const std::string body = Trim(line);
      const size_t colon = body.find(':');
      if (colon == std::string::npos) {
        throw std::runtime_error("malformed config line: " + body);
      }
      const std::string key = Trim(body.substr(0, colon));
      const std::string value = Trim(body.substr(colon + 1));
      if (!indented) {
        in_solver = key == "solver" && value.empty();
        continue;
      }
      if (!in_solver) {
        continue;
      }
      if (known.count(key) == 0) {
        throw std::runtime_error("unknown config key 'solver." + key + "'");
      }
      if (key == "backend") {
        if (value != "lm" && value != "ceres" && value != "hip") {
          throw std::runtime_error("unknown solver.backend '" + value + "'");
        }
        backend_name = value;
      } else if (key == "linear_solver") {
        if (value == "eigen") {
          o.linear_solver = LinearSolver::kEigen;
        } else if (value == "cholmod") {
          o.linear_solver = LinearSolver::kCholmod;
        } else {
          throw std::runtime_error("unknown solver.linear_solver '" + value + "'");
        }
      } else if (key == "max_iterations") {
        o.max_iterations = std::stoi(value);
      } else if (key == "step_tolerance") {
        o.step_tolerance = std::stod(value);
      } else if (key == "lambda_init") {
        o.lambda_init = std::stod(value);
      } else if (key == "lambda_factor") {
        o.lambda_factor = std::stod(value);
      } else if (key == "reject_worse_steps") {
        o.reject_worse_steps = ParseBool(value);
      } else if (key == "marquardt_scaling") {
        o.marquardt_scaling = ParseBool(value);
      } else if (key == "trust_region") {
        o.trust_region = ParseBool(value);
      } else if (key == "num_threads") {
        o.num_threads = std::stoi(value);
      }
    }
    *out = ToC(o);
    std::strncpy(backend, backend_name.c_str(), static_cast<size_t>(backend_capacity) - 1);
    backend[backend_capacity - 1] = '\0';
    return SDFSLAM_OK;
  });
}

int sdfslam_host_dataset_create(const char* workload, uint64_t seed, sdfslam_host_dataset** out) {
  return Guard([&]() -> int {
    if (workload == nullptr || out == nullptr) throw std::invalid_argument("null argument");
    auto d = std::make_unique<sdfslam_host_dataset>();
    d->workload = MakeWorkload(workload, seed);
    *out = d.release();
    return SDFSLAM_OK;
  });
}

void sdfslam_host_dataset_destroy(sdfslam_host_dataset* dataset) { delete dataset; }

int sdfslam_host_dataset_info(const sdfslam_host_dataset* dataset, int32_t* num_frames,
                              sdfslam_grid* grid, sdfslam_problem_options* problem,
                              sdfslam_solver_options* step) {
  return Guard([&]() -> int {
    if (dataset == nullptr) throw std::invalid_argument("null argument");
    const Workload& wl = dataset->workload;
    if (num_frames != nullptr) *num_frames = static_cast<int32_t>(wl.scans.size());
    if (grid != nullptr) {
      *grid = {wl.map.nx(), wl.map.ny(), wl.map.min_corner().x, wl.map.min_corner().y, wl.map.dx(),
               wl.map.dy()};
    }
    if (problem != nullptr) ToCProblem(wl.problem, problem);
    if (step != nullptr) {
      SolverOptions o = wl.solver;
      o.max_iterations = wl.iterations_per_increment;
      *step = ToC(o);
    }
    return SDFSLAM_OK;
  });
}

int sdfslam_host_dataset_frame(const sdfslam_host_dataset* dataset, int32_t frame, double* scan_xy,
                               int64_t capacity, int64_t* num_points, double* pose,
                               double* odometry) {
  return Guard([&]() -> int {
    if (dataset == nullptr) throw std::invalid_argument("null argument");
    const Workload& wl = dataset->workload;
    if (frame < 0 || frame >= static_cast<int32_t>(wl.scans.size())) {
      throw std::invalid_argument("frame out of range");
    }
    const auto& pts = wl.scans[static_cast<size_t>(frame)].points;
    if (num_points != nullptr) *num_points = static_cast<int64_t>(pts.size());
    if (scan_xy != nullptr) {
      const size_t n = std::min(pts.size(), static_cast<size_t>(std::max<int64_t>(capacity, 0)));
      for (size_t i = 0; i < n; ++i) {
        scan_xy[2 * i] = pts[i].x;
        scan_xy[2 * i + 1] = pts[i].y;
      }
    }
    if (pose != nullptr) {
      const Pose2& p = wl.dataset_poses[static_cast<size_t>(frame)];
      pose[0] = p.x;
      pose[1] = p.y;
      pose[2] = p.theta;
    }
    if (odometry != nullptr) {
      const Pose2 o = frame == 0 ? Pose2{0.0, 0.0, 0.0} : wl.odometry[static_cast<size_t>(frame) - 1];
      odometry[0] = o.x;
      odometry[1] = o.y;
      odometry[2] = o.theta;
    }
    return SDFSLAM_OK;
  });
}

int sdfslam_server_create(const sdfslam_server_options* options, sdfslam_server** out) {
  return Guard([&]() -> int {
    if (options == nullptr || out == nullptr) throw std::invalid_argument("null argument");
    ServerOptions o;
    o.workers = options->workers;
    o.max_batch = options->max_batch;
    o.pipelines = options->pipelines;
    o.device_budget_bytes = options->device_budget_bytes;
    o.max_in_flight = options->max_in_flight;
    o.device_normals = options->device_normals != 0;
    auto s = std::make_unique<sdfslam_server>();
    s->server = std::make_unique<Server>(o);
    *out = s.release();
    return SDFSLAM_OK;
  });
}

void sdfslam_server_destroy(sdfslam_server* server) { delete server; }

int sdfslam_server_open_session(sdfslam_server* server, const sdfslam_grid* grid,
                                const double* map_values, const sdfslam_problem_options* problem,
                                const sdfslam_solver_options* step, const double* first_pose,
                                const double* scan_xy, int64_t num_points, int32_t* session) {
  return Guard([&]() -> int {
    if (server == nullptr || grid == nullptr || problem == nullptr || step == nullptr ||
        first_pose == nullptr || session == nullptr) {
      throw std::invalid_argument("null argument");
    }
    if (grid->nx < 2 || grid->ny < 2) throw std::invalid_argument("grid needs nx, ny >= 2");
    GridMap map(grid->nx, grid->ny, {grid->min_x, grid->min_y},
                {grid->min_x + grid->dx * (grid->nx - 1), grid->min_y + grid->dy * (grid->ny - 1)}, 0.0);
    if (map_values != nullptr) {
      std::copy(map_values, map_values + map.num_nodes(), map.values().begin());
    }
    Scan first;
    first.points = ScanFrom(scan_xy, num_points);
    try {
      *session = server->server->OpenSession(map, FromC(*problem), FromC(*step),
                                             {first_pose[0], first_pose[1], first_pose[2]}, first);
    } catch (const std::runtime_error& e) {
      if (std::string(e.what()).rfind("admission refused", 0) == 0) {
        return Fail(SDFSLAM_ERR_OOM, e.what());
      }
      throw;
    }
    return SDFSLAM_OK;
  });
}

int sdfslam_server_submit(sdfslam_server* server, int32_t session, const double* scan_xy,
                          int64_t num_points, const double* odometry, int64_t* ticket) {
  return Guard([&]() -> int {
    if (server == nullptr || odometry == nullptr || ticket == nullptr) {
      throw std::invalid_argument("null argument");
    }
    Scan scan;
    scan.points = ScanFrom(scan_xy, num_points);
    *ticket = server->server->Submit(session, scan, {odometry[0], odometry[1], odometry[2]},
                                     std::chrono::steady_clock::now());
    return SDFSLAM_OK;
  });
}

int sdfslam_server_wait(sdfslam_server* server, int32_t session, int64_t ticket,
                        sdfslam_increment_record* record, double* latency_seconds) {
  return Guard([&]() -> int {
    if (server == nullptr) throw std::invalid_argument("null argument");
    const IncrementRecord r = server->server->Wait(session, ticket);
    if (record != nullptr) {
      *record = {r.frames, r.iterations, r.initial_cost, r.final_cost, r.build_seconds, r.solve_seconds};
    }
    if (latency_seconds != nullptr) *latency_seconds = server->server->Latency(session, ticket);
    return SDFSLAM_OK;
  });
}

int sdfslam_server_close_session(sdfslam_server* server, int32_t session) {
  return Guard([&]() -> int {
    if (server == nullptr) throw std::invalid_argument("null argument");
    server->server->CloseSession(session);
    return SDFSLAM_OK;
  });
}

int sdfslam_server_session_info(sdfslam_server* server, int32_t session, int32_t* admitted,
                                int32_t* closed, int64_t* device_bytes, int64_t* submitted,
                                int64_t* finished) {
  return Guard([&]() -> int {
    if (server == nullptr) throw std::invalid_argument("null argument");
    const Server::SessionInfo i = server->server->Info(session);
    if (admitted != nullptr) *admitted = i.admitted ? 1 : 0;
    if (closed != nullptr) *closed = i.closed ? 1 : 0;
    if (device_bytes != nullptr) *device_bytes = i.device_bytes;
    if (submitted != nullptr) *submitted = i.submitted;
    if (finished != nullptr) *finished = i.finished;
    return SDFSLAM_OK;
  });
}

int sdfslam_server_state(sdfslam_server* server, int32_t session, int32_t capacity,
                         int32_t* frames, double* poses, double* map_values) {
  return Guard([&]() -> int {
    if (server == nullptr) throw std::invalid_argument("null argument");
    std::vector<Pose2> est;
    GridMap map;
    server->server->State(session, est, map);
    if (frames != nullptr) *frames = static_cast<int32_t>(est.size());
    if (poses != nullptr) {
      const size_t n = std::min(est.size(), static_cast<size_t>(std::max(capacity, 0)));
      for (size_t i = 0; i < n; ++i) {
        poses[3 * i] = est[i].x;
        poses[3 * i + 1] = est[i].y;
        poses[3 * i + 2] = est[i].theta;
      }
    }
    if (map_values != nullptr) std::copy(map.values().begin(), map.values().end(), map_values);
    return SDFSLAM_OK;
  });
}

int sdfslam_host_served_benchmark(const sdfslam_served_options* options, sdfslam_served_result* out) {
  return Guard([&]() -> int {
    if (options == nullptr || out == nullptr) throw std::invalid_argument("null argument");
    ServedBenchmark b;
    b.sessions = options->sessions;
    b.rate_hz = options->rate_hz;
    b.duration_s = options->duration_s;
    b.workers = options->workers;
    b.max_batch = options->max_batch;
    b.pipelines = options->pipelines;
    b.seed0 = options->seed0;
    b.max_in_flight = options->max_in_flight;
    b.device_normals = options->device_normals != 0;
    b.start_frame = std::max(options->start_frame, 1);
    const ServedResult r = RunServedBenchmark(b);
    *out = {r.increments,       r.lm_iterations,    r.wall_seconds,     r.increments_per_second,
            r.latency_p50_ms,   r.latency_p99_ms,   r.latency_mean_ms,  r.build_mean_ms,
            r.solve_mean_ms,    r.device_bytes,     r.batches,          r.mean_batch,
            r.prepare_batch_ms, r.solve_batch_ms,   r.collect_batch_ms, r.capture_batch_ms,
            r.analysis_batch_ms, r.epochs_rebuilt, r.latency_p50_after_first_ms,
            r.latency_p99_after_first_ms, r.first_increment_mean_ms, r.queue_p99_ms, r.prepare_p99_ms,
            r.solve_wait_p99_ms, r.solve_p99_ms};
    return SDFSLAM_OK;
  });
}

int sdfslam_host_batch_benchmark(const sdfslam_batch_options* options, sdfslam_batch_result* out) {
  return Guard([&]() -> int {
    if (options == nullptr || out == nullptr) throw std::invalid_argument("null argument");
    BatchBenchmark b;
    b.instances = options->instances;
    b.iterations = options->iterations;
    b.max_batch = options->max_batch;
    b.workers = options->workers;
    b.seed0 = options->seed0;
    const BatchResult r = RunBatchBenchmark(b);
    *out = {r.lm_iterations, r.unknowns, r.wall_seconds, r.lm_iterations_per_second};
    return SDFSLAM_OK;
  });
}

int sdfslam_host_batch_autotune(const sdfslam_batch_options* options, const int32_t* max_batches,
                                int32_t num_max_batches, const int32_t* workers, int32_t num_workers,
                                int32_t trial_iterations, sdfslam_batch_shape* trials,
                                sdfslam_batch_shape* best) {
  return Guard([&]() -> int {
    if (options == nullptr || best == nullptr || num_max_batches < 1 || num_workers < 1 || max_batches == nullptr ||
        workers == nullptr) {
      throw std::invalid_argument("null argument or empty candidate list");
    }
    BatchBenchmark b;
    b.instances = options->instances;
    b.seed0 = options->seed0;
    const std::vector<int> mb(max_batches, max_batches + num_max_batches);
    const std::vector<int> w(workers, workers + num_workers);
    const BatchTuneResult r = TuneBatchShapes(b, mb, w, trial_iterations);
    for (size_t i = 0; trials != nullptr && i < r.trials.size(); ++i) {
      trials[i] = {r.trials[i].max_batch, r.trials[i].workers, r.trials[i].lm_iterations_per_second};
    }
    *best = {r.best.max_batch, r.best.workers, r.best.lm_iterations_per_second};
    return SDFSLAM_OK;
  });
}

}  // extern "C"
