// Solve dispatch and the SolveHip shim over the C ABI. The shim owns no
// numerics: it passes the Problem's structure-of-arrays view and state to
// libsdfslam_hip.so and maps status codes to the reference's exceptions
// (lm_solver.cpp:354-355, 360-361, 443; problem.cpp:118-120).
#include "solver.hpp"

#include <stdexcept>
#include <string>

namespace sdf_slam {

sdfslam_solver_options ToC(const SolverOptions& options) {
  sdfslam_solver_options o{};
  o.max_iterations = options.max_iterations;
  o.step_tolerance = options.step_tolerance;
  o.lambda_init = options.lambda_init;
  o.lambda_factor = options.lambda_factor;
  o.reject_worse_steps = options.reject_worse_steps ? 1 : 0;
  o.marquardt_scaling = options.marquardt_scaling ? 1 : 0;
  o.trust_region = options.trust_region ? 1 : 0;
  o.num_threads = options.num_threads;
  o.linear_solver = static_cast<int32_t>(options.linear_solver);
  return o;
}

SolveResult SolveHip(Problem& problem, const SolverOptions& options) {
  const sdfslam_problem_view view = problem.View();
  const sdfslam_solver_options copts = ToC(options);
  std::vector<double> map_values;
  std::vector<double> poses;
  problem.GetState(map_values, poses);

  SolveResult result;
  result.cost_history.assign(static_cast<size_t>(std::max(0, options.max_iterations)) + 1, 0.0);
  sdfslam_solve_result cres{};
  cres.cost_history = result.cost_history.data();
  cres.cost_history_capacity = static_cast<int32_t>(result.cost_history.size());
  const int status = sdfslam_hip_solve(&view, &copts, map_values.data(), poses.data(), &cres);
  if (status == SDFSLAM_ERR_INVALID_ARGUMENT) {
    throw std::invalid_argument(sdfslam_hip_last_error());
  }
  if (status == SDFSLAM_ERR_NOT_SPD) {
    throw std::runtime_error("sparse Cholesky factorization failed");
  }
  if (status != SDFSLAM_OK) {
    throw std::runtime_error(std::string(sdfslam_hip_status_string(status)) + ": " +
                             sdfslam_hip_last_error());
  }
  problem.SetState(map_values.data(), poses.data());
  result.converged = cres.converged != 0;
  result.iterations = cres.iterations;
  result.initial_cost = cres.initial_cost;
  result.final_cost = cres.final_cost;
  result.duration_seconds = cres.duration_seconds;
  result.cost_history.resize(static_cast<size_t>(
      std::min(cres.cost_history_length, cres.cost_history_capacity)));
  return result;
}

SolveResult Solve(Problem& problem, const SolverOptions& options) {
  switch (options.backend) {
    case SolverBackend::kHip:
      return SolveHip(problem, options);
    case SolverBackend::kLm:
    case SolverBackend::kCeres:
      throw std::logic_error(
          "solver backend not built into this library (only 'hip' is available here)");
  }
  throw std::logic_error("unknown solver backend");
}

}  // namespace sdf_slam
