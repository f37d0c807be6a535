// Solver entry point of the host mirror: the reference's Solve/SolverOptions/
// SolveResult (src/solvers/solver.hpp:10-70) with one more backend, kHip,
// which runs the joint LM hot path on an MI355X through the C ABI in
// include/sdfslam_hip.h.
//
// The reference's own kLm (threaded Eigen / CHOLMOD) and kCeres backends live
// in the reference build; this library carries only the device backend and
// raises std::logic_error for the others, as the reference does for an
// unknown backend (lm_solver.cpp:443).
#pragma once

#include <cstdint>
#include <vector>

#include "problem.hpp"

namespace sdf_slam {

enum class SolverBackend : std::uint8_t {
  kLm,
  kCeres,
  /// MI355X joint-LM backend: device assembly, fixed-pattern J^T J, supernodal
  /// LDL^T and the LM loop on the GPU.
  kHip,
};

enum class LinearSolver : std::uint8_t { kEigen, kCholmod };

struct SolverOptions {
  SolverBackend backend{SolverBackend::kLm};
  int max_iterations{100};
  double step_tolerance{1e-6};
  double lambda_init{1.0};
  double lambda_factor{1.0};
  bool reject_worse_steps{false};
  bool marquardt_scaling{false};
  bool trust_region{false};
  int num_threads{0};
  LinearSolver linear_solver{LinearSolver::kEigen};
};

struct SolveResult {
  bool converged{false};
  int iterations{0};
  double initial_cost{0.0};
  double final_cost{0.0};
  double duration_seconds{0.0};
  std::vector<double> cost_history{};
};

/// Minimizes the problem in place with the selected backend (solver.hpp:67).
SolveResult Solve(Problem& problem, const SolverOptions& options);

/// The device backend. Throws what SolveLm throws: std::invalid_argument for
/// malformed input, std::runtime_error("sparse Cholesky factorization failed")
/// for a zero pivot, std::runtime_error for device errors.
SolveResult SolveHip(Problem& problem, const SolverOptions& options);

/// C-ABI form of the options.
sdfslam_solver_options ToC(const SolverOptions& options);

}  // namespace sdf_slam
