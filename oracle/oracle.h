/*
 * CPU oracle for sdf-slam's joint-LM hot path - TEST INFRASTRUCTURE ONLY.
 *
 * A plain C++ restatement of the reference's CPU path, used as the checker
 * for the device backend. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product (sdf-slam_amd/) never does.
 *
 * Each function cites the reference file:line it restates. Contraction
 * policy: built with -ffp-contract=off; the only fused operations are the
 * explicit std::fma calls the reference itself writes in its LDL^T row
 * kernel (src/solvers/parallel_ldlt.hpp:186, 196).
 *
 * Parity pins (what checks each function, in tests/test_oracle.py):
 *   oracle_wrap_angle/compose/relative  tests/test_pose.cpp known answers
 *   oracle_locate/interpolate           tests/test_grid_map.cpp known answers
 *   oracle_eval_point (exact gradient)  central finite differences (test_jacobians.cpp:33-76)
 *   oracle_eval_point (smooth modes)    closed form: exact on affine fields, incl. borders
 *   oracle_eval_eikonal, oracle_eval_odom  central finite differences (test_jacobians.cpp:104-153)
 *   oracle_assemble                     brute-force per-row restatement + FD of the cost
 *   oracle_normal_matrix/neg_atb        dense J^T J and -J^T r computed with numpy
 *   oracle_ldlt_solve                   dense solves (numpy), residual norms, singular -> error
 *   oracle_cost                         r^T r of the assembled residuals (Huber off);
 *                                       hand-computed Huber cost on a 3-row affine toy
 *   oracle_solve (LM control)           e2e thresholds (test_e2e.cpp:63-88), the
 *                                       trust-region monotonicity (test_e2e.cpp:168-184),
 *                                       and hand-computed fixed-factor lambda traces with
 *                                       and without reject_worse_steps (each iteration's
 *                                       state = ApplyStep(linear step at the traced lambda),
 *                                       reverts bitwise)
 *   oracle_eval_points_frozen           equals oracle_eval_point inside the cell
 *                                       (test_jacobians.cpp:78-102) + central FD off-cell
 *   oracle_init_from_scans              tests/test_map_init.cpp known answers, plus a
 *                                       vectorised numpy nearest-point search
 */
#ifndef SDFSLAM_ORACLE_H_
#define SDFSLAM_ORACLE_H_

#include <stdint.h>

#include "sdfslam_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* pose.hpp:9, 32-43 */
double oracle_wrap_angle(double angle);
void oracle_compose(const double a[3], const double rel[3], double out[3]);
void oracle_relative(const double a[3], const double b[3], double out[3]);

/* grid_map.hpp:58-70 (clamped = 0) and 75-83 (clamped = 1); returns 1 when
 * the point is inside (always 1 for clamped). */
int oracle_locate(const sdfslam_grid* grid, double x, double y, int clamped, int32_t* w,
                  int32_t* h, double* alpha, double* beta);
/* grid_map.hpp:86-93 */
double oracle_interpolate(const sdfslam_grid* grid, const double* values, double x, double y,
                          int clamped);

/* residual_math.hpp:75-106; returns valid (1) or 0 (out of domain). */
int oracle_eval_point(const sdfslam_grid* grid, const double* values, const double pose[3],
                      double px, double py, double delta, int smooth_gradient,
                      double smooth_gradient_step, double* residual, int32_t node_ids[4],
                      double d_nodes[4], double d_pose[3]);
/* residual_math.hpp:158-171 */
void oracle_eval_eikonal(const sdfslam_grid* grid, const double* values, int32_t w, int32_t h,
                         double nx, double ny, double* residual, int32_t node_ids[3],
                         double d_nodes[3]);
/* residual_math.hpp:182-205; jacobians row-major 3x3. */
void oracle_eval_odom(const double pose_i[3], const double pose_j[3], const double meas[3],
                      double residual[3], double d_pose_i[9], double d_pose_j[9]);

/* Sparse matrix owned by the oracle (compressed columns, sorted rows). */
typedef struct oracle_matrix oracle_matrix;
/* Copies a compressed-column matrix (rows sorted within each column). */
oracle_matrix* oracle_matrix_from_csc(int64_t rows, int64_t cols, const int64_t* col_ptr,
                                      const int32_t* row_idx, const double* values);
void oracle_matrix_free(oracle_matrix* m);
void oracle_matrix_shape(const oracle_matrix* m, int64_t* rows, int64_t* cols, int64_t* nnz);
void oracle_matrix_copy(const oracle_matrix* m, int64_t* col_ptr, int32_t* row_idx,
                        double* values);

/* Residual vector and J at the given state: AssemblePointRows,
 * AssembleEikonalRows, the odometry tail and setFromTriplets
 * (lm_solver.cpp:30-82, 275-313). residuals: num_residuals doubles. */
int oracle_assemble(const sdfslam_problem_view* problem, const double* map_values,
                    const double* poses, double* residuals, oracle_matrix** jacobian);
/* ParallelAtA (lm_solver.cpp:113-172): full symmetric J^T J. */
oracle_matrix* oracle_normal_matrix(const oracle_matrix* jacobian);
/* Damping with diagonal insertion (lm_solver.cpp:321-330), in place. */
void oracle_damp(oracle_matrix* a, double lambda, int marquardt_scaling);
/* ParallelNegAtb (lm_solver.cpp:178-197). */
void oracle_neg_atb(const oracle_matrix* jacobian, const double* residuals, double* rhs);

/* A geometric order (active grid nodes by recursive coordinate bisection,
 * then the pose columns; order[k] = column eliminated k-th). The solves use
 * oracle_amd_order instead. */
void oracle_order(const sdfslam_problem_view* problem, int32_t* order);
/* Up-looking simplicial LDL^T with the row kernel of parallel_ldlt.hpp:165-206
 * (rows over elimination subtrees on oracle_set_threads threads) and the
 * solve P^T L^-T D^-1 L^-1 P b. order may be NULL (natural order).
 * Returns SDFSLAM_OK or SDFSLAM_ERR_NOT_SPD (zero pivot, parallel_ldlt.hpp:205). */
int oracle_ldlt_solve(const oracle_matrix* a, const int32_t* order, const double* b, double* x,
                      int64_t* factor_nnz);

/* Problem::Cost (problem.cpp:132-156). */
double oracle_cost(const sdfslam_problem_view* problem, const double* map_values,
                   const double* poses);
/* Problem::ApplyStep (problem.cpp:117-130), in place. */
void oracle_apply_step(const sdfslam_problem_view* problem, const double* step,
                       double* map_values, double* poses);

/* One damped Gauss-Newton step at the given state (lm_solver.cpp:232-364),
 * with the approximate minimum degree order of its matrix. */
int oracle_linear_step(const sdfslam_problem_view* problem, const double* map_values,
                       const double* poses, double lambda, int marquardt_scaling, double* step);

/* SolveLm (lm_solver.cpp:201-434): map_values/poses updated in place. */
int oracle_solve(const sdfslam_problem_view* problem, const sdfslam_solver_options* options,
                 double* map_values, double* poses, sdfslam_solve_result* result);

/* oracle_solve that also copies the state at the end of iteration
 * record_iterations[q] (q < num_records) into record_map + q*nx*ny and
 * record_poses + q*3*num_poses (either may be NULL). An iteration count the
 * solve does not reach (convergence stop) is not written. */
int oracle_solve_recorded(const sdfslam_problem_view* problem, const sdfslam_solver_options* options,
                          double* map_values, double* poses, sdfslam_solve_result* result,
                          int32_t num_records, const int32_t* record_iterations, double* record_map,
                          double* record_poses);

/* EvalPointResidualFrozenCell (residual_math.hpp:113-147) for n point rows:
 * row k evaluates point (px[k], py[k]) of frame frame[k] (pose
 * poses[3*frame[k]..]) against the cell (cell_w[k], cell_h[k]) with the map's
 * values at that cell's four nodes: alpha, beta from the cell's corner node,
 * so a point outside the cell extrapolates its bilinear patch. Outputs per
 * row: residual, d_nodes[4k..4k+3], d_pose[3k..3k+2]. */
void oracle_eval_points_frozen(const sdfslam_grid* grid, const double* values, const double* poses,
                               int64_t n, const int32_t* frame, const double* px, const double* py,
                               const double* delta, const int32_t* cell_w, const int32_t* cell_h,
                               double* residual, double* d_nodes, double* d_pose);

/* InitializeFromScans (map_init.cpp:31-69): scan i (points
 * scan_xy[2*scan_ptr[i] .. 2*scan_ptr[i+1]), sensor frame) placed by pose i;
 * every node gets the distance to the nearest point by exhaustive search
 * (ties to the lower point index), negated when signed and the node is
 * farther from that point's sensor origin than the point. No points: the
 * map is left unchanged. */
void oracle_init_from_scans(const sdfslam_grid* grid, double* map_values, const double* scan_xy,
                            const int64_t* scan_ptr, int32_t num_scans, const double* poses,
                            int32_t num_poses, int signed_distance);

/* Host threads of the threaded phases (assembly chunks, J^T J column ranges,
 * -J^T r, LDL^T elimination subtrees), default 1; the results do not depend
 * on it (every phase gives the serial result bit for bit). */
void oracle_set_threads(int threads);
int oracle_threads(void);
/* Nonzeros of L below the diagonal and sum over columns of c^2 + 3c of the
 * last LDL^T factorization (any thread). */
void oracle_last_factor_stats(int64_t* factor_nnz, double* factor_flops);
/* The approximate minimum degree order the solves use for matrix a (order[k]
 * = the column eliminated k-th), oracle/ordering.cpp. */
void oracle_amd_order(const oracle_matrix* a, int32_t* order);

/* Per-phase wall time (seconds) of the last oracle_solve on this thread:
 * assembly, normal equations, factorization+solve, step+cost. */
void oracle_last_phase_seconds(double out[4]);

#ifdef __cplusplus
}
#endif

#endif /* SDFSLAM_ORACLE_H_ */
