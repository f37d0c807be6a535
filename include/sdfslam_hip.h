/*
 * sdfslam_hip.h - C ABI of the MI355X (gfx950) backend for sdf-slam's joint
 * Levenberg-Marquardt hot path (libsdfslam_hip.so).
 *
 * Plain C: pointers, sizes and status codes only. No exception crosses this
 * boundary; the C++ shim (sdf-slam_amd/host/solver.cpp, SolveHip) maps the
 * status codes back to the exception types the reference throws.
 *
 * Reference interfaces replaced (reference file:line):
 *   sdfslam_hip_solve            <- SolveResult Solve(Problem&, const SolverOptions&)
 *                                   src/solvers/solver.hpp:67 dispatching to
 *                                   SolveLm src/solvers/lm_solver.cpp:201-434
 *   sdfslam_problem_view         <- Problem getters src/problem/problem.hpp:104-142
 *                                   (map(), poses(), point_specs(), eikonal_specs(),
 *                                   odom_specs(), huber_delta(), smooth_gradient*(),
 *                                   NodeColumn(), PoseColumn())
 *   sdfslam_solver_options       <- SolverOptions src/solvers/solver.hpp:29-55
 *   sdfslam_solve_result         <- SolveResult src/solvers/solver.hpp:57-64
 *   stage entry points           <- the phases of one LM iteration,
 *                                   src/solvers/lm_solver.cpp:232-364
 */
#ifndef SDFSLAM_HIP_H_
#define SDFSLAM_HIP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. The C++ shim rethrows:
 *   SDFSLAM_ERR_INVALID_ARGUMENT -> std::invalid_argument (problem.cpp:118-120)
 *   SDFSLAM_ERR_NOT_SPD          -> std::runtime_error("sparse Cholesky factorization failed")
 *                                   (lm_solver.cpp:354-355, 360-361)
 *   everything else              -> std::runtime_error(sdfslam_hip_status_string(code)) */
enum sdfslam_status {
  SDFSLAM_OK = 0,
  SDFSLAM_ERR_INVALID_ARGUMENT = 1,
  SDFSLAM_ERR_NOT_SPD = 2,
  SDFSLAM_ERR_HIP = 3,
  SDFSLAM_ERR_OOM = 4,
  SDFSLAM_ERR_NO_DEVICE = 5,
  SDFSLAM_ERR_INTERNAL = 6
};

/* Grid geometry of GridMap (src/core/grid_map.hpp:11-33): node (w, h) has flat
 * id h * nx + w; node position = (min_x + w * dx, min_y + h * dy). */
typedef struct sdfslam_grid {
  int32_t nx;
  int32_t ny;
  double min_x;
  double min_y;
  double dx;
  double dy;
} sdfslam_grid;

/* Read-only view of a Problem (src/problem/problem.hpp:92-161) as structure of
 * arrays. Rows keep the reference order: point rows (spec order), eikonal
 * rows, then 3 rows per odometry/relation spec (lm_solver.cpp:209-211, 281).
 * State layout: [active nodes in node-id order; 3 per pose for frames >= 1]
 * (problem.hpp:125-139). */
typedef struct sdfslam_problem_view {
  sdfslam_grid grid;
  int32_t num_poses;
  int32_t num_active_nodes;
  const int32_t* node_to_active; /* nx*ny entries; -1 = constant node (NodeColumn) */

  int64_t num_points; /* PointSpec: problem.hpp:28-33 */
  const int32_t* point_frame;
  const double* point_x; /* point in the sensor frame */
  const double* point_y;
  const double* point_expected_sdf;
  const double* point_sqrt_weight;

  int64_t num_eikonal; /* EikonalSpec: problem.hpp:36-41 */
  const int32_t* eik_w;
  const int32_t* eik_h;
  const double* eik_nx;
  const double* eik_ny;
  const double* eik_sqrt_weight;

  int64_t num_odom; /* OdomSpec: problem.hpp:45-50 (odometry then relations) */
  const int32_t* odom_i;
  const int32_t* odom_j;
  const double* odom_dx;
  const double* odom_dy;
  const double* odom_dtheta;
  const double* odom_sqrt_weight;

  double huber_delta;          /* problem.hpp:70, 115-121 */
  int32_t smooth_gradient;     /* problem.hpp:76 */
  double smooth_gradient_step; /* problem.hpp:81 */
} sdfslam_problem_view;

/* SolverOptions (solver.hpp:29-55). num_threads and linear_solver are accepted
 * for symmetry with the CPU backends and do not change the device path. */
typedef struct sdfslam_solver_options {
  int32_t max_iterations;
  double step_tolerance;
  double lambda_init;
  double lambda_factor;
  int32_t reject_worse_steps;
  int32_t marquardt_scaling;
  int32_t trust_region;
  int32_t num_threads;
  int32_t linear_solver;
} sdfslam_solver_options;

/* SolveResult (solver.hpp:57-64). cost_history is caller-owned storage of
 * cost_history_capacity doubles; the backend writes iterations + 1 entries
 * (truncated to the capacity) and the full count to cost_history_length. */
typedef struct sdfslam_solve_result {
  int32_t converged;
  int32_t iterations;
  double initial_cost;
  double final_cost;
  double duration_seconds;
  double* cost_history;
  int32_t cost_history_capacity;
  int32_t cost_history_length;
} sdfslam_solve_result;

/* Free and total memory of the current HIP device (hipMemGetInfo), for
 * admission decisions of callers that keep many contexts. */
int sdfslam_hip_device_memory(int64_t* free_bytes, int64_t* total_bytes);

/* Human-readable text for a status code (static storage). */
const char* sdfslam_hip_status_string(int status);

/* Message of the last error on the calling thread (static per thread). */
const char* sdfslam_hip_last_error(void);

/* Minimizes the problem in place (Solve/SolveLm, solver.hpp:67,
 * lm_solver.cpp:201-434). map_values: nx*ny doubles (only active nodes
 * change); poses: 3*num_poses doubles (x, y, theta; frame 0 is fixed). */
int sdfslam_hip_solve(const sdfslam_problem_view* problem, const sdfslam_solver_options* options,
                      double* map_values, double* poses, sdfslam_solve_result* result);

/* ---- Stage entry points (device-resident context) --------------------------
 * A context holds the problem, its state, the symbolic analysis and the
 * workspaces on the current HIP device. The stage calls expose the phases of
 * one LM iteration so each can be compared against the CPU path. */
typedef struct sdfslam_hip_context sdfslam_hip_context;

int sdfslam_hip_context_create(const sdfslam_problem_view* problem,
                               const sdfslam_solver_options* options, const double* map_values,
                               const double* poses, sdfslam_hip_context** out);
int sdfslam_hip_context_destroy(sdfslam_hip_context* ctx);
/* Frees the device and pinned host buffers a context outgrew (its pools keep
 * them when a layout grows, because freeing device memory waits for the
 * whole device). No counterpart in the reference; a server calls it when it
 * needs the memory. Waits for the device; the context stays usable. */
int sdfslam_hip_context_trim(sdfslam_hip_context* ctx);
/* A context update whose problem's eikonal normals are computed on the
 * device, at the start of the context's next call, instead of taken from
 * the view: the per-increment normal estimation of the problem build
 * (ComputeScanNormals over all scans and AssignNodeNormals,
 * src/core/normals.cpp:128-220, called by Problem::Problem,
 * src/problem/problem.cpp:72-94) without its host cost. The scan points are
 * the view's point rows with expected SDF 0 (in row order), placed by the
 * problem's poses with frame_cos_sin (2 per frame: cos and sin of the pose
 * angle as the caller's own point placement computes them; NULL: both from
 * one ::sincos call per frame, as the host library's Pose2 does); every
 * eikonal spec takes
 * its node's normal. Only the pca and hybrid methods, whose eikonal scale is
 * 1 (the specs do not depend on the normals), are accepted; the values
 * equal the host builder's bit for bit. No counterpart in the reference. */
typedef struct sdfslam_normals_request {
  int32_t k_neighbors;
  int32_t method; /* SDFSLAM_NORMALS_PCA or SDFSLAM_NORMALS_HYBRID */
  double corner_threshold;
  const double* frame_cos_sin;
} sdfslam_normals_request;
int sdfslam_hip_context_update_device_normals(sdfslam_hip_context* ctx, const sdfslam_problem_view* problem,
                                              const double* map_values, const double* poses,
                                              const sdfslam_normals_request* normals);

/* Replaces the context's problem and state with a new one (the next
 * increment of an incremental run: the problem rebuilt over more frames, as
 * runner.cpp:249-260 rebuilds it before each Solve), keeping the context's
 * stream and device memory. The next context_solve behaves as a Solve on a
 * fresh context created with the new problem and the context's options, to
 * rounding: when the new problem extends the old one (same grid, more
 * poses), the earlier poses keep their footprint margins and the appended
 * ones start wider, so the J^T J pattern superset (and with it the order of
 * the factorization's sums) can differ from a fresh context's. */
int sdfslam_hip_context_update(sdfslam_hip_context* ctx, const sdfslam_problem_view* problem,
                               const double* map_values, const double* poses);
/* Extends a context's problem by the next frames of an incremental run, as
 * the reference's runner rebuilds the Problem over all frames so far before
 * each Solve (runner.cpp:245-258 -> Problem::Problem,
 * src/problem/problem.cpp:11-115), with the per-frame construction on the
 * device and nothing of the earlier frames rebuilt or uploaded again:
 *  - each frame's point rows (its scan points, then their hallucinated
 *    samples, src/core/hallucination.cpp:5-28) are generated on the device
 *    from the raw scan and appended to the rows the context holds;
 *  - one odometry spec per frame joins it to the frame before;
 *  - every eikonal spec's normal is recomputed on the device from all scan
 *    points at the given poses (src/core/normals.cpp:128-220, as in
 *    sdfslam_hip_context_update_device_normals);
 *  - with the active region the active node set is recomputed on the device
 *    from every point row at the given poses (problem.cpp:39-70) and read
 *    back (the one device round trip of this call; without it the call
 *    does host work only and defers its device work to the context's next
 *    call, as an update does), and the eikonal specs follow it.
 * The resulting problem equals the one Problem::Problem builds from all
 * frames so far (the tests solve both and compare the results bit for bit).
 * The context's problem must be such a session problem: consecutive
 * odometry only (no relations), eikonal specs from the pca or hybrid
 * normals, the same options as below. scan_xy holds the frames' sensor-frame
 * points (x, y) one frame after another, points_per_frame their counts,
 * odometry 3 values (dx, dy, dtheta) per frame; map_values and poses are
 * the state the problem is built at (num_poses + num_frames poses: the
 * earlier poses, then the new frames' initial poses). No counterpart in the
 * reference (it rebuilds the whole problem on the host). */
typedef struct sdfslam_frame_options {
  int32_t hallucination_points;  /* per scan point (HallucinationOptions) */
  double hallucination_step;     /* > 0 */
  int32_t hallucination_both_directions;
  double weight_scan;  /* Weights: sqrt taken as Problem::Problem does */
  double weight_hallucination;
  double weight_eikonal;
  double weight_odometry;
  int32_t active_region;
  int32_t active_margin;
  sdfslam_normals_request normals; /* frame_cos_sin must be NULL */
} sdfslam_frame_options;
int sdfslam_hip_context_append_frames(sdfslam_hip_context* ctx, int32_t num_frames,
                                      const int64_t* points_per_frame, const double* scan_xy,
                                      const double* odometry, const sdfslam_frame_options* options,
                                      const double* map_values, const double* poses);
/* The point rows the context holds, as the device holds them: the row
 * count into *num_points (may be NULL) and the first min(capacity, count)
 * rows into the arrays (any may be NULL). For checking appended rows. */
int sdfslam_hip_context_point_rows(sdfslam_hip_context* ctx, int64_t capacity, int64_t* num_points,
                                   int32_t* frame, double* x, double* y, double* expected_sdf,
                                   double* sqrt_weight);
/* Process-wide: LM iterations as captured HIP graphs (enabled != 0, the
 * default) or as direct launches of the same kernels in the same order
 * (the same results; for kernel profilers and debuggers that do not follow
 * graph launches). No counterpart in the reference. */
int sdfslam_hip_set_graphs(int32_t enabled);
/* Process-wide: the pose-block and panel extend-add kernels use their LDS
 * forms (enabled, the default) or their global-memory forms, which give the
 * same results bit for bit (the forms a frame with more than 192 occupied
 * cells or a front of more than 2048 rows takes anyway). Launches made or
 * captured after the call follow it. No counterpart in the reference (its
 * CPU path has no kernels); for tests that exercise both forms. */
int sdfslam_hip_set_lds_forms(int32_t enabled);
/* Footprint margin (cells) of every pose in the current pattern epoch:
 * min(capacity, num_poses) entries into margins (may be NULL), the pose
 * count into *num_poses (may be NULL). */
int sdfslam_hip_pose_margins(sdfslam_hip_context* ctx, int32_t* margins, int32_t capacity,
                             int32_t* num_poses);
int sdfslam_hip_set_state(sdfslam_hip_context* ctx, const double* map_values, const double* poses);
int sdfslam_hip_get_state(sdfslam_hip_context* ctx, double* map_values, double* poses);

/* Sizes of the analysis: num_parameters, num_residuals, nnz of the lower
 * triangle of the fixed J^T J pattern, supernodes, nnz of the supernodal factor
 * (incl. explicit zeros), and the factorization flop count sum_j (c_j^2 + 3 c_j);
 * pattern_epochs counts the pattern epochs so far (speculative_epochs of them
 * analysed on a host thread while the device iterated) and analysis_seconds
 * the host time they took on the calling thread (analysis, uploads,
 * allocation). */
typedef struct sdfslam_hip_analysis_info {
  int64_t num_parameters;
  int64_t num_residuals;
  int64_t pattern_nnz_lower;
  int64_t num_supernodes;
  int64_t factor_nnz;
  double factor_flops;
  int64_t update_doubles;
  int32_t tree_levels;
  int32_t pattern_epochs;
  double analysis_seconds;
  int64_t device_bytes; /* device memory the context holds */
  /* Launches of the first 64-column panel step per factorization (one per
   * level; later steps are fused into the previous step's trailing update)
   * and their algorithmic FP64 flops (the fronts' first diagonal blocks,
   * 2/3 nb^3 each, and nb^2 per TRSM row below them). */
  int32_t panel_step_launches;
  double panel_step_flops;
  int32_t speculative_epochs;
  /* Launches of the trailing-update / update-matrix kernels (FP64 MFMA)
   * per factorization and their algorithmic flops (2 K per lower-triangle
   * output entry, plus the diagonal block and TRSM rows of the step that a
   * trailing update prepares, counted once). */
  int32_t gemm_launches;
  double gemm_flops;
  /* 64-column diagonal blocks of the factor (one kept 64 x 64 inverse
   * each, read by the solves). */
  int64_t diag_blocks;
} sdfslam_hip_analysis_info;
int sdfslam_hip_analysis(sdfslam_hip_context* ctx, sdfslam_hip_analysis_info* info);

/* The factorization's size without this backend's padding: the symbolic
 * LDL^T of the reference's J^T J pattern at the current state (every pair of
 * columns sharing a Jacobian row, plus the diagonal; lm_solver.cpp:113-172,
 * 321-330) under the context's elimination order. factor_nnz: entries of L
 * below the diagonal; factor_flops: sum over columns of c^2 + 3c (the
 * count of the analysis' factor_flops, which includes the superset
 * pattern's explicit zeros and the supernodes' dense blocks). */
int sdfslam_hip_reference_factor_counts(sdfslam_hip_context* ctx, int64_t* factor_nnz, double* factor_flops);
/* The context's current elimination order: order[k] = the column (in the
 * reference's column layout) eliminated k-th; num_parameters entries. */
int sdfslam_hip_elimination_order(sdfslam_hip_context* ctx, int32_t* order);

/* Residual vector and Jacobian rows at the current state
 * (AssemblePointRows/AssembleEikonalRows/odometry tail, lm_solver.cpp:30-82,
 * 275-296). Rows are in reference order; each row has 7 slots of
 * (column, value) with column -1 for an empty slot, in the reference's triplet
 * order within the row. residuals: num_residuals; j_cols/j_vals: 7*num_residuals. */
int sdfslam_hip_assemble(sdfslam_hip_context* ctx, double* residuals, int32_t* j_cols,
                         double* j_vals);

/* Damped normal equations at the current state (ParallelAtA + damping +
 * ParallelNegAtb, lm_solver.cpp:113-197, 319-331), lower triangle in the
 * original column order as CSC over the fixed pattern: col_ptr (n+1), row_idx
 * and values (pattern_nnz_lower), structural (pattern_nnz_lower bytes: 1 where
 * the reference's J^T J has a structural entry, i.e. some row couples the two
 * columns, or on the diagonal), rhs = -J^T r (n). The pattern is first made
 * valid for the current state (which may start a new pattern epoch); its
 * size is stored in *nnz. When it exceeds `capacity`, nothing else is
 * written and SDFSLAM_ERR_INVALID_ARGUMENT is returned: call again with
 * buffers of *nnz entries. */
int sdfslam_hip_normal_equations(sdfslam_hip_context* ctx, double lambda, int64_t capacity,
                                 int64_t* nnz, int32_t* col_ptr, int32_t* row_idx, double* values,
                                 uint8_t* structural, double* rhs);

/* Solution of (J^T J + damping) step = -J^T r at the current state
 * (lm_solver.cpp:351-364). Returns SDFSLAM_ERR_NOT_SPD on a zero pivot. */
int sdfslam_hip_linear_step(sdfslam_hip_context* ctx, double lambda, double* step);

/* Total weighted squared error at the current state (Problem::Cost,
 * problem.cpp:132-156). */
int sdfslam_hip_cost(sdfslam_hip_context* ctx, double* cost);

/* Runs the LM loop on the context's device state (SolveLm,
 * lm_solver.cpp:201-434); the state stays on the device. */
int sdfslam_hip_context_solve(sdfslam_hip_context* ctx, const sdfslam_solver_options* options,
                              sdfslam_solve_result* result);

/* Runs the LM loop (as sdfslam_hip_context_solve) on `count` contexts at
 * once: every stage of an iteration is one launch over all of them, so many
 * small problems share each launch (the batched and served forms). Each
 * context keeps its own LM state, pattern epochs and result, identical to
 * solving it alone. statuses[i] (may be NULL) receives SDFSLAM_OK or
 * SDFSLAM_ERR_NOT_SPD per context; results (may be NULL) one result each.
 * The contexts must be distinct, on one device, and idle. info (may be NULL)
 * receives where the call spent its time. No counterpart in the reference
 * (one Solve per call, solver.hpp:67). */
typedef struct {
  int32_t graph_captures;   /* captured iterations (one per change of the batch's buffers) */
  int32_t graph_launches;   /* launches of the captured iteration */
  int32_t epochs_rebuilt;   /* pattern re-analyses of any context during the run */
  double setup_seconds;     /* LM initialisation of every context */
  double capture_seconds;   /* slot table, launch plans, capture and instantiation */
  double analysis_seconds;  /* re-analyses (host) with their uploads */
  double run_seconds;       /* the whole call */
} sdfslam_hip_batch_info;

int sdfslam_hip_batch_solve(sdfslam_hip_context* const* contexts, int32_t count,
                            const sdfslam_solver_options* options, sdfslam_solve_result* results,
                            int32_t* statuses, sdfslam_hip_batch_info* info);

/* Node normals of a problem build on the device, for incremental and served
 * runs whose problem is rebuilt every increment (problem.cpp:72-94 via
 * normals.cpp:128-220, the all-scans neighbourhood; the per-scan variant
 * stays on the host): the kNN PCA normal of every scan point (points_xy and
 * their sensor origins origins_xy are x, y pairs in world coordinates, k =
 * k_neighbors, at most 32), then per grid node the normal of the nearest
 * scan point and its eikonal scale for `method`. Exact neighbour sets with
 * ties broken by point index and the same arithmetic as the host builder,
 * so the outputs equal AssignNodeNormals(ComputeScanNormals(...)) bit for
 * bit. node_normal_xy receives 2 * nx * ny values, node_scale nx * ny. */
#define SDFSLAM_NORMALS_PCA 0
#define SDFSLAM_NORMALS_HYBRID 1
#define SDFSLAM_NORMALS_WEIGHTED 2
int sdfslam_hip_node_normals(const double* points_xy, const double* origins_xy, int64_t num_points,
                             int32_t k_neighbors, const sdfslam_grid* grid, int32_t method,
                             double corner_threshold, double* node_normal_xy, double* node_scale);

/* Map seeding from scans, replacing InitializeFromScans
 * (src/core/map_init.hpp:27-28, map_init.cpp:31-69; called by
 * runner.cpp:219-221 over all scans of a batch run and runner.cpp:231-233
 * over the first scan of an incremental run). points_xy are the scan points
 * in world coordinates and origins_xy their sensor positions (x, y pairs,
 * num_points of each). Every grid node receives the distance to its nearest
 * point, sqrt(dx*dx + dy*dy); with signed_distance != 0 it is negated when
 * the node is farther from that point's origin than the point is. Ties in
 * distance go to the lower point index. num_points == 0 leaves map_values
 * (nx * ny, node-major) unchanged. */
int sdfslam_hip_init_from_scans(const double* points_xy, const double* origins_xy, int64_t num_points,
                                const sdfslam_grid* grid, int32_t signed_distance, double* map_values);

/* sdfslam_hip_init_from_scans without any device allocation. No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments and results, but points_xy, origins_xy and
 * map_values are DEVICE pointers, and the work is enqueued on `stream` (a
 * hipStream_t; NULL = the null stream) without host synchronisation: the
 * values are in map_values once the stream reaches that point. Every
 * intermediate buffer lives in one file-scope device array per GPU
 * (SDFSLAM_NO_ALLOC_SCRATCH_MIB, held until the process ends), so at most
 * SDFSLAM_NO_ALLOC_MAX_POINTS points are accepted (more: INVALID_ARGUMENT,
 * nothing enqueued). Calls that overlap on one GPU share that array and
 * race: serialise them (one stream, or wait between calls). */
#define SDFSLAM_NO_ALLOC_MAX_POINTS 524288
#define SDFSLAM_NO_ALLOC_SCRATCH_MIB 24
int sdfslam_hip_init_from_scans_noalloc(const double* points_xy, const double* origins_xy,
                                        int64_t num_points, const sdfslam_grid* grid,
                                        int32_t signed_distance, double* map_values, void* stream);

/* Runs exactly `iterations` LM iterations from the current device state,
 * carrying the LM state (lambda, previous cost) in the context and ignoring
 * max_iterations; used to time iterations. With collect_phase_times != 0
 * every iteration is bracketed by HIP events per phase (and synchronized),
 * otherwise iterations are enqueued back to back. */
int sdfslam_hip_context_iterate(sdfslam_hip_context* ctx, int32_t iterations,
                                int32_t collect_phase_times);

/* Average device time in milliseconds per iteration of each phase over the
 * last context_iterate with collect_phase_times (HIP events on the context's
 * stream): assembly, normal equations, factorization, triangular solves,
 * step+cost+control, their total, the point-row assembly kernel alone, and
 * the average panel-step launch (diagonal block + TRSM kernels of one
 * 64-column step, bracketed by their own events). */
typedef struct sdfslam_hip_phase_times {
  double assembly_ms;
  double normal_ms;
  double factor_ms;
  double solve_ms;
  double update_ms;
  double total_ms;
  double point_assembly_kernel_ms;
  double panel_step_kernel_ms; /* average duration of one panel-step launch */
  double gemm_kernel_ms;       /* average duration of one trailing-update launch */
} sdfslam_hip_phase_times;
int sdfslam_hip_phase_times_get(sdfslam_hip_context* ctx, sdfslam_hip_phase_times* out);

#ifdef __cplusplus
}
#endif

#endif /* SDFSLAM_HIP_H_ */
