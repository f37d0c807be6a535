/*
 * sdfslam_host.h - C API of the host library (libsdfslam_host.so): the
 * mirror of the reference's Problem builder and Solve entry point, plus the
 * seeded synthetic workloads. It is how Python tests and the benchmark build
 * problems and reach Solve(problem, {backend: hip}).
 *
 * Reference interfaces mirrored (reference file:line):
 *   sdfslam_host_problem_create  <- Problem::Problem, src/problem/problem.cpp:11-115
 *   sdfslam_host_solve           <- Solve, src/solvers/solver.hpp:67 (+ kHip)
 *   sdfslam_host_parse_solver_yaml <- LoadConfig's solver block, src/pipeline/config.cpp:170-190
 *   sdfslam_host_run_incremental <- Run's incremental branch, src/pipeline/runner.cpp:226-285
 *   sdfslam_host_run             <- Run, src/pipeline/runner.cpp:196-303 (+ output writers 73-113, 172-192)
 *   sdfslam_host_init_from_scans <- InitializeFromScans, src/core/map_init.cpp:31-69
 */
#ifndef SDFSLAM_HOST_H_
#define SDFSLAM_HOST_H_

#include <stdint.h>

#include "sdfslam_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ProblemOptions (src/problem/problem.hpp:14-87). normals_method: 0 pca,
 * 1 hybrid, 2 weighted. */
typedef struct sdfslam_problem_options {
  double weight_scan;
  double weight_hallucination;
  double weight_eikonal;
  double weight_odometry;
  double weight_relation;
  int32_t hallucination_points_per_scan_point;
  double hallucination_step_size;
  int32_t hallucination_both_directions;
  int32_t normals_method;
  int32_t normals_k_neighbors;
  int32_t normals_per_scan;
  double normals_corner_threshold;
  double huber_delta;
  int32_t smooth_gradient;
  double smooth_gradient_step;
  int32_t active_region;
  int32_t active_margin;
} sdfslam_problem_options;

typedef struct sdfslam_host_problem sdfslam_host_problem;

/* Defaults of ProblemOptions / SolverOptions as the reference declares them. */
void sdfslam_host_default_problem_options(sdfslam_problem_options* out);
void sdfslam_host_default_solver_options(sdfslam_solver_options* out);

/* Message of the last failed call on this thread. */
const char* sdfslam_host_last_error(void);

/* Builds a Problem. scans: num_frames scans, scan k has points
 * scan_xy[2*scan_offsets[k] .. 2*scan_offsets[k+1]) in the sensor frame;
 * poses: 3*num_frames; odometry: 3*(num_frames-1) or NULL; relations:
 * rel_frames (2 per relation), rel_meas (3 per relation), rel_sqrt_weight.
 * map_values: nx*ny initial values (NULL = 0). */
int sdfslam_host_problem_create(const sdfslam_grid* grid, const double* map_values,
                                int32_t num_frames, const double* poses,
                                const int64_t* scan_offsets, const double* scan_xy,
                                const double* odometry, int32_t num_relations,
                                const int32_t* rel_frames, const double* rel_meas,
                                const double* rel_sqrt_weight,
                                const sdfslam_problem_options* options,
                                sdfslam_host_problem** out);

/* Node normals on the host (ComputeScanNormals' all-scans neighbourhood and
 * AssignNodeNormals, normals.cpp:128-220), with the arguments and outputs
 * of sdfslam_hip_node_normals: the reference values the device ones are
 * checked against. */
int sdfslam_host_node_normals(const double* points_xy, const double* origins_xy, int64_t num_points,
                              int32_t k_neighbors, const sdfslam_grid* grid, int32_t method,
                              double corner_threshold, double* node_normal_xy, double* node_scale);

/* Builds the Problem of a seeded synthetic workload at its initial state:
 * "room", "room_small", "office150", "office200", "office_small", "corner". */
int sdfslam_host_workload_create(const char* name, uint64_t seed, sdfslam_host_problem** out);

void sdfslam_host_problem_destroy(sdfslam_host_problem* problem);

/* View of the specs; pointers stay valid while the problem lives. */
int sdfslam_host_problem_view(const sdfslam_host_problem* problem, sdfslam_problem_view* out);
int sdfslam_host_problem_get_state(const sdfslam_host_problem* problem, double* map_values,
                                   double* poses);
int sdfslam_host_problem_set_state(sdfslam_host_problem* problem, const double* map_values,
                                   const double* poses);

/* Workload metadata: the solver options of the workload, whether it is an
 * incremental run, its iterations per increment, and ground-truth poses
 * (3*num_poses, may be NULL). */
int sdfslam_host_workload_info(const sdfslam_host_problem* problem,
                               sdfslam_solver_options* solver, int32_t* incremental,
                               int32_t* iterations_per_increment, double* ground_truth);

/* Solve(problem, options) with backend "hip" (the only backend built here);
 * any other name fails with the reference's logic_error text. Returns a
 * status code; the exception message is in sdfslam_host_last_error(). */
int sdfslam_host_solve(sdfslam_host_problem* problem, const char* backend,
                       const sdfslam_solver_options* options, sdfslam_solve_result* result);

/* One increment of an incremental run. */
typedef struct sdfslam_increment_record {
  int32_t frames;       /* frames in the increment's problem */
  int32_t iterations;
  double initial_cost;
  double final_cost;
  double build_seconds; /* Problem construction (hallucination, normals, specs) */
  double solve_seconds; /* the increment's Solve, transfers included */
} sdfslam_increment_record;

/* Solve function of one increment, with sdfslam_hip_solve's contract. */
typedef int (*sdfslam_solve_fn)(void* user, const sdfslam_problem_view* problem,
                                const sdfslam_solver_options* options, double* map_values,
                                double* poses, sdfslam_solve_result* result);

/* Incremental run of a synthetic workload (runner.cpp:226-285): frames are
 * appended `increment_size` at a time by composing odometry onto the last
 * optimized pose, the Problem is rebuilt over all frames so far and refined
 * with `options` (NULL: the workload's solver options with max_iterations =
 * its iterations per increment). Runs over the first max_frames frames (all
 * when <= 0). solve_fn NULL solves on the device through one reused context;
 * otherwise solve_fn(user, ...) solves each increment. total: accumulated as
 * runner.cpp:262-270 (cost_history is caller storage, may be NULL). records:
 * up to records_capacity increments, *num_records = increments run.
 * poses_out (3 * frames) and map_out (nx * ny) receive the final estimate
 * (either may be NULL). */
int sdfslam_host_run_incremental(const char* workload, uint64_t seed,
                                 const sdfslam_solver_options* options, int32_t increment_size,
                                 int32_t max_frames, sdfslam_solve_fn solve_fn, void* user,
                                 sdfslam_solve_result* total, sdfslam_increment_record* records,
                                 int32_t records_capacity, int32_t* num_records, double* poses_out,
                                 double* map_out);

/* A whole run of a synthetic workload, replacing Run
 * (src/pipeline/runner.cpp:196-303, declared in runner.hpp): batch (one
 * solve from the workload's initial poses) or incremental (as
 * sdfslam_host_run_incremental), with map seeding from the scans when
 * init_from_scans != 0 (all scans at the initial poses for a batch run, the
 * first scan for an incremental one; init_signed as map.init_signed) and,
 * when output_dir is not NULL, the reference's output files there:
 * poses_odometry.csv, poses_estimated.csv, map.txt, scan_points_global.csv,
 * poses_ground_truth.csv, metrics.json, and snapshots/map_NNNNNN.txt +
 * poses_NNNNNN.csv every snapshot_every increments and after the last.
 * options NULL: the workload's solver options (max_iterations per increment
 * for incremental runs). solve_fn as in sdfslam_host_run_incremental. */
typedef struct sdfslam_run_options {
  int32_t incremental;     /* 0: batch, 1: incremental */
  int32_t increment_size;  /* frames per increment (incremental) */
  int32_t max_frames;      /* incremental: frames used, all when <= 0 */
  int32_t snapshot_every;  /* incremental: increments between snapshots, 0: none */
  int32_t init_from_scans; /* map.init_from_scans */
  int32_t init_signed;     /* map.init_signed */
} sdfslam_run_options;
int sdfslam_host_run(const char* workload, uint64_t seed, const sdfslam_run_options* run,
                     const sdfslam_solver_options* options, sdfslam_solve_fn solve_fn, void* user,
                     const char* output_dir, sdfslam_solve_result* total);

/* Map seeding from scans in the sensor frame (InitializeFromScans,
 * src/core/map_init.hpp:27-28): scan i holds the points
 * scan_xy[2 * scan_ptr[i] .. 2 * scan_ptr[i + 1]) and is placed by pose i
 * (x, y, theta); scans beyond num_poses are ignored. The points are placed on
 * the host as the problem builder places them and the distances are taken on
 * the device (sdfslam_hip_init_from_scans). map_values (nx * ny) in/out. */
int sdfslam_host_init_from_scans(const sdfslam_grid* grid, double* map_values, const double* scan_xy,
                                 const int64_t* scan_ptr, int32_t num_scans, const double* poses,
                                 int32_t num_poses, int32_t signed_distance);

/* ---- Datasets: the frames of a synthetic workload -------------------------
 * A dataset holds a workload's scans (sensor frame), its odometry
 * (RelativeOdometry of the dataset poses, runner.cpp:52-59), its dataset
 * poses, map grid and options, without building a Problem. */
typedef struct sdfslam_host_dataset sdfslam_host_dataset;
int sdfslam_host_dataset_create(const char* workload, uint64_t seed, sdfslam_host_dataset** out);
void sdfslam_host_dataset_destroy(sdfslam_host_dataset* dataset);
/* Frame count, grid, the workload's problem options and its per-increment
 * solver options (max_iterations = iterations per increment). */
int sdfslam_host_dataset_info(const sdfslam_host_dataset* dataset, int32_t* num_frames,
                              sdfslam_grid* grid, sdfslam_problem_options* problem,
                              sdfslam_solver_options* step);
/* Frame k: its scan (2 * *num_points doubles into scan_xy, at most
 * capacity points), the dataset pose (3) and the odometry from frame k-1
 * to k (3; zeros for frame 0). Any output may be NULL. */
int sdfslam_host_dataset_frame(const sdfslam_host_dataset* dataset, int32_t frame, double* scan_xy,
                               int64_t capacity, int64_t* num_points, double* pose,
                               double* odometry);

/* ---- Served form: many incremental sessions on one GPU --------------------
 * No counterpart in the reference (one session per process, runner.cpp:
 * 226-285). Each session's increments run in submission order on the
 * session's own device context. Continuous batching: a scheduler takes every
 * idle session with a waiting frame (up to max_batch), prepares them on
 * `workers` host threads and solves them in one sdfslam_hip_batch_solve;
 * `pipelines` batches are in flight at once; sessions with different step
 * options are solved in different batches. A session is admitted while the
 * device memory of the admitted sessions stays within device_budget_bytes
 * (0: 80 % of the device): each session's exact context bytes once it has
 * solved, a reservation of 64 MiB + 4 KiB per grid node before. A session
 * that does not fit is queued, its frames accepted, and admitted in order
 * as closed sessions free their memory. */
typedef struct sdfslam_server sdfslam_server;
typedef struct sdfslam_server_options {
  int32_t workers;
  int32_t max_batch;
  int32_t pipelines;
  int64_t device_budget_bytes;
  int32_t max_in_flight;  /* increments prepared or solving at once; 0: no limit,
                             1: one increment at a time (the plain path) */
  int32_t device_normals; /* 1: problem rebuilds take their node normals from the
                             device (sdfslam_hip_node_normals, same values) */
} sdfslam_server_options;
int sdfslam_server_create(const sdfslam_server_options* options, sdfslam_server** out);
/* Waits for the increments in flight, then frees every session. */
void sdfslam_server_destroy(sdfslam_server* server);
/* Opens a session at its first frame. map_values: nx*ny initial values
 * (NULL = 0). The session is admitted at once or queued (see above); a
 * session whose reservation alone exceeds the budget is refused with
 * SDFSLAM_ERR_OOM. */
int sdfslam_server_open_session(sdfslam_server* server, const sdfslam_grid* grid,
                                const double* map_values, const sdfslam_problem_options* problem,
                                const sdfslam_solver_options* step, const double* first_pose,
                                const double* scan_xy, int64_t num_points, int32_t* session);
/* Queues the session's next frame: its scan and the odometry from the
 * previous frame (3). *ticket numbers the session's increments 1, 2, ... */
int sdfslam_server_submit(sdfslam_server* server, int32_t session, const double* scan_xy,
                          int64_t num_points, const double* odometry, int64_t* ticket);
/* Blocks until increment `ticket` of the session is done; latency_seconds
 * (may be NULL) receives completion minus submission time. */
int sdfslam_server_wait(sdfslam_server* server, int32_t session, int64_t ticket,
                        sdfslam_increment_record* record, double* latency_seconds);
/* Closes a session: waits for its submitted increments (a session still
 * queued for admission drops them, and waiting on them fails), frees its
 * device memory and admits queued sessions that now fit. Its state stays
 * readable; later submits fail. */
int sdfslam_server_close_session(sdfslam_server* server, int32_t session);
/* Admission and progress of a session: admitted (0: queued), closed, the
 * device bytes counted for it, increments submitted and finished. Any
 * output may be NULL. */
int sdfslam_server_session_info(sdfslam_server* server, int32_t session, int32_t* admitted,
                                int32_t* closed, int64_t* device_bytes, int64_t* submitted,
                                int64_t* finished);
/* The session's estimate after its finished increments: *frames poses
 * (3 each, at most capacity frames) and nx*ny map values (may be NULL). */
int sdfslam_server_state(sdfslam_server* server, int32_t session, int32_t capacity,
                         int32_t* frames, double* poses, double* map_values);

/* Served benchmark: `sessions` sessions ("served<i>" workloads, seed seed0)
 * with Poisson frame arrivals of rate_hz each for duration_s seconds. */
typedef struct sdfslam_served_options {
  int32_t sessions;
  double rate_hz;
  double duration_s;
  int32_t workers;
  int32_t max_batch;
  int32_t pipelines;
  uint64_t seed0;
  int32_t max_in_flight;  /* as in sdfslam_server_options */
  int32_t device_normals; /* as in sdfslam_server_options */
  /* Frames 1 .. start_frame - 1 of every session are solved before the
   * clock starts (steady state: the timed frames extend sessions that hold
   * start_frame poses); 1 or less: timed from the first increment. */
  int32_t start_frame;
} sdfslam_served_options;
typedef struct sdfslam_served_result {
  int64_t increments;
  int64_t lm_iterations;
  double wall_seconds;
  double increments_per_second;
  double latency_p50_ms;
  double latency_p99_ms;
  double latency_mean_ms;
  double build_mean_ms;
  double solve_mean_ms;
  int64_t device_bytes;
  /* Where the server spent its time, per batched solve (means). */
  int64_t batches;
  double mean_batch;         /* increments per batched solve */
  double prepare_batch_ms;   /* host problems + context loads (parallel) */
  double solve_batch_ms;     /* the batched solve call */
  double collect_batch_ms;   /* states back into the sessions (parallel) */
  double capture_batch_ms;   /* inside the solve: slot table, plans, graph capture */
  double analysis_batch_ms;  /* inside the solve: pattern re-analyses */
  int64_t epochs_rebuilt;    /* re-analyses inside batched solves */
  /* Latency over every increment but each session's first (which creates
   * the session's device context). */
  double latency_p50_after_first_ms;
  double latency_p99_after_first_ms;
  double first_increment_mean_ms;
  /* 99th percentiles of the parts of an increment's latency: waiting for a
   * preparer (the session's previous increment included), the preparation,
   * waiting for a solver thread, and the batched solve call. */
  double queue_p99_ms;
  double prepare_p99_ms;
  double solve_wait_p99_ms;
  double solve_p99_ms;
} sdfslam_served_result;
int sdfslam_host_served_benchmark(const sdfslam_served_options* options, sdfslam_served_result* out);

/* Batched benchmark: `instances` uneven batch problems ("batch<i>", seed
 * seed0), `iterations` LM iterations each. The instances, largest first, are
 * cut into batched solves of at most max_batch (1: one at a time), which
 * `workers` host threads run concurrently (each on its own stream). */
typedef struct sdfslam_batch_options {
  int32_t instances;
  int32_t iterations;
  int32_t max_batch;
  int32_t workers;
  uint64_t seed0;
} sdfslam_batch_options;
typedef struct sdfslam_batch_result {
  int64_t lm_iterations;
  int64_t unknowns;
  double wall_seconds;
  double lm_iterations_per_second;
} sdfslam_batch_result;
int sdfslam_host_batch_benchmark(const sdfslam_batch_options* options, sdfslam_batch_result* out);

/* Autotuner of the batched form's shapes (no counterpart in the reference):
 * builds the instances of `options` once, then times trial_iterations LM
 * iterations for every pair of max_batches[i] (instances per batched solve)
 * and workers[j] (batched solves in flight, each on its own stream), row
 * major, after one untimed warm-up trial. trials receives
 * num_max_batches * num_workers entries (may be NULL); *best the fastest.
 * options->max_batch, ->workers and ->iterations are not used. */
typedef struct sdfslam_batch_shape {
  int32_t max_batch;
  int32_t workers;
  double lm_iterations_per_second;
} sdfslam_batch_shape;
int sdfslam_host_batch_autotune(const sdfslam_batch_options* options, const int32_t* max_batches,
                                int32_t num_max_batches, const int32_t* workers, int32_t num_workers,
                                int32_t trial_iterations, sdfslam_batch_shape* trials,
                                sdfslam_batch_shape* best);

/* Parses the `solver:` block of a run-config YAML file with the reference's
 * strictness (unknown keys are errors, config.cpp:13-26, 170-190). backend
 * receives "lm", "ceres" or "hip" (capacity >= 8). */
int sdfslam_host_parse_solver_yaml(const char* path, sdfslam_solver_options* out, char* backend,
                                   int32_t backend_capacity);

#ifdef __cplusplus
}
#endif

#endif /* SDFSLAM_HOST_H_ */
