// Served and batched forms of the device backend.
//
// The reference solves one Problem per call (Solve, src/solvers/solver.hpp:67)
// and runs one incremental session per process (src/pipeline/runner.cpp:
// 226-285). The server keeps many incremental sessions on one GPU, each with
// its own device context (memory reused from one increment to the next).
// Continuous batching in two stages. Preparer threads (`workers`) take the
// idle session with the oldest waiting frame, rebuild its problem and load
// its context (host work only: the context's uploads run at the start of its
// next solve) and queue it as prepared, so a frame is prepared as soon as it
// arrives and its session is idle. Solver threads (`pipelines`) take every
// prepared increment (oldest first, up to max_batch), solve them together
// with sdfslam_hip_batch_solve (every stage of an LM iteration is one launch
// over all of them) on their own stream, and read the states back. While the
// GPU is busy, prepared increments accumulate and the next batch is larger;
// when it is idle, an increment is solved as soon as it is prepared. Each
// session's increments run one at a time in submission order, and a
// session's results do not depend on what else is being served.
//
// Admission: a session is admitted while the device memory of the admitted
// sessions plus its own stays within the budget: a session's exact device
// bytes (its context's pools) once it has solved, a per-node reservation
// before. A session that does not fit waits in a queue (its frames are
// accepted and wait with it) after the server has freed what the idle
// sessions' contexts outgrew; queued sessions are admitted in order as
// closed sessions free their memory. A session whose reservation alone
// exceeds the budget is refused. Sessions with different step options are
// never solved in the same batch.
#pragma once

#include <condition_variable>
#include <cstdint>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "runner.hpp"

namespace sdf_slam {

/// Totals over the batches a server ran.
struct ServerStats {
  int64_t batches{0};
  int64_t increments{0};
  int64_t graph_captures{0};
  int64_t epochs_rebuilt{0};
  double prepare_seconds{0.0};
  double solve_seconds{0.0};
  double collect_seconds{0.0};
  double capture_seconds{0.0};
  double analysis_seconds{0.0};
};

struct ServerOptions {
  int workers{12};                 // host threads preparing increments
  int max_batch{64};               // sessions per batched solve (1: one at a time)
  int pipelines{3};                // batched solves in flight (solver threads)
  int64_t device_budget_bytes{0};  // 0: 80 % of the device's memory
  int max_in_flight{0};            // increments prepared or solving at once (0: no limit;
                                   // 1: one increment at a time, the plain path)
  bool device_normals{false};      // rebuilds take node normals from the device (same values)
};

class Server {
 public:
  explicit Server(const ServerOptions& options);
  ~Server();
  Server(const Server&) = delete;
  Server& operator=(const Server&) = delete;

  /// Opens a session at its first frame (pose and scan, no solve yet),
  /// admitted at once or queued until memory frees up. Throws
  /// std::runtime_error("admission refused: ...") when the session alone
  /// would exceed the device memory budget.
  int OpenSession(GridMap map, const ProblemOptions& problem, const SolverOptions& step,
                  const Pose2& first_pose, const Scan& first_scan);
  /// Closes a session: waits for its submitted increments (a session still
  /// queued for admission drops them; their Wait fails), frees its device
  /// context and admits queued sessions that now fit. Its last estimate
  /// stays readable with State; Submit fails from now on.
  void CloseSession(int session);
  struct SessionInfo {
    bool admitted{false};
    bool closed{false};
    int64_t device_bytes{0};  // counted against the budget
    int64_t submitted{0};
    int64_t finished{0};
  };
  SessionInfo Info(int session);
  /// Queues the next frame of a session (scan and odometry relative to the
  /// previous frame); `arrival` is the steady-clock time latency counts from.
  /// Returns the increment's ticket (1, 2, ... per session).
  int64_t Submit(int session, const Scan& scan, const Pose2& odometry,
                 std::chrono::steady_clock::time_point arrival);
  /// Blocks until the increment with `ticket` of `session` is done and
  /// returns its record; rethrows the increment's failure.
  IncrementRecord Wait(int session, int64_t ticket);
  /// Latency of a finished increment (completion minus arrival), seconds.
  double Latency(int session, int64_t ticket);
  /// Current estimate of a session (after its last finished increment).
  void State(int session, std::vector<Pose2>& poses, GridMap& map);
  /// Device memory held by all sessions.
  int64_t DeviceBytes();
  /// Changes the limit on increments prepared or solving at once (as
  /// ServerOptions::max_in_flight) for the increments picked from now on.
  void SetMaxInFlight(int max_in_flight);
  /// Totals over the batches run so far.
  ServerStats Stats();

 private:
  struct Pending {
    Scan scan;
    Pose2 odometry;
    std::chrono::steady_clock::time_point arrival;
  };
  struct Session {
    GridMap map;
    std::vector<Pose2> estimated;
    std::vector<Scan> scans;
    std::vector<Pose2> odometry;
    ProblemOptions problem;
    SolverOptions step;
    HipSession hip;
    std::deque<Pending> pending;
    bool busy{false};
    int64_t submitted{0};
    std::vector<IncrementRecord> records;
    std::vector<double> latency;
    std::vector<std::string> errors;
    int64_t device_bytes{0};
    bool admitted{false};
    bool closing{false};
    bool closed{false};
  };

  /// One increment between its preparation and its solve.
  struct Prepared {
    Session* session{nullptr};
    std::chrono::steady_clock::time_point arrival;
    std::chrono::steady_clock::time_point prepared;
    std::unique_ptr<Problem> problem;
    sdfslam_hip_context* ctx{nullptr};
    IncrementRecord record;
    std::string error;
  };

  void PrepareLoop();
  void SolveLoop();
  /// Appends a session's next frame to its history, rebuilds its problem
  /// and loads its context (runner.cpp:240-260 for one appended frame).
  void Prepare(Prepared& item, Pending& frame);
  /// One batched solve of prepared increments and the read-back of their
  /// states; fills each item's record or error.
  void SolveBatch(std::vector<std::unique_ptr<Prepared>>& batch);
  /// Records a finished increment (under mu_).
  void Finish(Prepared& item, std::chrono::steady_clock::time_point done, int64_t device_bytes);
  /// Device bytes counted for the admitted, open sessions (under mu_).
  int64_t Used() const;
  /// Admits queued sessions in order while they fit (under mu_).
  void AdmitWaiting();
  /// Frees what the idle admitted sessions' contexts outgrew; takes and
  /// releases `lock` (on mu_) around the device work.
  void TrimIdle(std::unique_lock<std::mutex>& lock);

  ServerOptions options_;
  std::mutex mu_;
  std::condition_variable work_cv_;
  std::condition_variable solve_cv_;
  std::condition_variable done_cv_;
  std::vector<std::unique_ptr<Session>> sessions_;
  std::deque<Session*> waiting_;  // sessions queued for admission, in order
  std::deque<std::unique_ptr<Prepared>> prepared_;
  std::vector<std::thread> preparers_;
  std::vector<std::thread> solvers_;
  bool stop_{false};
  bool stop_solvers_{false};
  int in_flight_{0};
  std::mutex stats_mu_;
  ServerStats stats_;
};

/// Served benchmark: `sessions` incremental sessions (synthetic floorplans,
/// seeds seed0 + i), each receiving frames as a Poisson process of
/// `rate_hz` for `duration_s` seconds of wall time (from frame start_frame
/// on, the earlier frames solved before the clock starts), served by
/// `workers`.
struct ServedBenchmark {
  int sessions{32};
  double rate_hz{10.0};
  double duration_s{10.0};
  int workers{12};
  int max_batch{64};
  int pipelines{3};
  uint64_t seed0{2000};
  int max_in_flight{0};
  bool device_normals{false};
  // Frames 1 .. start_frame - 1 of every session are solved before the
  // clock starts (all submitted at once, no limit on increments in flight),
  // so the timed frames measure sessions that already hold start_frame
  // poses; arrivals then continue from frame start_frame.
  int start_frame{1};
};
struct ServedResult {
  int64_t increments{0};
  int64_t lm_iterations{0};
  double wall_seconds{0.0};  // first arrival to last completion
  double increments_per_second{0.0};
  double latency_p50_ms{0.0};
  double latency_p99_ms{0.0};
  double latency_mean_ms{0.0};
  double build_mean_ms{0.0};
  double solve_mean_ms{0.0};
  int64_t device_bytes{0};
  int64_t batches{0};
  double mean_batch{0.0};
  double prepare_batch_ms{0.0};  // host preparation of a batch's increments (overlaps other solves)
  double solve_batch_ms{0.0};
  double collect_batch_ms{0.0};
  double capture_batch_ms{0.0};
  double analysis_batch_ms{0.0};
  int64_t epochs_rebuilt{0};
  double latency_p50_after_first_ms{0.0};  // each session's first increment left out
  double latency_p99_after_first_ms{0.0};
  double first_increment_mean_ms{0.0};
  double queue_p99_ms{0.0};       // arrival to the start of its preparation
  double prepare_p99_ms{0.0};     // the preparation
  double solve_wait_p99_ms{0.0};  // prepared to the start of its batched solve
  double solve_p99_ms{0.0};       // the batched solve call
};
ServedResult RunServedBenchmark(const ServedBenchmark& options);

/// Batched benchmark: `instances` independent batch problems (seeds
/// seed0 + i, uneven sizes), each solved for `iterations` LM iterations.
/// The instances, largest first, are cut into batched solves of at most
/// `max_batch` (similar sizes share a launch, so few workgroups idle), and
/// `workers` host threads run the batches concurrently, each on its own
/// stream. Host problem construction is done before the clock starts; the
/// clock covers device setup (one context per instance) and the iterations.
struct BatchBenchmark {
  int instances{64};
  int iterations{30};
  int max_batch{8};
  int workers{8};
  uint64_t seed0{1000};
};
struct BatchResult {
  int64_t lm_iterations{0};
  int64_t unknowns{0};
  double wall_seconds{0.0};
  double lm_iterations_per_second{0.0};
};
BatchResult RunBatchBenchmark(const BatchBenchmark& options);

/// Autotuner of the batched form's shapes: times `trial_iterations` LM
/// iterations of the benchmark's instances for every (max_batch, workers)
/// candidate pair (instances per batched solve, batched solves in flight on
/// their own streams) and keeps the fastest. The instances are built once
/// and shared by the trials; each trial creates its contexts afresh, as a
/// run does.
struct BatchShape {
  int max_batch{0};
  int workers{0};
  double lm_iterations_per_second{0.0};
};
struct BatchTuneResult {
  BatchShape best;
  std::vector<BatchShape> trials;
};
BatchTuneResult TuneBatchShapes(const BatchBenchmark& options, const std::vector<int>& max_batches,
                                const std::vector<int>& workers, int trial_iterations);

}  // namespace sdf_slam
