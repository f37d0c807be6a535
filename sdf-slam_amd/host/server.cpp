// Session server and the served / batched benchmarks (server.hpp).
#include "server.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <stdexcept>
#include <string>

#include "scene.hpp"

namespace sdf_slam {

namespace {

using Clock = std::chrono::steady_clock;

double Seconds(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double>(b - a).count();
}

/// Device memory a new session is assumed to need before its first solve:
/// per-node grid arrays and a first epoch's factor storage.
int64_t AdmissionEstimate(const GridMap& map) {
  return (64LL << 20) + 4096LL * map.num_nodes();
}

double Percentile(std::vector<double> v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const double pos = q * static_cast<double>(v.size() - 1);
  const auto lo = static_cast<size_t>(std::floor(pos));
  const size_t hi = std::min(lo + 1, v.size() - 1);
  return v[lo] + (v[hi] - v[lo]) * (pos - static_cast<double>(lo));
}

}  // namespace

Server::Server(const ServerOptions& options) : options_(options) {
  if (options_.workers < 1) throw std::invalid_argument("server needs at least one worker");
  if (options_.device_budget_bytes <= 0) {
    int64_t free_b = 0;
    int64_t total_b = 0;
    if (sdfslam_hip_device_memory(&free_b, &total_b) != SDFSLAM_OK) {
      throw std::runtime_error("no HIP device for the server");
    }
    options_.device_budget_bytes = static_cast<int64_t>(0.8 * static_cast<double>(total_b));
  }
  for (int w = 0; w < options_.workers; ++w) workers_.emplace_back([this] { WorkerLoop(); });
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lock(mu_);
    stop_ = true;
  }
  work_cv_.notify_all();
  for (std::thread& t : workers_) t.join();
}

int Server::OpenSession(GridMap map, const ProblemOptions& problem, const SolverOptions& step,
                        const Pose2& first_pose, const Scan& first_scan) {
  auto s = std::make_unique<Session>();
  const int64_t estimate = AdmissionEstimate(map);
  s->device_bytes = estimate;
  s->map = std::move(map);
  s->estimated = {first_pose};
  s->scans = {first_scan};
  s->problem = problem;
  s->step = step;
  s->step.backend = SolverBackend::kHip;
  std::lock_guard<std::mutex> lock(mu_);
  int64_t used = 0;
  for (const auto& o : sessions_) used += o->device_bytes;
  if (used + estimate > options_.device_budget_bytes) {
    throw std::runtime_error("admission refused: " + std::to_string(used + estimate) +
                             " bytes of device memory would exceed the budget of " +
                             std::to_string(options_.device_budget_bytes));
  }
  sessions_.push_back(std::move(s));
  return static_cast<int>(sessions_.size()) - 1;
}

int64_t Server::Submit(int session, const Scan& scan, const Pose2& odometry, Clock::time_point arrival) {
  int64_t ticket = 0;
  {
    std::lock_guard<std::mutex> lock(mu_);
    Session& s = *sessions_.at(static_cast<size_t>(session));
    s.pending.push_back({scan, odometry, arrival});
    ticket = ++s.submitted;
  }
  work_cv_.notify_one();
  return ticket;
}

IncrementRecord Server::Wait(int session, int64_t ticket) {
  std::unique_lock<std::mutex> lock(mu_);
  Session& s = *sessions_.at(static_cast<size_t>(session));
  if (ticket < 1 || ticket > s.submitted) throw std::invalid_argument("unknown ticket");
  done_cv_.wait(lock, [&] { return static_cast<int64_t>(s.records.size()) >= ticket; });
  const std::string& err = s.errors[static_cast<size_t>(ticket - 1)];
  if (!err.empty()) {
    if (err == "sparse Cholesky factorization failed") throw std::runtime_error(err);
    throw std::runtime_error("increment failed: " + err);
  }
  return s.records[static_cast<size_t>(ticket - 1)];
}

double Server::Latency(int session, int64_t ticket) {
  std::lock_guard<std::mutex> lock(mu_);
  const Session& s = *sessions_.at(static_cast<size_t>(session));
  return s.latency.at(static_cast<size_t>(ticket - 1));
}

void Server::State(int session, std::vector<Pose2>& poses, GridMap& map) {
  std::unique_lock<std::mutex> lock(mu_);
  Session& s = *sessions_.at(static_cast<size_t>(session));
  done_cv_.wait(lock, [&] { return !s.busy; });
  poses = s.estimated;
  map = s.map;
}

int64_t Server::DeviceBytes() {
  std::lock_guard<std::mutex> lock(mu_);
  int64_t used = 0;
  for (const auto& s : sessions_) used += s->device_bytes;
  return used;
}

void Server::RunIncrement(Session& s, Pending p, IncrementRecord& rec) {
  // runner.cpp:240-260 for one appended frame.
  s.estimated.push_back(s.estimated.back().Compose(p.odometry));
  s.scans.push_back(std::move(p.scan));
  s.odometry.push_back(p.odometry);
  rec.frames = static_cast<int>(s.scans.size());
  auto t0 = Clock::now();
  Problem problem(std::move(s.map), s.estimated, s.scans, s.odometry, s.problem);
  rec.build_seconds = Seconds(t0, Clock::now());
  t0 = Clock::now();
  const SolveResult step = s.hip.Solve(problem, s.step);
  rec.solve_seconds = Seconds(t0, Clock::now());
  s.estimated = problem.poses();
  s.map = problem.map();
  rec.iterations = step.iterations;
  rec.initial_cost = step.initial_cost;
  rec.final_cost = step.final_cost;
}

void Server::WorkerLoop() {
  std::unique_lock<std::mutex> lock(mu_);
  while (true) {
    // The idle session whose oldest pending frame arrived first.
    Session* next = nullptr;
    work_cv_.wait(lock, [&] {
      if (stop_) return true;
      next = nullptr;
      for (const auto& s : sessions_) {
        if (s->busy || s->pending.empty()) continue;
        if (next == nullptr || s->pending.front().arrival < next->pending.front().arrival) next = s.get();
      }
      return next != nullptr;
    });
    if (stop_) return;
    Session& s = *next;
    s.busy = true;
    Pending p = std::move(s.pending.front());
    s.pending.pop_front();
    const Clock::time_point arrival = p.arrival;
    lock.unlock();
    IncrementRecord rec;
    std::string error;
    try {
      RunIncrement(s, std::move(p), rec);
    } catch (const std::exception& e) {
      error = e.what();
    }
    const auto done = Clock::now();
    const int64_t bytes = s.hip.DeviceBytes();
    lock.lock();
    s.records.push_back(rec);
    s.latency.push_back(Seconds(arrival, done));
    s.errors.push_back(error);
    if (bytes > 0) s.device_bytes = bytes;
    s.busy = false;
    done_cv_.notify_all();
    // Another frame of this session may be waiting for it to become idle.
    work_cv_.notify_one();
  }
}

ServedResult RunServedBenchmark(const ServedBenchmark& o) {
  if (o.sessions < 1 || o.rate_hz <= 0.0 || o.duration_s <= 0.0) {
    throw std::invalid_argument("served benchmark needs sessions, a rate and a duration");
  }
  std::vector<Workload> wls(static_cast<size_t>(o.sessions));
  {
    std::vector<std::thread> gen;
    std::atomic<int> next{0};
    const int threads = std::max(1, std::min(o.sessions, 8));
    for (int t = 0; t < threads; ++t) {
      gen.emplace_back([&] {
        for (int i = next++; i < o.sessions; i = next++) {
          wls[static_cast<size_t>(i)] = MakeWorkload("served" + std::to_string(i), o.seed0);
        }
      });
    }
    for (std::thread& t : gen) t.join();
  }
  // Poisson arrivals per session: exponential gaps, seeded per session.
  struct Arrival {
    double t;
    int session;
    int frame;
  };
  std::vector<Arrival> arrivals;
  for (int i = 0; i < o.sessions; ++i) {
    Rng rng(o.seed0 * 7919ULL + static_cast<uint64_t>(i));
    double t = 0.0;
    const auto frames = static_cast<int>(wls[static_cast<size_t>(i)].scans.size());
    for (int f = 1; f < frames; ++f) {
      t += -std::log(1.0 - rng.Uniform()) / o.rate_hz;
      if (t >= o.duration_s) break;
      arrivals.push_back({t, i, f});
    }
  }
  std::sort(arrivals.begin(), arrivals.end(),
            [](const Arrival& a, const Arrival& b) { return a.t != b.t ? a.t < b.t : a.session < b.session; });
  ServerOptions so;
  so.workers = o.workers;
  Server server(so);
  for (int i = 0; i < o.sessions; ++i) {
    const Workload& wl = wls[static_cast<size_t>(i)];
    SolverOptions step = wl.solver;
    step.max_iterations = wl.iterations_per_increment;
    server.OpenSession(wl.map, wl.problem, step, wl.dataset_poses[0], wl.scans[0]);
  }
  std::vector<std::pair<int, int64_t>> tickets;
  tickets.reserve(arrivals.size());
  const auto start = Clock::now();
  for (const Arrival& a : arrivals) {
    const auto when = start + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(a.t));
    std::this_thread::sleep_until(when);
    const Workload& wl = wls[static_cast<size_t>(a.session)];
    const int64_t ticket = server.Submit(a.session, wl.scans[static_cast<size_t>(a.frame)],
                                         wl.odometry[static_cast<size_t>(a.frame) - 1], when);
    tickets.emplace_back(a.session, ticket);
  }
  ServedResult r;
  std::vector<double> lat;
  double build = 0.0;
  double solve = 0.0;
  for (const auto& [sid, ticket] : tickets) {
    const IncrementRecord rec = server.Wait(sid, ticket);
    lat.push_back(1e3 * server.Latency(sid, ticket));
    build += rec.build_seconds;
    solve += rec.solve_seconds;
    r.lm_iterations += rec.iterations;
  }
  const auto end = Clock::now();
  r.increments = static_cast<int64_t>(tickets.size());
  r.wall_seconds = Seconds(start, end);
  r.increments_per_second = r.increments / std::max(r.wall_seconds, 1e-9);
  r.latency_p50_ms = Percentile(lat, 0.50);
  r.latency_p99_ms = Percentile(lat, 0.99);
  double sum = 0.0;
  for (const double v : lat) sum += v;
  r.latency_mean_ms = lat.empty() ? 0.0 : sum / static_cast<double>(lat.size());
  r.build_mean_ms = r.increments > 0 ? 1e3 * build / static_cast<double>(r.increments) : 0.0;
  r.solve_mean_ms = r.increments > 0 ? 1e3 * solve / static_cast<double>(r.increments) : 0.0;
  r.device_bytes = server.DeviceBytes();
  return r;
}

BatchResult RunBatchBenchmark(const BatchBenchmark& o) {
  if (o.instances < 1 || o.workers < 1) throw std::invalid_argument("batch benchmark needs instances and workers");
  std::vector<std::unique_ptr<Problem>> problems(static_cast<size_t>(o.instances));
  {
    std::vector<std::thread> gen;
    std::atomic<int> next{0};
    const int threads = std::max(1, std::min(o.instances, 8));
    for (int t = 0; t < threads; ++t) {
      gen.emplace_back([&] {
        for (int i = next++; i < o.instances; i = next++) {
          const Workload wl = MakeWorkload("batch" + std::to_string(i), o.seed0);
          problems[static_cast<size_t>(i)] = std::make_unique<Problem>(
              wl.map, wl.initial, wl.scans, wl.odometry, wl.problem, wl.relations);
        }
      });
    }
    for (std::thread& t : gen) t.join();
  }
  SolverOptions opts;
  opts.backend = SolverBackend::kHip;
  opts.max_iterations = o.iterations;
  opts.step_tolerance = 0.0;
  opts.lambda_init = 1.0;
  opts.lambda_factor = 1.0;
  BatchResult r;
  for (const auto& p : problems) r.unknowns += p->num_parameters();
  std::atomic<int> next{0};
  std::atomic<int64_t> iters{0};
  std::vector<std::string> errors(static_cast<size_t>(o.instances));
  const auto start = Clock::now();
  std::vector<std::thread> pool;
  for (int w = 0; w < std::min(o.workers, o.instances); ++w) {
    pool.emplace_back([&] {
      HipSession session;
      for (int i = next++; i < o.instances; i = next++) {
        try {
          const SolveResult res = session.Solve(*problems[static_cast<size_t>(i)], opts);
          iters += res.iterations;
        } catch (const std::exception& e) {
          errors[static_cast<size_t>(i)] = e.what();
        }
      }
    });
  }
  for (std::thread& t : pool) t.join();
  r.wall_seconds = Seconds(start, Clock::now());
  for (const std::string& e : errors) {
    if (!e.empty()) throw std::runtime_error("batch instance failed: " + e);
  }
  r.lm_iterations = iters;
  r.lm_iterations_per_second = static_cast<double>(r.lm_iterations) / std::max(r.wall_seconds, 1e-9);
  return r;
}

}  // namespace sdf_slam
