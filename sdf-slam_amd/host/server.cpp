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

/// Device memory reserved for a session before its first solve (its
/// context does not exist yet): per-node grid arrays and a first epoch's
/// factor storage. After a solve the context's exact bytes count instead.
int64_t AdmissionEstimate(const GridMap& map) {
  return (64LL << 20) + 4096LL * map.num_nodes();
}

/// Step options under which two sessions' increments give the same results
/// as each alone when solved in one batch (every field SolveLm reads).
bool SameStep(const SolverOptions& a, const SolverOptions& b) {
  return a.max_iterations == b.max_iterations && a.step_tolerance == b.step_tolerance &&
         a.lambda_init == b.lambda_init && a.lambda_factor == b.lambda_factor &&
         a.reject_worse_steps == b.reject_worse_steps && a.marquardt_scaling == b.marquardt_scaling &&
         a.trust_region == b.trust_region;
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
  if (options_.max_batch < 1 || options_.pipelines < 1 || options_.max_in_flight < 0) {
    throw std::invalid_argument("server needs max_batch >= 1, pipelines >= 1 and max_in_flight >= 0");
  }
  for (int w = 0; w < options_.workers; ++w) preparers_.emplace_back([this] { PrepareLoop(); });
  for (int w = 0; w < options_.pipelines; ++w) solvers_.emplace_back([this] { SolveLoop(); });
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lock(mu_);
    stop_ = true;
  }
  work_cv_.notify_all();
  for (std::thread& t : preparers_) t.join();
  // Solvers drain what the preparers queued before they stopped.
  {
    std::lock_guard<std::mutex> lock(mu_);
    stop_solvers_ = true;
  }
  solve_cv_.notify_all();
  for (std::thread& t : solvers_) t.join();
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
  if (estimate > options_.device_budget_bytes) {
    throw std::runtime_error("admission refused: the session's " + std::to_string(estimate) +
                             " bytes of device memory exceed the budget of " +
                             std::to_string(options_.device_budget_bytes));
  }
  std::unique_lock<std::mutex> lock(mu_);
  if (waiting_.empty() && Used() + estimate > options_.device_budget_bytes) TrimIdle(lock);
  Session* raw = s.get();
  sessions_.push_back(std::move(s));
  const int id = static_cast<int>(sessions_.size()) - 1;
  waiting_.push_back(raw);
  AdmitWaiting();
  if (raw->admitted && AppendApplies(raw->problem, false)) {
    // The session's device context is made now, over its first frame, so
    // that its increments append frames (no context setup on the first).
    raw->busy = true;
    lock.unlock();
    int64_t bytes = 0;
    try {
      Problem first(raw->map, raw->estimated, raw->scans, {}, raw->problem);
      (void)raw->hip.Load(first, raw->step);
      bytes = raw->hip.DeviceBytes();
    } catch (const std::exception&) {
      raw->hip.Release();  // the first increment builds it instead
    }
    lock.lock();
    raw->device_bytes = std::max(raw->device_bytes, bytes);
    raw->busy = false;
    work_cv_.notify_all();
    done_cv_.notify_all();
  }
  return id;
}

int64_t Server::Used() const {
  int64_t used = 0;
  for (const auto& o : sessions_) {
    if (o->admitted && !o->closed) used += o->device_bytes;
  }
  return used;
}

void Server::AdmitWaiting() {
  bool any = false;
  while (!waiting_.empty() && Used() + waiting_.front()->device_bytes <= options_.device_budget_bytes) {
    waiting_.front()->admitted = true;
    waiting_.pop_front();
    any = true;
  }
  if (any) work_cv_.notify_all();
}

void Server::TrimIdle(std::unique_lock<std::mutex>& lock) {
  std::vector<Session*> idle;
  for (const auto& o : sessions_) {
    if (o->admitted && !o->closed && !o->busy && o->device_bytes > 0) {
      o->busy = true;  // keeps the scheduler off the session meanwhile
      idle.push_back(o.get());
    }
  }
  lock.unlock();
  std::vector<int64_t> bytes(idle.size(), 0);
  for (size_t i = 0; i < idle.size(); ++i) {
    try {
      idle[i]->hip.Trim();
      bytes[i] = idle[i]->hip.DeviceBytes();
    } catch (const std::exception&) {
      bytes[i] = 0;  // keep the previous count
    }
  }
  lock.lock();
  for (size_t i = 0; i < idle.size(); ++i) {
    if (bytes[i] > 0) idle[i]->device_bytes = bytes[i];
    idle[i]->busy = false;
  }
  done_cv_.notify_all();
  work_cv_.notify_all();
}

void Server::CloseSession(int session) {
  std::unique_lock<std::mutex> lock(mu_);
  Session& s = *sessions_.at(static_cast<size_t>(session));
  if (s.closing) throw std::invalid_argument("session already closed");
  s.closing = true;
  if (!s.admitted) {
    // Still queued: nothing ran; its frames are dropped.
    waiting_.erase(std::remove(waiting_.begin(), waiting_.end(), &s), waiting_.end());
    while (!s.pending.empty()) {
      s.pending.pop_front();
      s.records.push_back({});
      s.latency.push_back(0.0);
      s.errors.push_back("session closed before it was admitted");
    }
    s.closed = true;
    done_cv_.notify_all();
    return;
  }
  done_cv_.wait(lock, [&] { return !s.busy && s.pending.empty(); });
  s.busy = true;
  lock.unlock();
  s.hip.Release();
  lock.lock();
  s.busy = false;
  s.closed = true;
  s.device_bytes = 0;
  s.scans.clear();
  s.scans.shrink_to_fit();
  s.odometry.clear();
  s.odometry.shrink_to_fit();
  AdmitWaiting();
  done_cv_.notify_all();
}

Server::SessionInfo Server::Info(int session) {
  std::lock_guard<std::mutex> lock(mu_);
  const Session& s = *sessions_.at(static_cast<size_t>(session));
  SessionInfo i;
  i.admitted = s.admitted;
  i.closed = s.closed;
  i.device_bytes = s.device_bytes;
  i.submitted = s.submitted;
  i.finished = static_cast<int64_t>(s.records.size());
  return i;
}

int64_t Server::Submit(int session, const Scan& scan, const Pose2& odometry, Clock::time_point arrival) {
  int64_t ticket = 0;
  {
    std::lock_guard<std::mutex> lock(mu_);
    Session& s = *sessions_.at(static_cast<size_t>(session));
    if (s.closing) throw std::invalid_argument("session is closed");
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
  return Used();
}

void Server::Prepare(Prepared& item, Pending& frame) {
  Session& s = *item.session;
  IncrementRecord& rec = item.record;
  try {
    s.estimated.push_back(s.estimated.back().Compose(frame.odometry));
    s.scans.push_back(std::move(frame.scan));
    s.odometry.push_back(frame.odometry);
    rec.frames = static_cast<int>(s.estimated.size());
    const auto t0 = Clock::now();
    if (s.hip.HasContext() && AppendApplies(s.problem, false)) {
      // The context extends its own problem by the frame (rows, odometry and
      // normals built on the device). A failed append leaves the context's
      // problem undefined: it is dropped and the problem rebuilt below.
      try {
        item.ctx = s.hip.Append({&s.scans.back()}, {frame.odometry}, s.problem, s.map, s.estimated);
        rec.build_seconds = Seconds(t0, Clock::now());
        return;
      } catch (const std::exception&) {
        s.hip.Release();
      }
    }
    // With device normals, increments after the first take their eikonal
    // normals from the device inside the batched solve (no host kNN, no GPU
    // call from this thread).
    const bool deferred = options_.device_normals && s.hip.HasContext() && DeviceNormalsApply(s.problem);
    item.problem = std::make_unique<Problem>(s.map, s.estimated, s.scans, s.odometry, s.problem,
                                             std::vector<RelationMeasurement>(),
                                             deferred ? NodeNormalsFn(DeferredNodeNormals{}) : NodeNormalsFn());
    rec.build_seconds = Seconds(t0, Clock::now());
    item.ctx = s.hip.Load(*item.problem, s.step, deferred ? &s.problem.normals : nullptr);
  } catch (const std::exception& e) {
    item.error = e.what();
  }
}

void Server::SolveBatch(std::vector<std::unique_ptr<Prepared>>& batch) {
  const auto nl = static_cast<int>(batch.size());
  // The sessions of a batch share their step options (SolveLoop).
  const SolverOptions& step = batch[0]->session->step;
  const sdfslam_solver_options copts = ToC(step);
  std::vector<sdfslam_hip_context*> ctxs(static_cast<size_t>(nl));
  std::vector<sdfslam_solve_result> results(static_cast<size_t>(nl));
  std::vector<std::vector<double>> hist(static_cast<size_t>(nl),
                                        std::vector<double>(static_cast<size_t>(std::max(0, step.max_iterations)) + 1));
  std::vector<int32_t> status(static_cast<size_t>(nl), SDFSLAM_OK);
  for (int j = 0; j < nl; ++j) {
    ctxs[static_cast<size_t>(j)] = batch[static_cast<size_t>(j)]->ctx;
    results[static_cast<size_t>(j)].cost_history = hist[static_cast<size_t>(j)].data();
    results[static_cast<size_t>(j)].cost_history_capacity = static_cast<int32_t>(hist[static_cast<size_t>(j)].size());
  }
  const auto t0 = Clock::now();
  for (auto& item : batch) item->record.solve_wait_seconds = Seconds(item->prepared, t0);
  sdfslam_hip_batch_info info{};
  const int st = sdfslam_hip_batch_solve(ctxs.data(), nl, &copts, results.data(), status.data(), &info);
  const double solve_s = Seconds(t0, Clock::now());
  if (st != SDFSLAM_OK) {
    const std::string msg = std::string(sdfslam_hip_status_string(st)) + ": " + sdfslam_hip_last_error();
    for (auto& item : batch) item->error = msg;
    return;
  }
  // Collect the states back into the sessions, on this thread (its GPU
  // stream has just run the batch; other threads' streams may be busy).
  const auto t_collect = Clock::now();
  for (int j = 0; j < nl; ++j) {
    Prepared& item = *batch[static_cast<size_t>(j)];
    Session& s = *item.session;
    try {
      SolveResult r;
      if (item.problem != nullptr) {
        r = s.hip.Collect(*item.problem, results[static_cast<size_t>(j)], status[static_cast<size_t>(j)],
                          hist[static_cast<size_t>(j)]);
        s.estimated = item.problem->poses();
        s.map = item.problem->map();
      } else {
        r = s.hip.CollectState(s.map, s.estimated, results[static_cast<size_t>(j)],
                               status[static_cast<size_t>(j)], hist[static_cast<size_t>(j)]);
      }
      item.record.iterations = r.iterations;
      item.record.initial_cost = r.initial_cost;
      item.record.final_cost = r.final_cost;
      item.record.solve_seconds = solve_s;
    } catch (const std::exception& e) {
      item.error = e.what();
    }
  }
  const double collect_s = Seconds(t_collect, Clock::now());
  std::lock_guard<std::mutex> lock(stats_mu_);
  stats_.batches += 1;
  stats_.increments += nl;
  stats_.graph_captures += info.graph_captures;
  stats_.epochs_rebuilt += info.epochs_rebuilt;
  stats_.solve_seconds += solve_s;
  stats_.collect_seconds += collect_s;
  stats_.capture_seconds += info.capture_seconds;
  stats_.analysis_seconds += info.analysis_seconds;
}

void Server::Finish(Prepared& item, Clock::time_point done, int64_t device_bytes) {
  Session& s = *item.session;
  s.records.push_back(item.record);
  s.latency.push_back(Seconds(item.arrival, done));
  s.errors.push_back(item.error);
  if (device_bytes > 0) s.device_bytes = device_bytes;  // exact bytes replace the reservation
  s.busy = false;
  --in_flight_;
  AdmitWaiting();
}

void Server::SetMaxInFlight(int max_in_flight) {
  if (max_in_flight < 0) throw std::invalid_argument("max_in_flight must be >= 0");
  {
    std::lock_guard<std::mutex> lock(mu_);
    options_.max_in_flight = max_in_flight;
  }
  work_cv_.notify_all();
}

ServerStats Server::Stats() {
  std::lock_guard<std::mutex> lock(stats_mu_);
  return stats_;
}

void Server::PrepareLoop() {
  std::unique_lock<std::mutex> lock(mu_);
  while (true) {
    Session* pick = nullptr;
    work_cv_.wait(lock, [&] {
      if (stop_) return true;
      pick = nullptr;
      if (options_.max_in_flight > 0 && in_flight_ >= options_.max_in_flight) return false;
      // The idle session with the oldest waiting frame.
      for (const auto& s : sessions_) {
        if (s->admitted && !s->busy && !s->pending.empty() &&
            (pick == nullptr || s->pending.front().arrival < pick->pending.front().arrival)) {
          pick = s.get();
        }
      }
      return pick != nullptr;
    });
    if (stop_) return;
    pick->busy = true;
    ++in_flight_;
    Pending frame = std::move(pick->pending.front());
    pick->pending.pop_front();
    lock.unlock();
    auto item = std::make_unique<Prepared>();
    item->session = pick;
    item->arrival = frame.arrival;
    const auto t0 = Clock::now();
    item->record.queue_seconds = Seconds(frame.arrival, t0);
    Prepare(*item, frame);
    item->prepared = Clock::now();
    item->record.prepare_seconds = Seconds(t0, item->prepared);
    {
      std::lock_guard<std::mutex> stats_lock(stats_mu_);
      stats_.prepare_seconds += item->record.prepare_seconds;
    }
    lock.lock();
    if (!item->error.empty()) {
      Finish(*item, Clock::now(), 0);
      done_cv_.notify_all();
      work_cv_.notify_all();
    } else {
      prepared_.push_back(std::move(item));
      solve_cv_.notify_one();
    }
  }
}

void Server::SolveLoop() {
  std::unique_lock<std::mutex> lock(mu_);
  while (true) {
    solve_cv_.wait(lock, [&] { return stop_solvers_ || !prepared_.empty(); });
    if (prepared_.empty()) return;  // stopping, nothing left to solve
    // The oldest prepared increment and every later one with the same step
    // options, in preparation order, up to max_batch; the others wait for
    // the next batch.
    std::vector<std::unique_ptr<Prepared>> batch;
    for (auto it = prepared_.begin();
         it != prepared_.end() && static_cast<int>(batch.size()) < options_.max_batch;) {
      if (batch.empty() || SameStep(batch[0]->session->step, (*it)->session->step)) {
        batch.push_back(std::move(*it));
        it = prepared_.erase(it);
      } else {
        ++it;
      }
    }
    if (!prepared_.empty()) solve_cv_.notify_one();
    lock.unlock();
    SolveBatch(batch);
    const auto done = Clock::now();
    std::vector<int64_t> bytes(batch.size());
    for (size_t i = 0; i < batch.size(); ++i) bytes[i] = batch[i]->session->hip.DeviceBytes();
    lock.lock();
    for (size_t i = 0; i < batch.size(); ++i) Finish(*batch[i], done, bytes[i]);
    done_cv_.notify_all();
    // Sessions of this batch may have more frames waiting.
    work_cv_.notify_all();
  }
}

ServedResult RunServedBenchmark(const ServedBenchmark& o) {
  if (o.sessions < 1 || o.rate_hz <= 0.0 || o.duration_s <= 0.0 || o.start_frame < 1) {
    throw std::invalid_argument("served benchmark needs sessions, a rate, a duration and start_frame >= 1");
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
    for (int f = o.start_frame; f < frames; ++f) {
      t += -std::log(1.0 - rng.Uniform()) / o.rate_hz;
      if (t >= o.duration_s) break;
      arrivals.push_back({t, i, f});
    }
  }
  std::sort(arrivals.begin(), arrivals.end(),
            [](const Arrival& a, const Arrival& b) { return a.t != b.t ? a.t < b.t : a.session < b.session; });
  ServerOptions so;
  so.workers = o.workers;
  so.max_batch = o.max_batch;
  so.pipelines = o.pipelines;
  so.max_in_flight = o.max_in_flight;
  so.device_normals = o.device_normals;
  Server server(so);
  for (int i = 0; i < o.sessions; ++i) {
    const Workload& wl = wls[static_cast<size_t>(i)];
    SolverOptions step = wl.solver;
    step.max_iterations = wl.iterations_per_increment;
    server.OpenSession(wl.map, wl.problem, step, wl.dataset_poses[0], wl.scans[0]);
  }
  ServerStats before;
  if (o.start_frame > 1) {
    // Sessions advanced to start_frame before the clock starts.
    server.SetMaxInFlight(0);
    std::vector<std::pair<int, int64_t>> warm;
    const auto now = Clock::now();
    for (int f = 1; f < o.start_frame; ++f) {
      for (int i = 0; i < o.sessions; ++i) {
        const Workload& wl = wls[static_cast<size_t>(i)];
        if (f >= static_cast<int>(wl.scans.size())) continue;
        warm.emplace_back(i, server.Submit(i, wl.scans[static_cast<size_t>(f)],
                                           wl.odometry[static_cast<size_t>(f) - 1], now));
      }
    }
    for (const auto& [sid, ticket] : warm) (void)server.Wait(sid, ticket);
    server.SetMaxInFlight(o.max_in_flight);
    before = server.Stats();
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
  std::vector<double> lat_after_first;
  double first_sum = 0.0;
  int64_t first_count = 0;
  double build = 0.0;
  double solve = 0.0;
  std::vector<double> queue_ms, prepare_ms, wait_ms, solve_ms;
  for (const auto& [sid, ticket] : tickets) {
    const IncrementRecord rec = server.Wait(sid, ticket);
    queue_ms.push_back(1e3 * rec.queue_seconds);
    prepare_ms.push_back(1e3 * rec.prepare_seconds);
    wait_ms.push_back(1e3 * rec.solve_wait_seconds);
    solve_ms.push_back(1e3 * rec.solve_seconds);
    lat.push_back(1e3 * server.Latency(sid, ticket));
    if (ticket == 1) {
      first_sum += lat.back();
      ++first_count;
    } else {
      lat_after_first.push_back(lat.back());
    }
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
  r.queue_p99_ms = Percentile(queue_ms, 0.99);
  r.prepare_p99_ms = Percentile(prepare_ms, 0.99);
  r.solve_wait_p99_ms = Percentile(wait_ms, 0.99);
  r.solve_p99_ms = Percentile(solve_ms, 0.99);
  r.latency_p50_after_first_ms = Percentile(lat_after_first, 0.50);
  r.latency_p99_after_first_ms = Percentile(lat_after_first, 0.99);
  r.first_increment_mean_ms = first_count > 0 ? first_sum / static_cast<double>(first_count) : 0.0;
  double sum = 0.0;
  for (const double v : lat) sum += v;
  r.latency_mean_ms = lat.empty() ? 0.0 : sum / static_cast<double>(lat.size());
  r.build_mean_ms = r.increments > 0 ? 1e3 * build / static_cast<double>(r.increments) : 0.0;
  r.solve_mean_ms = r.increments > 0 ? 1e3 * solve / static_cast<double>(r.increments) : 0.0;
  r.device_bytes = server.DeviceBytes();
  ServerStats st = server.Stats();
  st.batches -= before.batches;
  st.increments -= before.increments;
  st.graph_captures -= before.graph_captures;
  st.epochs_rebuilt -= before.epochs_rebuilt;
  st.prepare_seconds -= before.prepare_seconds;
  st.solve_seconds -= before.solve_seconds;
  st.collect_seconds -= before.collect_seconds;
  st.capture_seconds -= before.capture_seconds;
  st.analysis_seconds -= before.analysis_seconds;
  const double nb = static_cast<double>(std::max<int64_t>(st.batches, 1));
  r.batches = st.batches;
  r.mean_batch = static_cast<double>(st.increments) / nb;
  r.prepare_batch_ms = 1e3 * st.prepare_seconds / nb;
  r.solve_batch_ms = 1e3 * st.solve_seconds / nb;
  r.collect_batch_ms = 1e3 * st.collect_seconds / nb;
  r.capture_batch_ms = 1e3 * st.capture_seconds / nb;
  r.analysis_batch_ms = 1e3 * st.analysis_seconds / nb;
  r.epochs_rebuilt = st.epochs_rebuilt;
  return r;
}

namespace {

/// The batched benchmark's instances (built before any clock starts).
std::vector<std::unique_ptr<Problem>> BuildBatchInstances(const BatchBenchmark& o) {
  std::vector<std::unique_ptr<Problem>> problems(static_cast<size_t>(o.instances));
  std::vector<std::thread> gen;
  std::atomic<int> next{0};
  const int threads = std::max(1, std::min(o.instances, 8));
  for (int t = 0; t < threads; ++t) {
    gen.emplace_back([&] {
      for (int i = next++; i < o.instances; i = next++) {
        const Workload wl = MakeWorkload("batch" + std::to_string(i), o.seed0);
        problems[static_cast<size_t>(i)] =
            std::make_unique<Problem>(wl.map, wl.initial, wl.scans, wl.odometry, wl.problem, wl.relations);
      }
    });
  }
  for (std::thread& t : gen) t.join();
  return problems;
}

/// Times `iterations` LM iterations of every instance: instances largest
/// first, cut into batched solves of at most `max_batch`, which `workers`
/// host threads run concurrently on their own streams. The clock covers
/// device setup (one context per instance) and the iterations.
BatchResult TimeBatches(const std::vector<std::unique_ptr<Problem>>& problems, int iterations, int max_batch,
                        int workers) {
  const auto instances = static_cast<int>(problems.size());
  SolverOptions opts;
  opts.backend = SolverBackend::kHip;
  opts.max_iterations = iterations;
  opts.step_tolerance = 0.0;
  opts.lambda_init = 1.0;
  opts.lambda_factor = 1.0;
  const sdfslam_solver_options copts = ToC(opts);
  BatchResult r;
  for (const auto& p : problems) r.unknowns += p->num_parameters();
  // Largest first, so each batch holds instances of similar size.
  std::vector<int> order(static_cast<size_t>(instances));
  for (int i = 0; i < instances; ++i) order[static_cast<size_t>(i)] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return problems[static_cast<size_t>(a)]->num_parameters() > problems[static_cast<size_t>(b)]->num_parameters();
  });
  const int batches = (instances + max_batch - 1) / max_batch;
  std::atomic<int> next_batch{0};
  std::atomic<int64_t> iters{0};
  std::mutex err_mu;
  std::string error;
  // Contexts live until the clock stops (freeing device memory waits for the
  // whole device).
  std::vector<std::unique_ptr<HipSession>> sessions(static_cast<size_t>(instances));
  const auto start = Clock::now();
  auto body = [&] {
    for (int b = next_batch++; b < batches; b = next_batch++) {
      try {
        const int first = b * max_batch;
        const int last = std::min(instances, first + max_batch);
        const int nb = last - first;
        std::vector<sdfslam_hip_context*> ctxs(static_cast<size_t>(nb));
        for (int q = first; q < last; ++q) {
          const int i = order[static_cast<size_t>(q)];
          sessions[static_cast<size_t>(i)] = std::make_unique<HipSession>();
          ctxs[static_cast<size_t>(q - first)] =
              sessions[static_cast<size_t>(i)]->Load(*problems[static_cast<size_t>(i)], opts);
        }
        std::vector<sdfslam_solve_result> res(static_cast<size_t>(nb));
        std::vector<int32_t> status(static_cast<size_t>(nb), SDFSLAM_OK);
        const int st = sdfslam_hip_batch_solve(ctxs.data(), nb, &copts, res.data(), status.data(), nullptr);
        if (st != SDFSLAM_OK) {
          throw std::runtime_error(std::string("batch solve failed: ") + sdfslam_hip_last_error());
        }
        for (int j = 0; j < nb; ++j) {
          if (status[static_cast<size_t>(j)] != SDFSLAM_OK) throw std::runtime_error("batch instance failed");
          iters += res[static_cast<size_t>(j)].iterations;
        }
      } catch (const std::exception& e) {
        std::lock_guard<std::mutex> lock(err_mu);
        if (error.empty()) error = e.what();
      }
    }
  };
  std::vector<std::thread> pool;
  for (int w = 1; w < std::max(1, workers); ++w) pool.emplace_back(body);
  body();
  for (std::thread& t : pool) t.join();
  r.wall_seconds = Seconds(start, Clock::now());
  if (!error.empty()) throw std::runtime_error(error);
  r.lm_iterations = iters;
  r.lm_iterations_per_second = static_cast<double>(r.lm_iterations) / std::max(r.wall_seconds, 1e-9);
  return r;
}

}  // namespace

BatchResult RunBatchBenchmark(const BatchBenchmark& o) {
  if (o.instances < 1 || o.max_batch < 1) {
    throw std::invalid_argument("batch benchmark needs instances and max_batch >= 1");
  }
  return TimeBatches(BuildBatchInstances(o), o.iterations, o.max_batch, o.workers);
}

BatchTuneResult TuneBatchShapes(const BatchBenchmark& o, const std::vector<int>& max_batches,
                                const std::vector<int>& workers, int trial_iterations) {
  if (o.instances < 1 || max_batches.empty() || workers.empty() || trial_iterations < 1) {
    throw std::invalid_argument("autotune needs instances, candidate shapes and trial iterations >= 1");
  }
  for (const int mb : max_batches) {
    if (mb < 1) throw std::invalid_argument("autotune: max_batch candidates must be >= 1");
  }
  for (const int w : workers) {
    if (w < 1) throw std::invalid_argument("autotune: worker candidates must be >= 1");
  }
  const std::vector<std::unique_ptr<Problem>> problems = BuildBatchInstances(o);
  BatchTuneResult out;
  // One untimed trial first, so the process's first device setup is not
  // charged to the first candidate.
  (void)TimeBatches(problems, 1, max_batches.front(), workers.front());
  for (const int mb : max_batches) {
    for (const int w : workers) {
      const BatchResult r = TimeBatches(problems, trial_iterations, mb, w);
      out.trials.push_back({mb, w, r.lm_iterations_per_second});
      if (r.lm_iterations_per_second > out.best.lm_iterations_per_second) {
        out.best = {mb, w, r.lm_iterations_per_second};
      }
    }
  }
  return out;
}

}  // namespace sdf_slam
