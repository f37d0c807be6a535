// Host driver of the device backend and the C ABI of libsdfslam_hip.so
// (include/sdfslam_hip.h). Owns device memory, runs the symbolic analysis
// per pattern epoch, enqueues LM iterations on one stream and reads back
// only the small LM state between batches of iterations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <climits>
#include <cstring>
#include <functional>
#include <future>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "analysis.hpp"
#include "kernels.hpp"
#include "normals.hpp"
#include "sdfslam_hip.h"

using namespace sdfslam;

namespace {

thread_local std::string g_last_error;

struct Failure : std::runtime_error {
  int code;
  Failure(int c, const std::string& what) : std::runtime_error(what), code(c) {}
};

void HipCheck(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    const int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? SDFSLAM_ERR_OOM
                                                                                 : SDFSLAM_ERR_HIP;
    throw Failure(code, std::string(what) + ": " + hipGetErrorString(e));
  }
}

#define HIPC(x) HipCheck((x), #x)

class DeviceArena {
 public:
  DeviceArena() = default;
  DeviceArena(const DeviceArena&) = delete;
  DeviceArena& operator=(const DeviceArena&) = delete;
  ~DeviceArena() { Clear(); }
  template <typename T>
  T* Alloc(size_t n) {
    void* p = nullptr;
    HIPC(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    ptrs_.push_back(p);
    bytes_ += std::max<size_t>(n, 1) * sizeof(T);
    return static_cast<T*>(p);
  }
  void Clear() {
    for (void* p : ptrs_) (void)hipFree(p);
    ptrs_.clear();
    bytes_ = 0;
  }
  size_t bytes() const { return bytes_; }

 private:
  std::vector<void*> ptrs_;
  size_t bytes_{0};
};

/// Device memory of one layout (the problem, or one pattern epoch). One
/// allocation holds every array and is kept while later layouts fit (it
/// grows with 50 % slack; an outgrown allocation is kept until the pool is
/// released, since freeing device memory waits for the whole device);
/// host arrays are staged in pinned memory and uploaded with one copy per
/// contiguous run. The layout is laid out twice by the same code: once to
/// measure (Upload/Alloc return null), once to place.
class EpochPool {
 public:
  EpochPool() = default;
  EpochPool(const EpochPool&) = delete;
  EpochPool& operator=(const EpochPool&) = delete;
  ~EpochPool() { Release(); }

  void BeginMeasure() {
    measuring_ = true;
    off_ = 0;
    upload_end_ = 0;
  }
  /// Ends measuring and makes room for what was measured.
  void BeginPlace() {
    const size_t need = off_;
    const size_t staged = upload_end_;
    if (need > cap_) {
      if (dev_ != nullptr) retired_dev_.push_back(dev_);
      retired_bytes_ += cap_;
      dev_ = nullptr;
      cap_ = 0;
      const size_t want = need + need / 2;
      HIPC(hipMalloc(&dev_, want));
      cap_ = want;
    }
    if (staged > host_cap_) {
      if (host_ != nullptr) retired_host_.push_back(host_);
      host_ = nullptr;
      host_cap_ = 0;
      const size_t want = staged + staged / 2;
      HIPC(hipHostMalloc(&host_, want, hipHostMallocDefault));
      host_cap_ = want;
    }
    measuring_ = false;
    off_ = 0;
    runs_.clear();
  }
  /// Enqueues the copies of the staged arrays to the device on `s`. Every
  /// call of the library synchronizes its stream before it returns, so the
  /// copies are done before the staging memory is written again.
  void Copy(hipStream_t s) {
    for (const auto& r : runs_) {
      HIPC(hipMemcpyAsync(static_cast<char*>(dev_) + r.first, static_cast<char*>(host_) + r.first,
                          r.second - r.first, hipMemcpyHostToDevice, s));
    }
  }
  bool measuring() const { return measuring_; }
  template <typename T>
  T* Upload(const T* data, size_t n) {
    const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    const size_t at = Take(bytes);
    if (measuring_) {
      upload_end_ = std::max(upload_end_, at + bytes);
      return nullptr;
    }
    if (n > 0) std::memcpy(static_cast<char*>(host_) + at, data, n * sizeof(T));
    if (!runs_.empty() && runs_.back().second == at) {
      runs_.back().second = at + bytes;
    } else {
      runs_.push_back({at, at + bytes});
    }
    return reinterpret_cast<T*>(static_cast<char*>(dev_) + at);
  }
  template <typename T>
  T* Upload(const std::vector<T>& v) {
    return Upload(v.data(), v.size());
  }
  /// Room for `cap` elements of which the first `n` are uploaded.
  template <typename T>
  T* UploadPrefix(const T* data, size_t n, size_t cap) {
    const size_t bytes = std::max<size_t>(cap, 1) * sizeof(T);
    const size_t at = Take(bytes);
    if (measuring_) {
      upload_end_ = std::max(upload_end_, at + std::max<size_t>(n, 1) * sizeof(T));
      return nullptr;
    }
    if (n > 0) {
      std::memcpy(static_cast<char*>(host_) + at, data, n * sizeof(T));
      if (!runs_.empty() && runs_.back().second == at) {
        runs_.back().second = at + n * sizeof(T);
      } else {
        runs_.push_back({at, at + n * sizeof(T)});
      }
    }
    return reinterpret_cast<T*>(static_cast<char*>(dev_) + at);
  }
  template <typename T>
  T* Alloc(size_t n) {
    const size_t at = Take(std::max<size_t>(n, 1) * sizeof(T));
    return measuring_ ? nullptr : reinterpret_cast<T*>(static_cast<char*>(dev_) + at);
  }
  /// Device memory held, outgrown allocations included.
  size_t bytes() const { return cap_ + retired_bytes_; }
  /// Frees the outgrown allocations (no work of the pool's owner may be in
  /// flight; hipFree waits for the device).
  void Trim() {
    for (void* p : retired_dev_) (void)hipFree(p);
    for (void* p : retired_host_) (void)hipHostFree(p);
    retired_dev_.clear();
    retired_host_.clear();
    retired_bytes_ = 0;
  }
  void Release() {
    if (dev_ != nullptr) (void)hipFree(dev_);
    if (host_ != nullptr) (void)hipHostFree(host_);
    for (void* p : retired_dev_) (void)hipFree(p);
    for (void* p : retired_host_) (void)hipHostFree(p);
    retired_dev_.clear();
    retired_host_.clear();
    dev_ = host_ = nullptr;
    cap_ = host_cap_ = retired_bytes_ = 0;
  }

 private:
  size_t Take(size_t bytes) {
    const size_t at = off_;
    off_ += (bytes + 255) & ~static_cast<size_t>(255);
    return at;
  }
  bool measuring_{true};
  size_t off_{0}, upload_end_{0};
  void* dev_{nullptr};
  void* host_{nullptr};
  size_t cap_{0}, host_cap_{0}, retired_bytes_{0};
  std::vector<void*> retired_dev_, retired_host_;
  std::vector<std::pair<size_t, size_t>> runs_;
};

template <typename T>
std::vector<T> CopyArray(const T* p, int64_t n) {
  if (n <= 0) return {};
  if (p == nullptr) throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "null spec array");
  return std::vector<T>(p, p + n);
}

/// Owned copy of a problem view.
struct HostProblem {
  sdfslam_problem_view view{};
  std::vector<int32_t> node_to_active, pt_frame, eik_w, eik_h, od_i, od_j;
  std::vector<double> pt_x, pt_y, pt_delta, pt_sw, eik_nx, eik_ny, eik_sw, od_dx, od_dy, od_dth,
      od_sw;

  explicit HostProblem(const sdfslam_problem_view& v) {
    const sdfslam_grid& g = v.grid;
    if (g.nx < 2 || g.ny < 2 || !(g.dx > 0.0) || !(g.dy > 0.0)) {
      throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "grid needs nx, ny >= 2 and dx, dy > 0");
    }
    if (v.num_poses < 1 || v.num_points < 0 || v.num_eikonal < 0 || v.num_odom < 0) {
      throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "negative sizes or no poses");
    }
    const int64_t nn = static_cast<int64_t>(g.nx) * g.ny;
    if (nn > (1LL << 30)) throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "grid too large");
    node_to_active = CopyArray(v.node_to_active, nn);
    int count = 0;
    for (const int32_t c : node_to_active) {
      if (c >= 0) ++count;
      if (c >= v.num_active_nodes) throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "node column out of range");
    }
    if (count != v.num_active_nodes) {
      throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "num_active_nodes does not match node_to_active");
    }
    pt_frame = CopyArray(v.point_frame, v.num_points);
    pt_x = CopyArray(v.point_x, v.num_points);
    pt_y = CopyArray(v.point_y, v.num_points);
    pt_delta = CopyArray(v.point_expected_sdf, v.num_points);
    pt_sw = CopyArray(v.point_sqrt_weight, v.num_points);
    for (const int32_t f : pt_frame) {
      if (f < 0 || f >= v.num_poses) throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "point frame out of range");
    }
    eik_w = CopyArray(v.eik_w, v.num_eikonal);
    eik_h = CopyArray(v.eik_h, v.num_eikonal);
    eik_nx = CopyArray(v.eik_nx, v.num_eikonal);
    eik_ny = CopyArray(v.eik_ny, v.num_eikonal);
    eik_sw = CopyArray(v.eik_sqrt_weight, v.num_eikonal);
    for (int64_t e = 0; e < v.num_eikonal; ++e) {
      if (eik_w[e] < 0 || eik_h[e] < 0 || eik_w[e] + 1 >= g.nx || eik_h[e] + 1 >= g.ny) {
        throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "eikonal spec outside the grid");
      }
    }
    od_i = CopyArray(v.odom_i, v.num_odom);
    od_j = CopyArray(v.odom_j, v.num_odom);
    od_dx = CopyArray(v.odom_dx, v.num_odom);
    od_dy = CopyArray(v.odom_dy, v.num_odom);
    od_dth = CopyArray(v.odom_dtheta, v.num_odom);
    od_sw = CopyArray(v.odom_sqrt_weight, v.num_odom);
    for (int64_t o = 0; o < v.num_odom; ++o) {
      if (od_i[o] < 0 || od_j[o] < 0 || od_i[o] >= v.num_poses || od_j[o] >= v.num_poses ||
          od_i[o] == od_j[o]) {
        throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "odometry frames out of range");
      }
    }
    view = v;
    Refresh();
  }
  /// Points the view at the arrays and sets its counts (after appending).
  void Refresh() {
    view.num_points = static_cast<int64_t>(pt_frame.size());
    view.num_eikonal = static_cast<int64_t>(eik_w.size());
    view.num_odom = static_cast<int64_t>(od_i.size());
    view.node_to_active = node_to_active.data();
    view.point_frame = pt_frame.data();
    view.point_x = pt_x.data();
    view.point_y = pt_y.data();
    view.point_expected_sdf = pt_delta.data();
    view.point_sqrt_weight = pt_sw.data();
    view.eik_w = eik_w.data();
    view.eik_h = eik_h.data();
    view.eik_nx = eik_nx.data();
    view.eik_ny = eik_ny.data();
    view.eik_sqrt_weight = eik_sw.data();
    view.odom_i = od_i.data();
    view.odom_j = od_j.data();
    view.odom_dx = od_dx.data();
    view.odom_dy = od_dy.data();
    view.odom_dtheta = od_dth.data();
    view.odom_sqrt_weight = od_sw.data();
  }
  int n() const { return view.num_active_nodes + 3 * (view.num_poses - 1); }
  int64_t rows() const { return view.num_points + view.num_eikonal + 3 * view.num_odom; }
};

constexpr int kPhases = 5;

}  // namespace

/// Every call runs on a stream of the calling thread, so calls from
/// different threads never queue behind each other's work; a context has no
/// work in flight between calls, so any thread may drive it next. A thread
/// borrows one stream per device from a process-wide pool on first use and
/// returns it when it ends: streams are never destroyed while the process
/// runs (destroying one can disturb a graph capture on another thread), and
/// short-lived threads reuse the streams of earlier ones.
class StreamPool {
 public:
  static StreamPool& Get() {
    static StreamPool* pool = new StreamPool();  // lives until the process ends
    return *pool;
  }
  hipStream_t Borrow(int device) {
    {
      std::lock_guard<std::mutex> lock(mu_);
      for (size_t i = 0; i < free_.size(); ++i) {
        if (free_[i].first == device) {
          const hipStream_t s = free_[i].second;
          free_.erase(free_.begin() + static_cast<std::ptrdiff_t>(i));
          return s;
        }
      }
    }
    hipStream_t stream = nullptr;
    HIPC(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    return stream;
  }
  void Return(int device, hipStream_t stream) {
    std::lock_guard<std::mutex> lock(mu_);
    free_.emplace_back(device, stream);
  }

 private:
  std::mutex mu_;
  std::vector<std::pair<int, hipStream_t>> free_;
};

class ThreadStreams {
 public:
  ThreadStreams() = default;
  ThreadStreams(const ThreadStreams&) = delete;
  ThreadStreams& operator=(const ThreadStreams&) = delete;
  ~ThreadStreams() {
    for (const auto& [device, stream] : streams_) StreamPool::Get().Return(device, stream);
  }
  hipStream_t Get(int device) {
    for (const auto& [d, stream] : streams_) {
      if (d == device) return stream;
    }
    const hipStream_t stream = StreamPool::Get().Borrow(device);
    streams_.emplace_back(device, stream);
    return stream;
  }

 private:
  std::vector<std::pair<int, hipStream_t>> streams_;
};
thread_local ThreadStreams t_streams;
// A second stream per thread and device: the forward solves of the lower
// levels run on it beside the factorization of the upper ones.
thread_local ThreadStreams t_aux_streams;
// Two more streams per thread and device: a batched solve runs its
// contexts' deferred work (uploads, device normals) on them and on the
// thread's stream side by side.
constexpr int kPendingStreams = 3;
// Iterations as captured graphs (the default) or as direct launches
// (sdfslam_hip_set_graphs; for kernel profilers).
std::atomic<bool> g_graphs{true};
thread_local ThreadStreams t_pending_streams[kPendingStreams - 1];
thread_local std::vector<hipEvent_t> t_pending_events;

/// `n` events without timing in `ev` (created on first use, kept).
const hipEvent_t* OverlapEvents(std::vector<hipEvent_t>& ev, size_t n) {
  while (ev.size() < n) {
    hipEvent_t e = nullptr;
    HIPC(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    ev.push_back(e);
  }
  return ev.data();
}

using SteadyTime = std::chrono::steady_clock::time_point;
static SteadyTime Now() { return std::chrono::steady_clock::now(); }
static double SecondsSince(SteadyTime t0) { return std::chrono::duration<double>(Now() - t0).count(); }

/// Buffers of the batches a context leads (grown, never shrunk): the slot
/// table and launch plans with their pinned staging, and the gathered LM
/// states.
struct BatchSpace {
  int capacity{0};
  Slot* table{nullptr};
  Slot* slots_host{nullptr};
  LmState* states_dev{nullptr};
  LmState* states_host{nullptr};
  size_t plan_capacity{0};
  ListRef* plan{nullptr};
  ListRef* plan_host{nullptr};

  void Reserve(int k) {
    if (k <= capacity) return;
    // Outgrown buffers are kept until the context ends: freeing device
    // memory waits for the whole device.
    retired_dev.push_back(table);
    retired_dev.push_back(states_dev);
    retired_host.push_back(slots_host);
    retired_host.push_back(states_host);
    table = nullptr;
    states_dev = nullptr;
    slots_host = nullptr;
    states_host = nullptr;
    const int want = std::max(k, 2 * capacity);
    HIPC(hipMalloc(&table, sizeof(Slot) * static_cast<size_t>(want)));
    HIPC(hipHostMalloc(&slots_host, sizeof(Slot) * static_cast<size_t>(want), hipHostMallocDefault));
    HIPC(hipMalloc(&states_dev, sizeof(LmState) * static_cast<size_t>(want)));
    HIPC(hipHostMalloc(&states_host, sizeof(LmState) * static_cast<size_t>(want), hipHostMallocDefault));
    capacity = want;
  }
  void ReservePlans(size_t entries) {
    if (entries <= plan_capacity) return;
    retired_dev.push_back(plan);
    retired_host.push_back(plan_host);
    plan = nullptr;
    plan_host = nullptr;
    const size_t want = std::max(entries, 2 * plan_capacity) + 16;
    HIPC(hipMalloc(&plan, sizeof(ListRef) * want));
    HIPC(hipHostMalloc(&plan_host, sizeof(ListRef) * want, hipHostMallocDefault));
    plan_capacity = want;
  }
  std::vector<void*> retired_dev, retired_host;

  /// Frees the outgrown buffers (no batch led by the owner in flight).
  void Trim() {
    for (void* p : retired_dev) {
      if (p != nullptr) (void)hipFree(p);
    }
    for (void* p : retired_host) {
      if (p != nullptr) (void)hipHostFree(p);
    }
    retired_dev.clear();
    retired_host.clear();
  }

  /// Frees everything; errors are ignored so that this can run from a
  /// destructor.
  void Release() {
    retired_dev.push_back(table);
    retired_dev.push_back(states_dev);
    retired_dev.push_back(plan);
    retired_host.push_back(slots_host);
    retired_host.push_back(states_host);
    retired_host.push_back(plan_host);
    for (void* p : retired_dev) {
      if (p != nullptr) (void)hipFree(p);
    }
    for (void* p : retired_host) {
      if (p != nullptr) (void)hipHostFree(p);
    }
    retired_dev.clear();
    retired_host.clear();
    table = nullptr;
    slots_host = nullptr;
    states_dev = nullptr;
    states_host = nullptr;
    plan = nullptr;
    plan_host = nullptr;
    capacity = 0;
    plan_capacity = 0;
  }
};

struct sdfslam_hip_context {
  int device{0};
  hipStream_t stream{nullptr};  // the calling thread's stream, set at every call
  hipStream_t aux_stream{nullptr};  // the calling thread's second stream
  size_t solve_sync_words{0};        // of dw.solve_sync
  std::vector<hipEvent_t> overlap_events;  // level marks of the overlapped solves
  // Device work of a context update, deferred to the context's next call
  // (which runs it first, on its own stream): an update does host work only.
  std::vector<std::function<void(hipStream_t)>> pending;
  bool deferring{false};
  std::unique_ptr<HostProblem> hp;
  sdfslam_solver_options opts{};
  AnalysisOptions aopt;
  Analysis an;
  int pattern_epochs{0};
  bool short_epoch{false};  // the ending epoch lasted at most two iterations
  double analysis_seconds{0.0};  // host time spent building pattern epochs
  EpochPool base;   // problem specs, state and per-row work arrays
  EpochPool epoch;  // symbolic analysis and factor storage
  EpochPool normals;  // device eikonal normals: cos/sin, work
  // Point rows (the point specs) with room for appended frames, and the
  // list of rows that are scan points (for the device normals); rows_cap
  // rows and list entries fit before the layout has to grow.
  EpochPool rows;
  int64_t rows_cap{0}, list_cap{0};
  int32_t* pt_frame_d{nullptr};
  double *pt_x_d{nullptr}, *pt_y_d{nullptr}, *pt_delta_d{nullptr}, *pt_sw_d{nullptr};
  int32_t* scan_list_d{nullptr};
  std::vector<int32_t> scan_list;  // host copy: rows with expected SDF 0 (appended frames: their scan rows)
  // Staging of appended frames' raw scans (pinned) and device copies.
  EpochPool frames;
  DeviceArena hist;  // cost history
  DeviceProblem dp{};
  DeviceAnalysis da{};
  DeviceWork dw{};
  double* j_eik{nullptr};
  double* scalar{nullptr};
  int history_capacity{0};
  // Captured iteration: valid for one buffer generation and one set of LM options.
  hipGraphExec_t graph{nullptr};
  int graph_generation{-1};
  LmOptions graph_options{};
  int generation{0};
  Slot slot{};            // this context's problem as a launch slot
  Slot* slot_dev{nullptr};  // device copy (a table of one, in the epoch pool)
  LaunchShape shape;      // grid sizes of this context alone
  // Level schedule: fronts and update-matrix tiles per level, on the device
  // with their host offsets, for launch plans.
  const int32_t* level_sn_d{nullptr};
  const int32_t* tiles_d{nullptr};  // update-matrix tiles, level by level (pairs)
  std::vector<int64_t> level_off_h, tile_off_h;

  int max_f{0};
  hipEvent_t events[kPhases + 3]{};  // phase marks, then the point-row kernel pair
  double phase_ms[kPhases + 1]{};
  std::vector<hipEvent_t> step_events;  // pairs around panel steps (timed runs)
  int step_events_used{0};
  double panel_step_ms{0.0};
  int64_t panel_steps_timed{0};
  std::vector<hipEvent_t> gemm_events;  // pairs around trailing updates (timed runs)
  int gemm_events_used{0};
  double gemm_ms{0.0};
  int64_t gemms_timed{0};
  int timed_iterations{0};
  bool lm_ready{false};
  // An update's LM initialisation (InitLm with this lambda), run after its
  // deferred work by the context's next call; a batched solve, which
  // initialises every slot itself, drops it.
  bool init_pending{false};
  double init_lambda{0.0};
  BatchSpace batch;  // buffers of the batches this context leads
  // Speculative pattern epochs: the poses the current epoch's footprints
  // were computed at, how far each pose's points reach from it (metres),
  // and a re-analysis running on a host thread.
  std::vector<double> anchor_poses;
  std::vector<double> pose_reach;
  struct Speculation {
    std::future<Analysis> job;
    std::vector<double> map, poses;
    AnalysisOptions aopt;
  };
  std::unique_ptr<Speculation> spec;
  int speculative_epochs{0};

  ~sdfslam_hip_context() {
    spec.reset();  // waits for a running re-analysis (it reads the problem)
    batch.Release();
    if (graph != nullptr) (void)hipGraphExecDestroy(graph);

    for (hipEvent_t& e : events) {
      if (e != nullptr) (void)hipEventDestroy(e);
    }
    for (hipEvent_t e : step_events) (void)hipEventDestroy(e);
    for (hipEvent_t e : gemm_events) (void)hipEventDestroy(e);
    for (hipEvent_t e : overlap_events) (void)hipEventDestroy(e);
    epoch.Release();
    base.Release();
    normals.Release();
    rows.Release();
    frames.Release();
    hist.Clear();
  }
};

namespace {

LmOptions ToLm(const sdfslam_solver_options& o, int capacity) {
  LmOptions l{};
  l.max_iterations = o.max_iterations;
  l.history_capacity = capacity;
  l.step_tolerance = o.step_tolerance;
  l.lambda_factor = o.lambda_factor;
  l.reject_worse_steps = o.reject_worse_steps;
  l.marquardt_scaling = o.marquardt_scaling;
  l.trust_region = o.trust_region;
  return l;
}

/// Runs `work` on the context's stream now, or, while an update defers its
/// device work, at the start of the context's next call.
void Enqueue(sdfslam_hip_context& c, std::function<void(hipStream_t)> work) {
  if (c.deferring) {
    c.pending.push_back(std::move(work));
  } else {
    work(c.stream);
  }
}

/// Makes the calling thread's stream the context's stream and runs the
/// context's deferred work on it.
void InitLm(sdfslam_hip_context& c, double lambda_init);

/// Runs the context's deferred work on `s` (in order).
void RunPending(sdfslam_hip_context& c, hipStream_t s) {
  std::vector<std::function<void(hipStream_t)>> work;
  work.swap(c.pending);
  for (auto& w : work) w(s);
}

void Bind(sdfslam_hip_context& c) {
  HIPC(hipSetDevice(c.device));
  c.stream = t_streams.Get(c.device);
  c.aux_stream = nullptr;  // the calling thread's second stream, taken when needed
  RunPending(c, c.stream);
  if (c.init_pending) {
    c.init_pending = false;
    InitLm(c, c.init_lambda);
  }
}

/// How far each pose's points reach from it (metres), over rows [from, end)
/// (from > 0 extends the current values).
void ComputeReach(sdfslam_hip_context& c, size_t from) {
  const HostProblem& hp = *c.hp;
  if (from == 0) c.pose_reach.clear();
  c.pose_reach.resize(static_cast<size_t>(hp.view.num_poses), 0.0);
  for (size_t r = from; r < hp.pt_frame.size(); ++r) {
    double& reach = c.pose_reach[static_cast<size_t>(hp.pt_frame[r])];
    reach = std::max(reach, std::hypot(hp.pt_x[r], hp.pt_y[r]));
  }
}

/// Lays out the row pool for the problem's rows with room to append (half
/// again as many), uploading rows [0, upload_rows) and scan-list entries
/// [0, upload_list) from the host copy; the rest is written on the device.
void PlaceRows(sdfslam_hip_context& c, int64_t upload_rows, int64_t upload_list) {
  const HostProblem& hp = *c.hp;
  const auto np = static_cast<int64_t>(hp.pt_frame.size());
  const auto nl = static_cast<int64_t>(c.scan_list.size());
  c.rows_cap = np + np / 2 + 1024;
  c.list_cap = nl + nl / 2 + 1024;
  EpochPool& R = c.rows;
  const auto up = static_cast<size_t>(upload_rows);
  const auto cap = static_cast<size_t>(c.rows_cap);
  auto layout = [&]() {
    c.pt_frame_d = R.UploadPrefix(hp.pt_frame.data(), up, cap);
    c.pt_x_d = R.UploadPrefix(hp.pt_x.data(), up, cap);
    c.pt_y_d = R.UploadPrefix(hp.pt_y.data(), up, cap);
    c.pt_delta_d = R.UploadPrefix(hp.pt_delta.data(), up, cap);
    c.pt_sw_d = R.UploadPrefix(hp.pt_sw.data(), up, cap);
    c.scan_list_d = R.UploadPrefix(c.scan_list.data(), static_cast<size_t>(upload_list),
                                   static_cast<size_t>(c.list_cap));
  };
  R.BeginMeasure();
  layout();
  R.BeginPlace();
  layout();
  Enqueue(c, [&R](hipStream_t s) { R.Copy(s); });
}

/// The problem's scan rows for the device normals: the rows with expected
/// SDF 0, in row order.
void ComputeScanList(sdfslam_hip_context& c) {
  c.scan_list.clear();
  const HostProblem& hp = *c.hp;
  for (size_t r = 0; r < hp.pt_delta.size(); ++r) {
    if (hp.pt_delta[r] == 0.0) c.scan_list.push_back(static_cast<int32_t>(r));
  }
}

void AllocateBase(sdfslam_hip_context& c, const double* map_values, const double* poses) {
  const sdfslam_problem_view& v = c.hp->view;
  HostProblem& hp = *c.hp;
  DeviceProblem& P = c.dp;
  P.grid = {v.grid.nx, v.grid.ny, v.grid.min_x, v.grid.min_y, v.grid.dx, v.grid.dy};
  P.num_cells = (v.grid.nx - 1) * (v.grid.ny - 1);
  P.num_active = v.num_active_nodes;
  P.num_poses = v.num_poses;
  P.n = hp.n();
  P.num_points = v.num_points;
  P.num_eikonal = v.num_eikonal;
  P.num_odom = v.num_odom;
  std::vector<int32_t> active_to_node(static_cast<size_t>(v.num_active_nodes));
  for (size_t g = 0; g < hp.node_to_active.size(); ++g) {
    if (hp.node_to_active[g] >= 0) active_to_node[static_cast<size_t>(hp.node_to_active[g])] = static_cast<int32_t>(g);
  }
  P.huber_delta = v.huber_delta;
  P.smooth_gradient = v.smooth_gradient;
  P.smooth_gradient_step = v.smooth_gradient_step;
  EpochPool& A = c.base;
  DeviceWork& W = c.dw;
  const size_t nmap = static_cast<size_t>(v.grid.nx) * v.grid.ny;
  const size_t npose = 3 * static_cast<size_t>(v.num_poses);
  const auto np = static_cast<size_t>(v.num_points);
  auto layout = [&]() {
    P.node_to_active = A.Upload(hp.node_to_active);
    P.active_to_node = A.Upload(active_to_node);
    P.eik_w = A.Upload(hp.eik_w);
    P.eik_h = A.Upload(hp.eik_h);
    P.eik_nx = A.Upload(hp.eik_nx);
    P.eik_ny = A.Upload(hp.eik_ny);
    P.eik_sw = A.Upload(hp.eik_sw);
    P.od_i = A.Upload(hp.od_i);
    P.od_j = A.Upload(hp.od_j);
    P.od_dx = A.Upload(hp.od_dx);
    P.od_dy = A.Upload(hp.od_dy);
    P.od_dth = A.Upload(hp.od_dth);
    P.od_sw = A.Upload(hp.od_sw);
    W.map = A.Upload(map_values, nmap);
    W.poses = A.Upload(poses, npose);
    W.map_backup = A.Alloc<double>(nmap);
    W.poses_backup = A.Alloc<double>(npose);
    W.r_pt = A.Alloc<double>(np);
    W.jw = A.Alloc<double>(4 * np);
    W.jp = A.Alloc<double>(3 * np);
    W.cellkey = A.Alloc<uint32_t>(np);
    W.r_eik = A.Alloc<double>(static_cast<size_t>(v.num_eikonal));
    c.j_eik = A.Alloc<double>(3 * static_cast<size_t>(v.num_eikonal));
    W.od_r = A.Alloc<double>(3 * static_cast<size_t>(v.num_odom));
    W.od_ji = A.Alloc<double>(9 * static_cast<size_t>(v.num_odom));
    W.od_jj = A.Alloc<double>(9 * static_cast<size_t>(v.num_odom));
    W.seg_cell = A.Alloc<int32_t>(np);
    W.seg_part = A.Alloc<double>(14 * np);
    W.seg_jp = A.Alloc<double>(12 * np);
    W.seg_count = A.Alloc<int32_t>(static_cast<size_t>(v.num_poses));
    W.nn = A.Alloc<double>(10 * static_cast<size_t>(P.num_cells));
    W.rn = A.Alloc<double>(4 * static_cast<size_t>(P.num_cells));
    W.rhs = A.Alloc<double>(static_cast<size_t>(P.n));
    W.step = A.Alloc<double>(static_cast<size_t>(P.n));
    W.xnew = A.Alloc<double>(static_cast<size_t>(P.n));
    W.partials = A.Alloc<double>(3 * static_cast<size_t>(kPartialBlocks));
    W.state = A.Alloc<LmState>(1);
    W.pose_overflow = A.Alloc<int32_t>(static_cast<size_t>(v.num_poses));
    W.solve_tmp = A.Alloc<double>(static_cast<size_t>(P.n));
    c.scalar = A.Alloc<double>(1);
  };
  A.BeginMeasure();
  layout();
  A.BeginPlace();
  layout();
  P.pt_frame = c.pt_frame_d;
  P.pt_x = c.pt_x_d;
  P.pt_y = c.pt_y_d;
  P.pt_delta = c.pt_delta_d;
  P.pt_sw = c.pt_sw_d;
  int32_t* overflow = W.pose_overflow;
  LmState* state = W.state;
  const size_t poses_n = static_cast<size_t>(v.num_poses);
  Enqueue(c, [&A, overflow, state, poses_n](hipStream_t s) {
    A.Copy(s);
    HIPC(hipMemsetAsync(overflow, 0, sizeof(int32_t) * poses_n, s));
    HIPC(hipMemsetAsync(state, 0, sizeof(LmState), s));  // a zero LmState
  });
  const int want = std::max(c.opts.max_iterations, 0) + 1;
  if (W.cost_history == nullptr || want > c.history_capacity) {
    c.history_capacity = std::max(want, c.history_capacity);
    W.cost_history = c.hist.Alloc<double>(static_cast<size_t>(c.history_capacity));
  }
  LaunchShape& sh = c.shape;
  sh.slots = 1;
  sh.points = P.num_points;
  sh.eikonal = P.num_eikonal;
  sh.odom = P.num_odom;
  sh.cells = P.num_cells;
  sh.nodes = static_cast<int64_t>(nmap);
  sh.n = P.n;
  sh.state = static_cast<int64_t>(nmap + npose);
  sh.poses = P.num_poses;
  sh.step_items = P.num_active + P.num_poses;
  ++c.generation;
}

void ReadState(sdfslam_hip_context& c, LmState& st) {
  HIPC(hipMemcpyAsync(&st, c.dw.state, sizeof(LmState), hipMemcpyDeviceToHost, c.stream));
  HIPC(hipStreamSynchronize(c.stream));
}

void WriteState(sdfslam_hip_context& c, const LmState& st) {
  HIPC(hipMemcpyAsync(c.dw.state, &st, sizeof(LmState), hipMemcpyHostToDevice, c.stream));
  HIPC(hipStreamSynchronize(c.stream));
}

void DownloadState(sdfslam_hip_context& c, std::vector<double>& map, std::vector<double>& poses) {
  const sdfslam_problem_view& v = c.hp->view;
  map.resize(static_cast<size_t>(v.grid.nx) * v.grid.ny);
  poses.resize(3 * static_cast<size_t>(v.num_poses));
  HIPC(hipMemcpyAsync(map.data(), c.dw.map, map.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  HIPC(hipMemcpyAsync(poses.data(), c.dw.poses, poses.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  HIPC(hipStreamSynchronize(c.stream));
}

/// The context's current buffers as its launch slot.
void FillSlot(sdfslam_hip_context& c) {
  c.slot.P = c.dp;
  c.slot.A = c.da;
  c.slot.W = c.dw;
  c.slot.panel_total = c.an.panel_total;
  c.slot.nmap = c.hp->view.grid.nx * c.hp->view.grid.ny;
}

/// Refreshes the slot and its device copy (which lives in the epoch pool)
/// after a work buffer moved outside a pattern epoch change.
void SyncSlot(sdfslam_hip_context& c) {
  FillSlot(c);
  HIPC(hipMemcpyAsync(c.slot_dev, &c.slot, sizeof(Slot), hipMemcpyHostToDevice, c.stream));
  HIPC(hipStreamSynchronize(c.stream));
}

/// Launch plans of a launch over `ctxs` (slot b = ctxs[b]) for the merged
/// shape `sh`: per level and slot the level's fronts and its update-matrix
/// tiles. `all` receives the level and tile plans in that order (the layout
/// of their device copy, see SetPlans).
void BuildPlans(const std::vector<const sdfslam_hip_context*>& ctxs, LaunchShape& sh, std::vector<ListRef>& all) {
  const size_t k = ctxs.size();
  const size_t levels = sh.levels.size();
  std::vector<ListRef>& level_plan = sh.level_plan_h;
  std::vector<ListRef>& tile_plan = sh.tile_plan_h;
  level_plan.assign(levels * k, ListRef{nullptr, 0});
  tile_plan.assign(levels * k, ListRef{nullptr, 0});
  for (size_t b = 0; b < k; ++b) {
    const sdfslam_hip_context& c = *ctxs[b];
    const size_t own_levels = c.level_off_h.empty() ? 0 : c.level_off_h.size() - 1;
    for (size_t L = 0; L < std::min(levels, own_levels); ++L) {
      const int64_t l0 = c.level_off_h[L];
      level_plan[L * k + b] = {c.level_sn_d + l0, static_cast<int32_t>(c.level_off_h[L + 1] - l0)};
      const int64_t t0 = c.tile_off_h[L];
      tile_plan[L * k + b] = {c.tiles_d + 2 * t0, static_cast<int32_t>(c.tile_off_h[L + 1] - t0)};
    }
  }
  all = level_plan;
  all.insert(all.end(), tile_plan.begin(), tile_plan.end());
}

/// Points the shape's device plans into `dev`, a copy of BuildPlans' `all`.
void SetPlans(LaunchShape& sh, const ListRef* dev) {
  sh.level_plan = dev;
  sh.tile_plan = dev + sh.level_plan_h.size();
}

/// Widens the footprint dilation of the poses whose points left their
/// footprint in the ending epoch: by one cell (two when that epoch lasted at
/// most two iterations), and at least to twice the pose's motion bound over
/// the epoch (|dt| + reach * |dtheta|, in cells), since a pose that moved
/// fast tends to keep moving. Poses that moved more than half their margin
/// over the epoch grow the same way before their points leave (a warm
/// start's transient moves a few poses by whole cells per iteration, and
/// they would otherwise end the next epochs one after another). Up to
/// kMaxMargin. Poses that stay put keep the initial margin, so the pattern
/// grows only where the state moves.
constexpr int kMaxMargin = 12;
// Margin of the poses an incremental update appends; while a run has few
// frames (its map still barely constrained, its poses moving most, its
// pattern small) they start at kEarlyMargin.
constexpr int kFreshPoseMargin = 5;
constexpr int kEarlyMargin = 8;
constexpr int kEarlyFrames = 10;

void GrowMargins(sdfslam_hip_context& c, const std::vector<double>& poses) {
  const sdfslam_problem_view& v = c.hp->view;
  const int N = v.num_poses;
  if (static_cast<int>(c.aopt.pose_margin.size()) != N) {
    c.aopt.pose_margin.assign(static_cast<size_t>(N), c.aopt.margin);
  }
  const bool motion = static_cast<int>(c.anchor_poses.size()) == 3 * N &&
                      static_cast<int>(c.pose_reach.size()) == N && static_cast<int>(poses.size()) == 3 * N;
  const double h = std::min(v.grid.dx, v.grid.dy);
  std::vector<int32_t> flags(static_cast<size_t>(N));
  HIPC(hipMemcpyAsync(flags.data(), c.dw.pose_overflow, sizeof(int32_t) * flags.size(),
                      hipMemcpyDeviceToHost, c.stream));
  HIPC(hipStreamSynchronize(c.stream));
  const int step = c.short_epoch ? 2 : 1;
  for (int P = 0; P < N; ++P) {
    int32_t& m = c.aopt.pose_margin[static_cast<size_t>(P)];
    double cells = 0.0;  // motion bound over the epoch, in cells
    if (motion) {
      const size_t i = 3 * static_cast<size_t>(P);
      const double dth = std::remainder(poses[i + 2] - c.anchor_poses[i + 2], 2.0 * M_PI);
      cells = (std::hypot(poses[i] - c.anchor_poses[i], poses[i + 1] - c.anchor_poses[i + 1]) +
               c.pose_reach[static_cast<size_t>(P)] * std::fabs(dth)) / h;
    }
    if (flags[static_cast<size_t>(P)] != 0) {
      m = std::min(std::max(m + step, static_cast<int>(std::ceil(2.0 * cells))), kMaxMargin);
    } else if (cells > 0.5 * m) {
      m = std::min(std::max(m + 1, static_cast<int>(std::ceil(2.0 * cells))), kMaxMargin);
    }
  }
  HIPC(hipMemsetAsync(c.dw.pose_overflow, 0, sizeof(int32_t) * flags.size(), c.stream));
}

/// Symbolic analysis and the device buffers that depend on it (a new
/// pattern epoch): at the given host state of a new problem (its poses keep
/// the margins they have), or at the current device state (a point left its
/// footprint: the flagged poses' margins grow).
void InstallEpoch(sdfslam_hip_context& c, Analysis&& analysis, const double* anchor_poses,
                  SteadyTime t0);

void BuildEpoch(sdfslam_hip_context& c, const double* host_map = nullptr, const double* host_poses = nullptr) {
  std::vector<double> map;
  std::vector<double> poses;
  const auto t0 = std::chrono::steady_clock::now();
  if (host_map != nullptr && host_poses != nullptr) {
    const int N = c.hp->view.num_poses;
    if (static_cast<int>(c.aopt.pose_margin.size()) != N) {
      c.aopt.pose_margin.assign(static_cast<size_t>(N), c.aopt.margin);
    }
    InstallEpoch(c, Analyze(c.hp->view, host_map, host_poses, c.aopt), host_poses, t0);
  } else {
    DownloadState(c, map, poses);
    GrowMargins(c, poses);
    InstallEpoch(c, Analyze(c.hp->view, map.data(), poses.data(), c.aopt), poses.data(), t0);
  }
}

/// Makes `analysis` (of the current problem, with the context's margins,
/// at a state whose poses were `anchor_poses`) the context's pattern epoch:
/// device copies, launch shape and plans, slot.
void InstallEpoch(sdfslam_hip_context& c, Analysis&& analysis, const double* anchor_poses,
                  SteadyTime t0) {
  c.an = std::move(analysis);
  c.anchor_poses.assign(anchor_poses, anchor_poses + 3 * static_cast<size_t>(c.hp->view.num_poses));
  ++c.pattern_epochs;
  const Analysis& an = c.an;
  if (an.max_points_per_pose > 4096) {
    throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "more than 4096 point rows in one frame");
  }
  c.max_f = 0;
  for (const int f : an.sn_f) c.max_f = std::max(c.max_f, f);
  if (static_cast<size_t>(c.max_f) * sizeof(double) > 150 * 1024) {
    throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "front larger than the solve kernels' LDS budget");
  }
  std::vector<int64_t> cptr(static_cast<size_t>(an.num_sn) + 1, 0);
  std::vector<int32_t> clist;
  // Diagonal-block inverses: 64 x 64 per block column of every supernode.
  std::vector<int64_t> inv_off(static_cast<size_t>(an.num_sn) + 1, 0);
  for (int s = 0; s < an.num_sn; ++s) {
    for (const int ch : an.children[static_cast<size_t>(s)]) clist.push_back(ch);
    cptr[static_cast<size_t>(s) + 1] = static_cast<int64_t>(clist.size());
    const int ns = an.sn_first[static_cast<size_t>(s) + 1] - an.sn_first[static_cast<size_t>(s)];
    inv_off[static_cast<size_t>(s) + 1] = inv_off[static_cast<size_t>(s)] + static_cast<int64_t>((ns + 63) / 64) * 64 * 64;
  }
  // Per child whose entries reach its parent's update matrix: the map from
  // the parent's update rows to the child's (rel minus the parent's own
  // columns, inverted; -1 where the child has no row), for the gather in the
  // parent's update-matrix kernel.
  std::vector<int64_t> uinv_off(static_cast<size_t>(an.num_sn), -1);
  std::vector<int32_t> uinv;
  for (int ch = 0; ch < an.num_sn; ++ch) {
    const int par = an.sn_parent[static_cast<size_t>(ch)];
    const int mc = an.sn_m[static_cast<size_t>(ch)];
    const int split = an.rel_split[static_cast<size_t>(ch)];
    if (par < 0 || split >= mc) continue;
    const int np = an.sn_first[static_cast<size_t>(par) + 1] - an.sn_first[static_cast<size_t>(par)];
    const int mp = an.sn_m[static_cast<size_t>(par)];
    uinv_off[static_cast<size_t>(ch)] = static_cast<int64_t>(uinv.size());
    uinv.resize(uinv.size() + static_cast<size_t>(mp), -1);
    const int32_t* rel = an.rel.data() + an.rel_off[static_cast<size_t>(ch)];
    for (int a = split; a < mc; ++a) {
      uinv[static_cast<size_t>(uinv_off[static_cast<size_t>(ch)] + rel[a] - np)] = a;
    }
  }
  // Per child whose entries reach its parent's own columns: the map from the
  // parent's own columns to the child's columns (rel below the split,
  // inverted; -1 where no child column lands), for the panel extend-add.
  std::vector<int64_t> pinv_off(static_cast<size_t>(an.num_sn), -1);
  std::vector<int32_t> pinv;
  for (int ch = 0; ch < an.num_sn; ++ch) {
    const int par = an.sn_parent[static_cast<size_t>(ch)];
    const int split = an.rel_split[static_cast<size_t>(ch)];
    if (par < 0 || split <= 0) continue;
    const int np = an.sn_first[static_cast<size_t>(par) + 1] - an.sn_first[static_cast<size_t>(par)];
    pinv_off[static_cast<size_t>(ch)] = static_cast<int64_t>(pinv.size());
    pinv.resize(pinv.size() + static_cast<size_t>(np), -1);
    const int32_t* rel = an.rel.data() + an.rel_off[static_cast<size_t>(ch)];
    for (int a = 0; a < split; ++a) pinv[static_cast<size_t>(pinv_off[static_cast<size_t>(ch)] + rel[a])] = a;
  }
  // Level lists and update-matrix tiles, with the launch shape of this
  // context.
  std::vector<int32_t> lvl;
  std::vector<int64_t> level_off(static_cast<size_t>(an.num_levels) + 1, 0);
  std::vector<int32_t> tiles;
  std::vector<int64_t> tile_off(static_cast<size_t>(an.num_levels) + 1, 0);
  LaunchShape& sh = c.shape;
  sh.levels.assign(static_cast<size_t>(an.num_levels), {});
  for (int L = 0; L < an.num_levels; ++L) {
    LaunchShape::Level& lv = sh.levels[static_cast<size_t>(L)];
    for (const int s : an.level_sn[static_cast<size_t>(L)]) lvl.push_back(s);
    level_off[static_cast<size_t>(L) + 1] = static_cast<int64_t>(lvl.size());
    lv.count = static_cast<int>(an.level_sn[static_cast<size_t>(L)].size());
    lv.max_nb = an.level_max_nb[static_cast<size_t>(L)];
    lv.max_f = an.level_max_f[static_cast<size_t>(L)];
    lv.max_m = an.level_max_m[static_cast<size_t>(L)];
    lv.has_children = an.level_max_children[static_cast<size_t>(L)] > 0;
    for (const int s : an.level_sn[static_cast<size_t>(L)]) {
      const int t = (an.sn_m[static_cast<size_t>(s)] + 63) / 64;
      if (t == 0) lv.all_update = false;
      for (int ti = 0; ti < t; ++ti) {
        for (int tj = 0; tj <= ti; ++tj) {
          tiles.push_back(s);
          tiles.push_back((ti << 16) | tj);
        }
      }
    }
    tile_off[static_cast<size_t>(L) + 1] = static_cast<int64_t>(tiles.size() / 2);
    lv.tiles = static_cast<int>(tile_off[static_cast<size_t>(L) + 1] - tile_off[static_cast<size_t>(L)]);
  }
  sh.max_points_per_pose = std::max(1, an.max_points_per_pose);
  sh.panel_total = an.panel_total;
  sh.pairs = static_cast<int64_t>(an.op_pose.size() / 2);
  c.level_off_h = std::move(level_off);
  c.tile_off_h = std::move(tile_off);
  const size_t plan_entries = 2 * sh.levels.size();
  std::vector<ListRef> plans;
  EpochPool& E = c.epoch;
  DeviceAnalysis& D = c.da;
  auto layout = [&]() {
    D.perm = E.Upload(an.perm);
    D.iperm = E.Upload(an.iperm);
    D.sn_first = E.Upload(an.sn_first);
    D.sn_m = E.Upload(an.sn_m);
    D.sn_f = E.Upload(an.sn_f);
    D.sn_idx_ptr = E.Upload(an.sn_idx_ptr);
    D.sn_idx = E.Upload(an.sn_idx);
    D.panel_off = E.Upload(an.panel_off);
    D.upd_off = E.Upload(an.upd_off);
    D.sn_parent = E.Upload(an.sn_parent);
    D.rel_off = E.Upload(an.rel_off);
    D.rel = E.Upload(an.rel);
    D.rel_split = E.Upload(an.rel_split);
    D.uinv_off = E.Upload(uinv_off);
    D.uinv = E.Upload(uinv);
    D.pinv_off = E.Upload(pinv_off);
    D.pinv = E.Upload(pinv);
    D.nn_off = E.Upload(an.nn_off);
    D.eik_at = E.Upload(an.eik_at);
    D.fp_ptr = E.Upload(an.fp_ptr);
    D.fp_node = E.Upload(an.fp_node);
    D.fp_off = E.Upload(an.fp_off);
    D.fp_stride = E.Upload(an.fp_stride);
    D.words_per_pose = an.words_per_pose;
    D.fp_cellbits = E.Upload(an.fp_cellbits);
    D.pr_ptr = E.Upload(an.pr_ptr);
    D.pr_row = E.Upload(an.pr_row);
    D.cp_ptr = E.Upload(an.cp_ptr);
    D.cp_pose = E.Upload(an.cp_pose);
    D.cp_rank = E.Upload(an.cp_rank);
    D.dl_ptr = E.Upload(an.dl_ptr);
    D.fp_cellrank = E.Upload(an.fp_cellrank);
    D.pd_off = E.Upload(an.pd_off);
    D.po_ptr = E.Upload(an.po_ptr);
    D.po_spec = E.Upload(an.po_spec);
    D.num_pairs = static_cast<int>(an.op_pose.size() / 2);
    D.op_pose = E.Upload(an.op_pose);
    D.op_ptr = E.Upload(an.op_ptr);
    D.op_spec = E.Upload(an.op_spec);
    D.op_off = E.Upload(an.op_off);
    D.children_ptr = E.Upload(cptr);
    D.children = E.Upload(clist);
    D.inv_off = E.Upload(inv_off);
    c.level_sn_d = E.Upload(lvl);
    c.tiles_d = E.Upload(tiles);
    c.dw.panel = E.Alloc<double>(static_cast<size_t>(an.panel_total));
    c.dw.upd = E.Alloc<double>(static_cast<size_t>(an.upd_total));
    c.dw.diag_stage = E.Alloc<double>(static_cast<size_t>(an.num_sn) * 2 * 64 * 64);
    c.dw.diag_inv = E.Alloc<double>(static_cast<size_t>(inv_off.back()));
    c.dw.diag_d = E.Alloc<double>(static_cast<size_t>(inv_off.back() / 64));
    // Multi-workgroup solve words: 3 per supernode and one per 64-column
    // block, uploaded as zeros.
    c.solve_sync_words = static_cast<size_t>(inv_off.back() / (64 * 64) + 3 * an.num_sn);
    c.dw.solve_sync = E.Upload(std::vector<int32_t>(c.solve_sync_words, 0));
    c.dw.vec = E.Alloc<double>(static_cast<size_t>(an.sn_idx_ptr.back()));
    c.dw.seg_map = E.Alloc<int32_t>(an.cp_pose.size());
    // The launch plans and the slot come last: they hold addresses placed
    // above (known in the placing pass).
    if (E.measuring()) {
      E.Upload(static_cast<const ListRef*>(nullptr), plan_entries);
      E.Upload(static_cast<const Slot*>(nullptr), 1);
    } else {
      BuildPlans({&c}, c.shape, plans);
      SetPlans(c.shape, E.Upload(plans));
      FillSlot(c);
      c.slot_dev = E.Upload(&c.slot, 1);
    }
  };
  E.BeginMeasure();
  layout();
  E.BeginPlace();
  layout();
  Enqueue(c, [&E](hipStream_t s) { E.Copy(s); });
  ++c.generation;
  c.analysis_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

enum Phase { kAssembly = 0, kNormal = 1, kFactor = 2, kSolve = 3, kUpdate = 4 };

void Mark(sdfslam_hip_context& c, bool timing, int k) {
  if (timing) HIPC(hipEventRecord(c.events[k], c.stream));
}

void EnqueueAssemblyAndNormal(sdfslam_hip_context& c, const LmOptions& o, bool timing) {
  hipStream_t s = c.stream;
  LaunchAssemble(SlotArgs{c.slot_dev, &c.slot}, c.shape, true, s, timing ? c.events[kPhases + 1] : nullptr,
                 timing ? c.events[kPhases + 2] : nullptr);
  Mark(c, timing, 1);
  LaunchNormalSlots(SlotArgs{c.slot_dev, &c.slot}, c.shape, o, s);
}

// Factorizations of at least this many flops run their lower levels'
// forward solves on the calling thread's second stream (a long chain of
// dependent top-level steps leaves the device room for them). Smaller ones,
// and batched solves, keep one stream: several pipelines with two streams
// each share the device's few hardware queues, which then serializes
// unrelated work (the served form at 10 Hz went from p50 20 ms to 110 ms).
constexpr double kOverlapMinFlops = 1e9;

void EnqueueFactorAndSolve(sdfslam_hip_context& c, bool timing) {
  hipStream_t s = c.stream;
  const Analysis& an = c.an;
  StepEvents step_ev{nullptr, 0, 0, nullptr, 0, 0};
  if (timing) {
    const size_t want = 2 * static_cast<size_t>(std::min(an.panel_step_launches, 1024));
    while (c.step_events.size() < want) {
      hipEvent_t e = nullptr;
      HIPC(hipEventCreate(&e));
      c.step_events.push_back(e);
    }
    const size_t want_gemm = 2 * static_cast<size_t>(std::min(an.gemm_launches, 1024));
    while (c.gemm_events.size() < want_gemm) {
      hipEvent_t e = nullptr;
      HIPC(hipEventCreate(&e));
      c.gemm_events.push_back(e);
    }
    step_ev = {c.step_events.data(), static_cast<int>(c.step_events.size() / 2), 0,
               c.gemm_events.data(), static_cast<int>(c.gemm_events.size() / 2), 0};
  }
  if (!timing && an.factor_flops >= kOverlapMinFlops) {
    // The lower levels' forward solves beside the upper levels' factorization.
    if (c.aux_stream == nullptr) c.aux_stream = t_aux_streams.Get(c.device);
    HIPC(LaunchFactorSolveOverlapped(SlotArgs{c.slot_dev, &c.slot}, c.shape, s, c.aux_stream,
                                     OverlapEvents(c.overlap_events, c.shape.levels.size() + 1)));
    return;
  }
  // Timed phases: the factorization and the solves one after the other.
  LaunchFactor(SlotArgs{c.slot_dev, &c.slot}, c.shape, s, &step_ev);
  c.step_events_used = step_ev.used;
  c.gemm_events_used = step_ev.gemm_used;
  Mark(c, timing, 3);
  LaunchSolve(SlotArgs{c.slot_dev, &c.slot}, c.shape, s);
}

void EnqueueIteration(sdfslam_hip_context& c, const LmOptions& o, bool timing) {
  Mark(c, timing, 0);
  EnqueueAssemblyAndNormal(c, o, timing);
  Mark(c, timing, 2);
  EnqueueFactorAndSolve(c, timing);
  Mark(c, timing, 4);
  LaunchStepAndControl(SlotArgs{c.slot_dev, &c.slot}, c.shape, o, c.stream);
  Mark(c, timing, 5);
}

bool SameOptions(const LmOptions& a, const LmOptions& b) {
  return a.max_iterations == b.max_iterations && a.history_capacity == b.history_capacity &&
         a.step_tolerance == b.step_tolerance && a.lambda_factor == b.lambda_factor &&
         a.reject_worse_steps == b.reject_worse_steps && a.marquardt_scaling == b.marquardt_scaling &&
         a.trust_region == b.trust_region;
}

/// Runs shorter than this many iterations launch kernels directly unless a
/// captured graph for the current buffers already exists (a capture costs
/// about one direct enqueue of the iteration plus its instantiation).
constexpr int kGraphMinIterations = 2;

bool GraphValid(const sdfslam_hip_context& c, const LmOptions& o) {
  return c.graph != nullptr && c.graph_generation == c.generation && SameOptions(c.graph_options, o);
}

/// One LM iteration captured as a HIP graph (re-captured when buffers or
/// options change), so a batch of iterations is a batch of graph launches.
hipGraphExec_t IterationGraph(sdfslam_hip_context& c, const LmOptions& o) {
  if (GraphValid(c, o)) return c.graph;
  if (c.graph != nullptr) {
    HIPC(hipGraphExecDestroy(c.graph));
    c.graph = nullptr;
  }
  hipGraph_t g = nullptr;
  // Relaxed: other host threads driving their own contexts may call HIP
  // (allocations, copies) while this stream is being captured.
  HIPC(hipStreamBeginCapture(c.stream, hipStreamCaptureModeRelaxed));
  EnqueueIteration(c, o, false);
  HIPC(hipStreamEndCapture(c.stream, &g));
  const hipError_t e = hipGraphInstantiate(&c.graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  HIPC(e);
  c.graph_generation = c.generation;
  c.graph_options = o;
  return c.graph;
}

void AccumulateTimes(sdfslam_hip_context& c) {
  for (int k = 0; k < c.step_events_used; ++k) {
    float ms = 0.0f;
    HIPC(hipEventElapsedTime(&ms, c.step_events[2 * static_cast<size_t>(k)],
                             c.step_events[2 * static_cast<size_t>(k) + 1]));
    c.panel_step_ms += ms;
    ++c.panel_steps_timed;
  }
  c.step_events_used = 0;
  for (int k = 0; k < c.gemm_events_used; ++k) {
    float ms = 0.0f;
    HIPC(hipEventElapsedTime(&ms, c.gemm_events[2 * static_cast<size_t>(k)],
                             c.gemm_events[2 * static_cast<size_t>(k) + 1]));
    c.gemm_ms += ms;
    ++c.gemms_timed;
  }
  c.gemm_events_used = 0;
  for (int k = 0; k < kPhases; ++k) {
    float ms = 0.0f;
    HIPC(hipEventElapsedTime(&ms, c.events[k], c.events[k + 1]));
    c.phase_ms[k] += ms;
  }
  float ms = 0.0f;
  HIPC(hipEventElapsedTime(&ms, c.events[kPhases + 1], c.events[kPhases + 2]));
  c.phase_ms[kPhases] += ms;
  ++c.timed_iterations;
}

/// Clears the multi-workgroup solve words (normally clear between launches;
/// an aborted launch can leave them set).
void ClearSolveSync(sdfslam_hip_context& c, hipStream_t s) {
  if (c.dw.solve_sync != nullptr && c.solve_sync_words > 0) {
    HIPC(hipMemsetAsync(c.dw.solve_sync, 0, sizeof(int32_t) * c.solve_sync_words, s));
  }
}

/// The status of a stopped LM state: a zero pivot, or a multi-workgroup
/// solve fault.
int StateStatus(const LmState& st) {
  if (st.not_spd == kSolveSyncFault) {
    g_last_error = "triangular solve workgroups gave up waiting for each other";
    return SDFSLAM_ERR_INTERNAL;
  }
  g_last_error = "sparse Cholesky factorization failed";
  return SDFSLAM_ERR_NOT_SPD;
}

void InitLm(sdfslam_hip_context& c, double lambda_init) {
  ClearSolveSync(c, c.stream);
  LaunchCost(SlotArgs{c.slot_dev, &c.slot}, c.shape, c.stream);
  LaunchLmInit(SlotArgs{c.slot_dev, &c.slot}, c.shape, lambda_init, c.stream);
  c.lm_ready = true;
}

/// Clears the flags so stage calls run, optionally with a given lambda.
void PrepareStage(sdfslam_hip_context& c, const double* lambda) {
  LmState st{};
  ReadState(c, st);
  st.halt = 0;
  st.overflow = 0;
  st.not_spd = 0;
  st.revert = 0;
  if (lambda != nullptr) st.lambda = *lambda;
  WriteState(c, st);
  ClearSolveSync(c, c.stream);
}

/// Waits for a running speculative re-analysis and discards it.
void DropSpeculation(sdfslam_hip_context& c) {
  if (c.spec == nullptr) return;
  c.spec->job.wait();
  c.spec.reset();
}

/// Between chunks of iterations (the device idle): installs the
/// speculative re-analysis started at the previous chunk's end, or starts
/// one when a pose has moved so far from where its footprint was computed
/// that its points may soon leave it.
/// A point moves by at most |dt| + reach * |dtheta| when its pose moves by
/// (dt, dtheta), and its cell by at most that over the cell size: while
/// this stays below the pose's margin (in cells) no point can leave the
/// footprint. At one cell short of the margin the pose's margin grows by
/// one and the analysis of the current state runs on a host thread while
/// the device runs the next chunk in the current epoch; the new epoch takes
/// over at that chunk's end (waiting for the analysis if need be, so that
/// where epochs change depends on the device state alone and runs repeat
/// bit for bit). A point that leaves a footprint still ends the epoch at
/// once.
void Speculate(sdfslam_hip_context& c) {
  if (c.spec != nullptr) {
    const auto t0 = Now();
    std::unique_ptr<sdfslam_hip_context::Speculation> sp = std::move(c.spec);
    c.aopt = sp->aopt;
    InstallEpoch(c, sp->job.get(), sp->poses.data(), t0);
    ++c.speculative_epochs;
    return;
  }
  const sdfslam_problem_view& v = c.hp->view;
  const int N = v.num_poses;
  if (static_cast<int>(c.anchor_poses.size()) != 3 * N || static_cast<int>(c.pose_reach.size()) != N) return;
  std::vector<double> poses(3 * static_cast<size_t>(N));
  HIPC(hipMemcpyAsync(poses.data(), c.dw.poses, poses.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  HIPC(hipStreamSynchronize(c.stream));
  const double h = std::min(v.grid.dx, v.grid.dy);
  AnalysisOptions aopt = c.aopt;
  if (static_cast<int>(aopt.pose_margin.size()) != N) aopt.pose_margin.assign(static_cast<size_t>(N), aopt.margin);
  bool grow = false;
  for (int P = 0; P < N; ++P) {
    const size_t i = 3 * static_cast<size_t>(P);
    const double dth = std::remainder(poses[i + 2] - c.anchor_poses[i + 2], 2.0 * M_PI);
    const double d = std::hypot(poses[i] - c.anchor_poses[i], poses[i + 1] - c.anchor_poses[i + 1]) +
                     c.pose_reach[static_cast<size_t>(P)] * std::fabs(dth);
    int32_t& m = aopt.pose_margin[static_cast<size_t>(P)];
    if (d > (m - 1) * h && m < kMaxMargin) {
      ++m;
      grow = true;
    }
  }
  if (!grow) return;
  auto sp = std::make_unique<sdfslam_hip_context::Speculation>();
  sp->map.resize(static_cast<size_t>(v.grid.nx) * v.grid.ny);
  HIPC(hipMemcpyAsync(sp->map.data(), c.dw.map, sp->map.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  HIPC(hipStreamSynchronize(c.stream));
  sp->poses = std::move(poses);
  sp->aopt = std::move(aopt);
  const HostProblem* hp = c.hp.get();
  sdfslam_hip_context::Speculation* raw = sp.get();
  sp->job = std::async(std::launch::async,
                       [hp, raw]() { return Analyze(hp->view, raw->map.data(), raw->poses.data(), raw->aopt); });
  c.spec = std::move(sp);
}

/// Runs `iterations` LM iterations, re-analysing when a point leaves its
/// pose's footprint (and speculatively before that, see Speculate).
/// Returns the final LM state.
LmState RunIterations(sdfslam_hip_context& c, const LmOptions& o, int iterations, bool timing) {
  LmState st{};
  ReadState(c, st);
  const int target = st.iter + iterations;
  int last_overflow_iter = -1;
  int same_iter_overflow = 0;
  while (true) {
    if (st.halt != 0 || st.iter >= target) break;
    const int chunk = timing ? (target - st.iter) : std::min(target - st.iter, 8);
    if (timing) {
      for (int k = 0; k < chunk; ++k) {
        EnqueueIteration(c, o, true);
        HIPC(hipStreamSynchronize(c.stream));
        AccumulateTimes(c);
      }
    } else if (!g_graphs.load() || (target - st.iter < kGraphMinIterations && !GraphValid(c, o))) {
      // Too few iterations left to repay a capture: launch directly.
      for (int k = 0; k < chunk; ++k) EnqueueIteration(c, o, false);
    } else {
      hipGraphExec_t g = IterationGraph(c, o);
      for (int k = 0; k < chunk; ++k) HIPC(hipGraphLaunch(g, c.stream));
    }
    ReadState(c, st);
    if (st.not_spd != 0) break;
    if (st.overflow != 0) {
      if (st.iter == last_overflow_iter && same_iter_overflow++ > 0) {
        throw Failure(SDFSLAM_ERR_INTERNAL, "pattern re-analysis did not cover the point cells");
      }
      // Re-centre the footprints on the current state and widen those of
      // the poses that left theirs.
      DropSpeculation(c);
      c.short_epoch = last_overflow_iter >= 0 && st.iter - last_overflow_iter <= 2;
      last_overflow_iter = st.iter;
      BuildEpoch(c);
      c.short_epoch = false;
      st.overflow = 0;
      st.halt = 0;
      WriteState(c, st);
    } else if (!timing && st.halt == 0 && st.iter < target) {
      Speculate(c);
    }
  }
  // A re-analysis started for iterations this call does not run is dropped.
  DropSpeculation(c);
  return st;
}

/// LM loops of several contexts with every stage of an iteration launched
/// once over all of them (their slots in one table, one captured graph on
/// the lead context's stream, the lead being the first context). Each
/// context keeps its own LM state, pattern epochs and results: the same as
/// solving it alone. The table, plans and state buffers live in the lead
/// context and are reused by its later batches.
/// The calling thread's stream becomes every context's stream; their
/// deferred work runs on it and on kPendingStreams - 1 more streams side by
/// side (each context's in order on one stream; forked from and joined into
/// the thread's stream with events). Their LM initialisation is dropped:
/// the batched solve initialises every slot.
void RunBatchPending(sdfslam_hip_context* const* ctxs, int count) {
  const int dev = ctxs[0]->device;
  HIPC(hipSetDevice(dev));
  const hipStream_t main = t_streams.Get(dev);
  std::vector<sdfslam_hip_context*> work;
  for (int b = 0; b < count; ++b) {
    sdfslam_hip_context& c = *ctxs[b];
    c.stream = main;
    c.aux_stream = nullptr;
    c.init_pending = false;
    DropSpeculation(c);
    if (!c.pending.empty()) work.push_back(&c);
  }
  if (work.size() <= 1) {
    for (sdfslam_hip_context* c : work) RunPending(*c, main);
    return;
  }
  std::vector<hipEvent_t>& ev = t_pending_events;
  OverlapEvents(ev, kPendingStreams);
  HIPC(hipEventRecord(ev[0], main));
  hipStream_t streams[kPendingStreams];
  streams[0] = main;
  for (int k = 1; k < kPendingStreams; ++k) {
    streams[k] = t_pending_streams[k - 1].Get(dev);
    HIPC(hipStreamWaitEvent(streams[k], ev[0], 0));
  }
  for (size_t i = 0; i < work.size(); ++i) RunPending(*work[i], streams[i % kPendingStreams]);
  for (int k = 1; k < kPendingStreams; ++k) {
    HIPC(hipEventRecord(ev[static_cast<size_t>(k)], streams[k]));
    HIPC(hipStreamWaitEvent(main, ev[static_cast<size_t>(k)], 0));
  }
}

class BatchRun {
 public:
  BatchRun(sdfslam_hip_context* const* ctxs, int count)
      : ctxs_(ctxs, ctxs + count), space_(ctxs[0]->batch) {}
  ~BatchRun() {
    if (graph_ != nullptr) (void)hipGraphExecDestroy(graph_);
  }
  BatchRun(const BatchRun&) = delete;
  BatchRun& operator=(const BatchRun&) = delete;

  const sdfslam_hip_batch_info& info() const { return info_; }

  /// Runs every context's LM loop to its end; status per context.
  void Run(const sdfslam_solver_options& options, std::vector<int>& status,
           std::vector<LmState>& states) {
    const int k = static_cast<int>(ctxs_.size());
    stream_ = ctxs_[0]->stream;
    const auto t_setup = Now();
    int capacity = INT_MAX;
    for (sdfslam_hip_context* c : ctxs_) {
      HIPC(hipStreamSynchronize(c->stream));
      if (options.max_iterations + 1 > c->history_capacity) {
        c->history_capacity = options.max_iterations + 1;
        c->dw.cost_history = c->hist.Alloc<double>(static_cast<size_t>(c->history_capacity));
        SyncSlot(*c);
        ++c->generation;
      }
      capacity = std::min(capacity, c->history_capacity);
      c->lm_ready = true;
    }
    options_ = ToLm(options, capacity);
    space_.Reserve(k);
    info_.setup_seconds += SecondsSince(t_setup);
    status.assign(static_cast<size_t>(k), SDFSLAM_OK);
    states.assign(static_cast<size_t>(k), LmState{});
    std::vector<int> last_overflow(static_cast<size_t>(k), -1);
    std::vector<int> repeats(static_cast<size_t>(k), 0);
    Rebuild();
    // LM initialisation of every slot: two launches over the table.
    for (sdfslam_hip_context* c : ctxs_) ClearSolveSync(*c, stream_);
    const SlotArgs all{space_.table, nullptr};
    LaunchCost(all, shape_, stream_);
    LaunchLmInit(all, shape_, options.lambda_init, stream_);
    ReadAll(states);
    while (true) {
      int remaining = 0;
      for (const LmState& st : states) {
        if (st.halt == 0) remaining = std::max(remaining, options.max_iterations - st.iter);
      }
      if (remaining <= 0) break;
      const int chunk = std::min(remaining, 8);
      for (int i = 0; i < chunk; ++i) {
        if (graph_ != nullptr) {
          HIPC(hipGraphLaunch(graph_, stream_));
        } else {
          EnqueueIteration();
        }
      }
      info_.graph_launches += chunk;
      ReadAll(states);
      bool changed = false;
      for (int b = 0; b < k; ++b) {
        LmState& st = states[static_cast<size_t>(b)];
        sdfslam_hip_context& c = *ctxs_[static_cast<size_t>(b)];
        if (st.not_spd != 0) {
          status[static_cast<size_t>(b)] = StateStatus(st);
          continue;
        }
        if (st.overflow == 0) {
          // The same speculative epochs as a solve of this context alone
          // (RunIterations): its chunks end at the same iterations, so the
          // pattern, and with it every rounding, follows the same sequence.
          if (st.halt == 0 && st.iter < options.max_iterations) {
            const int epochs = c.pattern_epochs;
            const auto t_spec = Now();
            Speculate(c);
            if (c.pattern_epochs != epochs) {
              info_.analysis_seconds += SecondsSince(t_spec);
              ++info_.epochs_rebuilt;
              changed = true;
            }
          }
          continue;
        }
        if (st.iter == last_overflow[static_cast<size_t>(b)] && repeats[static_cast<size_t>(b)]++ > 0) {
          throw Failure(SDFSLAM_ERR_INTERNAL, "pattern re-analysis did not cover the point cells");
        }
        c.short_epoch = last_overflow[static_cast<size_t>(b)] >= 0 &&
                        st.iter - last_overflow[static_cast<size_t>(b)] <= 2;
        last_overflow[static_cast<size_t>(b)] = st.iter;
        DropSpeculation(c);
        const auto t_epoch = Now();
        BuildEpoch(c);
        info_.analysis_seconds += SecondsSince(t_epoch);
        ++info_.epochs_rebuilt;
        c.short_epoch = false;
        st.overflow = 0;
        st.halt = 0;
        WriteState(c, st);
        changed = true;
      }
      if (changed) Rebuild();
    }
    // A re-analysis started for iterations this call does not run is dropped.
    for (sdfslam_hip_context* c : ctxs_) DropSpeculation(*c);
  }

 private:
  /// One LM iteration of every slot, enqueued on the batch's stream.
  void EnqueueIteration() {
    const int k = static_cast<int>(ctxs_.size());
    const SlotArgs args{space_.table, k == 1 ? &ctxs_[0]->slot : nullptr};
    LaunchAssemble(args, shape_, true, stream_);
    LaunchNormalSlots(args, shape_, options_, stream_);
    LaunchFactor(args, shape_, stream_);
    LaunchSolve(args, shape_, stream_);
    LaunchStepAndControl(args, shape_, options_, stream_);
  }

  /// Slot table, merged shape and the captured iteration for the current
  /// buffers of every context.
  void Rebuild() {
    const auto t_capture = Now();
    const int k = static_cast<int>(ctxs_.size());
    shape_ = LaunchShape{};
    for (int b = 0; b < k; ++b) {
      space_.slots_host[b] = ctxs_[static_cast<size_t>(b)]->slot;
      shape_.Merge(ctxs_[static_cast<size_t>(b)]->shape);
    }
    std::vector<const sdfslam_hip_context*> cs(ctxs_.begin(), ctxs_.end());
    std::vector<ListRef> plans;
    BuildPlans(cs, shape_, plans);
    space_.ReservePlans(plans.size());
    std::copy(plans.begin(), plans.end(), space_.plan_host);
    // Pinned staging: the two uploads are true async copies; the stream is
    // synchronised before the staging is touched again.
    HIPC(hipMemcpyAsync(space_.table, space_.slots_host, sizeof(Slot) * static_cast<size_t>(k),
                        hipMemcpyHostToDevice, stream_));
    HIPC(hipMemcpyAsync(space_.plan, space_.plan_host, sizeof(ListRef) * plans.size(),
                        hipMemcpyHostToDevice, stream_));
    SetPlans(shape_, space_.plan);
    HIPC(hipStreamSynchronize(stream_));
    if (graph_ != nullptr) {
      HIPC(hipGraphExecDestroy(graph_));
      graph_ = nullptr;
    }
    if (!g_graphs.load()) {
      info_.capture_seconds += SecondsSince(t_capture);
      return;
    }
    hipGraph_t g = nullptr;
    HIPC(hipStreamBeginCapture(stream_, hipStreamCaptureModeRelaxed));
    EnqueueIteration();
    HIPC(hipStreamEndCapture(stream_, &g));
    const hipError_t e = hipGraphInstantiate(&graph_, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPC(e);
    ++info_.graph_captures;
    info_.capture_seconds += SecondsSince(t_capture);
  }

  /// Every slot's LM state with one gather launch and one copy.
  void ReadAll(std::vector<LmState>& states) {
    LaunchGatherStates(SlotArgs{space_.table, nullptr}, shape_, space_.states_dev, stream_);
    HIPC(hipMemcpyAsync(space_.states_host, space_.states_dev, sizeof(LmState) * states.size(),
                        hipMemcpyDeviceToHost, stream_));
    HIPC(hipStreamSynchronize(stream_));
    std::copy(space_.states_host, space_.states_host + states.size(), states.begin());
  }

  std::vector<sdfslam_hip_context*> ctxs_;
  BatchSpace& space_;
  LmOptions options_{};
  hipStream_t stream_{nullptr};
  LaunchShape shape_;
  hipGraphExec_t graph_{nullptr};
  sdfslam_hip_batch_info info_{};
};

int Guard(const std::function<int()>& fn) {
  try {
    return fn();
  } catch (const Failure& f) {
    g_last_error = f.what();
    return f.code;
  } catch (const std::invalid_argument& e) {
    g_last_error = e.what();
    return SDFSLAM_ERR_INVALID_ARGUMENT;
  } catch (const std::bad_alloc& e) {
    g_last_error = e.what();
    return SDFSLAM_ERR_OOM;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return SDFSLAM_ERR_INTERNAL;
  }
}

void CreateContext(const sdfslam_problem_view* problem, const sdfslam_solver_options* options,
                   const double* map_values, const double* poses, sdfslam_hip_context** out) {
  if (problem == nullptr || options == nullptr || map_values == nullptr || poses == nullptr ||
      out == nullptr) {
    throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "null argument");
  }
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
    throw Failure(SDFSLAM_ERR_NO_DEVICE, "no HIP device");
  }
  auto c = std::make_unique<sdfslam_hip_context>();
  c->hp = std::make_unique<HostProblem>(*problem);
  c->opts = *options;
  HIPC(hipGetDevice(&c->device));
  Bind(*c);
  for (hipEvent_t& e : c->events) HIPC(hipEventCreate(&e));
  ComputeScanList(*c);
  ComputeReach(*c, 0);
  PlaceRows(*c, static_cast<int64_t>(c->hp->pt_frame.size()), static_cast<int64_t>(c->scan_list.size()));
  AllocateBase(*c, map_values, poses);
  BuildEpoch(*c, map_values, poses);
  InitLm(*c, options->lambda_init);
  HIPC(hipStreamSynchronize(c->stream));
  *out = c.release();
}

/// Replaces the context's problem and state, keeping its device memory
/// (pools grow only when the new problem needs more). Host work only: the
/// uploads and the LM initialisation run at the start of the context's next
/// call, on that call's stream.
/// Enqueues (deferred, like the rest of an update) the device computation
/// of the eikonal specs' normals of the context's problem from its scan
/// points (the rows with expected SDF 0) at its poses.
void StageDeviceNormals(sdfslam_hip_context& c, const sdfslam_normals_request& nr, const double* poses) {
  const HostProblem& hp = *c.hp;
  const sdfslam_problem_view& v = hp.view;
  if (nr.method != SDFSLAM_NORMALS_PCA && nr.method != SDFSLAM_NORMALS_HYBRID) {
    throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "device eikonal normals need the pca or hybrid method");
  }
  if (c.scan_list.empty()) throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "no scan points for the eikonal normals");
  std::vector<double> cs(2 * static_cast<size_t>(v.num_poses));
  if (nr.frame_cos_sin != nullptr) {
    std::copy(nr.frame_cos_sin, nr.frame_cos_sin + cs.size(), cs.begin());
  } else {
    // One ::sincos per frame angle (the host's Pose2 rotation does the same).
    for (int f = 0; f < v.num_poses; ++f) ::sincos(poses[3 * f + 2], &cs[2 * f + 1], &cs[2 * f]);
  }
  const int64_t nodes = static_cast<int64_t>(v.grid.nx) * v.grid.ny;
  const size_t work_bytes = EikonalNormalsWorkBytes(static_cast<int64_t>(c.scan_list.size()), nodes);
  EpochPool& E = c.normals;
  double* d_cs = nullptr;
  unsigned char* d_work = nullptr;
  auto layout = [&]() {
    d_cs = E.Upload(cs);
    d_work = E.Alloc<unsigned char>(work_bytes);
  };
  E.BeginMeasure();
  layout();
  E.BeginPlace();
  layout();
  const DeviceProblem& P = c.dp;
  EikonalNormalsJob job{};
  job.scan_rows = c.scan_list_d;
  job.scan_points = static_cast<int64_t>(c.scan_list.size());
  job.frame = P.pt_frame;
  job.sensor_x = P.pt_x;
  job.sensor_y = P.pt_y;
  job.cos_sin = d_cs;
  job.poses = c.dw.poses;
  job.nx = v.grid.nx;
  job.ny = v.grid.ny;
  job.min_x = v.grid.min_x;
  job.min_y = v.grid.min_y;
  job.dx = v.grid.dx;
  job.dy = v.grid.dy;
  job.k = nr.k_neighbors;
  job.method = nr.method;
  job.corner_threshold = nr.corner_threshold;
  job.eik_w = P.eik_w;
  job.eik_h = P.eik_h;
  job.num_eikonal = P.num_eikonal;
  // The spec arrays live in the context's own base pool.
  job.eik_nx = const_cast<double*>(P.eik_nx);
  job.eik_ny = const_cast<double*>(P.eik_ny);
  Enqueue(c, [&E, job, d_work](hipStream_t s) {
    E.Copy(s);
    const int st = EnqueueEikonalNormals(job, d_work, s);
    if (st != SDFSLAM_OK) throw Failure(st, NormalsLastError());
  });
}

void UpdateContext(sdfslam_hip_context& c, const sdfslam_problem_view* problem,
                   const double* map_values, const double* poses,
                   const sdfslam_normals_request* normals = nullptr) {
  if (problem == nullptr || map_values == nullptr || poses == nullptr) {
    throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "null argument");
  }
  HIPC(hipSetDevice(c.device));
  auto hp = std::make_unique<HostProblem>(*problem);
  DropSpeculation(c);
  // An appended frame (same grid, the poses extended): the earlier poses
  // keep the margins they grew to, and the new ones start wider, since a new
  // frame's pose moves most in its first increments.
  const sdfslam_grid& og = c.hp->view.grid;
  const sdfslam_grid& ng = problem->grid;
  const int old_poses = c.hp->view.num_poses;
  const bool extends = og.nx == ng.nx && og.ny == ng.ny && og.min_x == ng.min_x && og.min_y == ng.min_y &&
                       og.dx == ng.dx && og.dy == ng.dy && problem->num_poses >= old_poses;
  std::vector<int32_t> margins;
  if (extends) {
    margins = c.aopt.pose_margin;
    if (static_cast<int>(margins.size()) != old_poses) margins.assign(static_cast<size_t>(old_poses), c.aopt.margin);
    margins.resize(static_cast<size_t>(problem->num_poses),
                   problem->num_poses <= kEarlyFrames ? kEarlyMargin : kFreshPoseMargin);
  }
  c.hp = std::move(hp);
  c.aopt.margin = AnalysisOptions{}.margin;
  c.aopt.pose_margin = std::move(margins);
  // Work an earlier update deferred is superseded (the new layout uploads
  // everything again).
  c.pending.clear();
  c.init_pending = false;
  struct Deferring {
    sdfslam_hip_context& c;
    explicit Deferring(sdfslam_hip_context& ctx) : c(ctx) { c.deferring = true; }
    ~Deferring() { c.deferring = false; }
  } deferring(c);
  ComputeScanList(c);
  ComputeReach(c, 0);
  PlaceRows(c, static_cast<int64_t>(c.hp->pt_frame.size()), static_cast<int64_t>(c.scan_list.size()));
  AllocateBase(c, map_values, poses);
  if (normals != nullptr) StageDeviceNormals(c, *normals, poses);
  BuildEpoch(c, map_values, poses);
  c.init_pending = true;
  c.init_lambda = c.opts.lambda_init;
  c.lm_ready = true;
}

/// Extends the context's problem by `frames` frames built on the device
/// (sdfslam_hip_context_append_frames): the host copy gets the same rows (the
/// symbolic analysis reads them), the device generates its own from the raw
/// scans; the earlier rows stay where they are unless the row pool grows.
void AppendFrames(sdfslam_hip_context& c, int32_t frames, const int64_t* counts, const double* scan_xy,
                  const double* odometry, const sdfslam_frame_options& o, const double* map_values,
                  const double* poses) {
  if (frames < 1 || counts == nullptr || odometry == nullptr || map_values == nullptr || poses == nullptr) {
    throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "append needs frames, counts, odometry and the state");
  }
  if (o.normals.frame_cos_sin != nullptr) {
    throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "appended frames take cos and sin from ::sincos (frame_cos_sin NULL)");
  }
  if (o.hallucination_points > 0 && !(o.hallucination_step > 0.0)) {
    throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "appended frames need a positive hallucination step");
  }
  if (o.active_region != 0 && o.active_margin < 0) {
    throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "negative active margin");
  }
  HostProblem& hp = *c.hp;
  const int old_poses = hp.view.num_poses;
  for (int k = 0; k < static_cast<int>(hp.od_i.size()); ++k) {
    if (hp.od_i[static_cast<size_t>(k)] != k || hp.od_j[static_cast<size_t>(k)] != k + 1) {
      throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "appending frames needs consecutive odometry only (no relations)");
    }
  }
  if (static_cast<int>(hp.od_i.size()) != old_poses - 1) {
    throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "appending frames needs one odometry spec per earlier frame");
  }
  int64_t total_points = 0;
  for (int f = 0; f < frames; ++f) {
    if (counts[f] < 0) throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "negative point count");
    total_points += counts[f];
  }
  if (total_points > 0 && scan_xy == nullptr) throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "null scan points");
  if (o.active_region == 0 && static_cast<int64_t>(hp.view.num_active_nodes) !=
                                   static_cast<int64_t>(hp.view.grid.nx) * hp.view.grid.ny) {
    throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "a problem with an active region needs active_region set");
  }
  HIPC(hipSetDevice(c.device));
  DropSpeculation(c);
  // Deferred work of an earlier update that no call has run yet is
  // superseded; the rows it would have written are uploaded again.
  const bool rows_on_device = c.pending.empty();
  c.pending.clear();
  c.init_pending = false;
  const auto old_rows = static_cast<int64_t>(hp.pt_frame.size());
  const auto old_list = static_cast<int64_t>(c.scan_list.size());
  // Host copy of the new rows (Problem::Problem, problem.cpp:26-37, with
  // GenerateHallucinatedPoints, hallucination.cpp:5-28), the same
  // arithmetic as the device's KFrameRows.
  const double sw_scan = std::sqrt(o.weight_scan);
  const double sw_hall = std::sqrt(o.weight_hallucination);
  const int per_side = o.hallucination_points > 0
                           ? o.hallucination_points / (o.hallucination_both_directions != 0 ? 2 : 1)
                           : 0;
  std::vector<int64_t> scan_off(static_cast<size_t>(frames) + 1, 0);
  std::vector<int64_t> row_off(static_cast<size_t>(frames));
  std::vector<int64_t> list_off(static_cast<size_t>(frames));
  for (int f = 0; f < frames; ++f) {
    const int frame = old_poses + f;
    row_off[static_cast<size_t>(f)] = static_cast<int64_t>(hp.pt_frame.size());
    list_off[static_cast<size_t>(f)] = static_cast<int64_t>(c.scan_list.size());
    const double* xy = scan_xy + 2 * scan_off[static_cast<size_t>(f)];
    const int64_t n = counts[f];
    for (int64_t i = 0; i < n; ++i) {
      c.scan_list.push_back(static_cast<int32_t>(hp.pt_frame.size()));
      hp.pt_frame.push_back(frame);
      hp.pt_x.push_back(xy[2 * i]);
      hp.pt_y.push_back(xy[2 * i + 1]);
      hp.pt_delta.push_back(0.0);
      hp.pt_sw.push_back(sw_scan);
    }
    if (per_side > 0) {
      for (int64_t i = 0; i < n; ++i) {
        const double px = xy[2 * i];
        const double py = xy[2 * i + 1];
        const double norm = std::sqrt(px * px + py * py);
        if (norm < 1e-9) continue;
        const double tx = -px / norm * o.hallucination_step;
        const double ty = -py / norm * o.hallucination_step;
        for (int j = 1; j <= per_side; ++j) {
          const double dj = static_cast<double>(j);
          hp.pt_frame.push_back(frame);
          hp.pt_x.push_back(px + tx * dj);
          hp.pt_y.push_back(py + ty * dj);
          hp.pt_delta.push_back(o.hallucination_step * dj);
          hp.pt_sw.push_back(sw_hall);
          if (o.hallucination_both_directions != 0) {
            hp.pt_frame.push_back(frame);
            hp.pt_x.push_back(px - tx * dj);
            hp.pt_y.push_back(py - ty * dj);
            hp.pt_delta.push_back(-o.hallucination_step * dj);
            hp.pt_sw.push_back(sw_hall);
          }
        }
      }
    }
    scan_off[static_cast<size_t>(f) + 1] = scan_off[static_cast<size_t>(f)] + n;
  }
  if (static_cast<int64_t>(hp.pt_frame.size()) > (int64_t{1} << 31) - 1) {
    throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "too many point rows");
  }
  // Odometry (problem.cpp:96-100): frame k-1 -> k for every new frame.
  const double sw_odo = std::sqrt(o.weight_odometry);
  for (int f = 0; f < frames; ++f) {
    const int j = old_poses + f;
    hp.od_i.push_back(j - 1);
    hp.od_j.push_back(j);
    hp.od_dx.push_back(odometry[3 * f]);
    hp.od_dy.push_back(odometry[3 * f + 1]);
    hp.od_dth.push_back(odometry[3 * f + 2]);
    hp.od_sw.push_back(sw_odo);
  }
  const int num_poses = old_poses + frames;
  const sdfslam_grid g = hp.view.grid;
  const int64_t nodes = static_cast<int64_t>(g.nx) * g.ny;
  std::vector<double> cs(2 * static_cast<size_t>(num_poses));
  for (int f = 0; f < num_poses; ++f) ::sincos(poses[3 * f + 2], &cs[2 * f + 1], &cs[2 * f]);
  // Row pool: grow (uploading the earlier rows) or keep; the device writes
  // the new rows either way.
  const bool grow = static_cast<int64_t>(hp.pt_frame.size()) > c.rows_cap ||
                    static_cast<int64_t>(c.scan_list.size()) > c.list_cap;
  if (grow || !rows_on_device) {
    struct Deferring {
      sdfslam_hip_context& c;
      explicit Deferring(sdfslam_hip_context& ctx) : c(ctx) { c.deferring = true; }
      ~Deferring() { c.deferring = false; }
    } deferring(c);
    PlaceRows(c, old_rows, old_list);
  }
  // Raw scans and their offsets, staged for the frame-row kernel.
  EpochPool& F = c.frames;
  double* d_xy = nullptr;
  int64_t* d_scan_off = nullptr;
  int64_t* d_row_off = nullptr;
  int64_t* d_list_off = nullptr;
  double* d_cs = nullptr;
  double* d_poses = nullptr;
  uint8_t* d_marks = nullptr;
  auto layout = [&]() {
    d_xy = F.Upload(scan_xy, 2 * static_cast<size_t>(total_points));
    d_scan_off = F.Upload(scan_off);
    d_row_off = F.Upload(row_off);
    d_list_off = F.Upload(list_off);
    if (o.active_region != 0) {
      d_cs = F.Upload(cs);
      d_poses = F.Upload(poses, 3 * static_cast<size_t>(num_poses));
      d_marks = F.Alloc<uint8_t>(static_cast<size_t>(nodes));
    }
  };
  F.BeginMeasure();
  layout();
  F.BeginPlace();
  layout();
  FrameRowsJob job{};
  job.frames = frames;
  job.first_frame = old_poses;
  job.scan_xy = d_xy;
  job.scan_off = d_scan_off;
  job.row_off = d_row_off;
  job.list_off = d_list_off;
  job.per_side = per_side;
  job.both_directions = o.hallucination_both_directions;
  job.step = o.hallucination_step;
  job.sqrt_w_scan = sw_scan;
  job.sqrt_w_hallucination = sw_hall;
  job.frame = c.pt_frame_d;
  job.x = c.pt_x_d;
  job.y = c.pt_y_d;
  job.expected_sdf = c.pt_delta_d;
  job.sqrt_weight = c.pt_sw_d;
  job.scan_list = c.scan_list_d;
  auto rows_work = [&F, job](hipStream_t s) {
    F.Copy(s);
    const int st = EnqueueFrameRows(job, s);
    if (st != SDFSLAM_OK) throw Failure(st, NormalsLastError());
  };
  hp.view.num_poses = num_poses;
  if (o.active_region != 0) {
    // The active set at the given poses from every row on the device, read
    // back: the node columns and the eikonal specs follow it.
    Bind(c);
    rows_work(c.stream);
    const int64_t np = static_cast<int64_t>(hp.pt_frame.size());
    const int st = EnqueueActiveMarks(c.pt_frame_d, c.pt_x_d, c.pt_y_d, np, d_cs, d_poses, g, o.active_margin,
                                      d_marks, c.stream);
    if (st != SDFSLAM_OK) throw Failure(st, NormalsLastError());
    std::vector<uint8_t> marks(static_cast<size_t>(nodes));
    HIPC(hipMemcpyAsync(marks.data(), d_marks, marks.size(), hipMemcpyDeviceToHost, c.stream));
    HIPC(hipStreamSynchronize(c.stream));
    int active = 0;
    for (int64_t v = 0; v < nodes; ++v) {
      hp.node_to_active[static_cast<size_t>(v)] = marks[static_cast<size_t>(v)] != 0 ? active++ : -1;
    }
    hp.view.num_active_nodes = active;
  } else {
    // The dense problem keeps every node active; the device writes its rows
    // at the start of the next call.
    c.pending.push_back(rows_work);
  }
  // Eikonal specs (problem.cpp:72-94): every cell corner whose node and its
  // right and upper neighbours are active, h-major, with weight
  // sqrt(w_eikonal) (the pca and hybrid methods' eikonal scale is 1); their
  // normals come from the device.
  const double sw_eik = std::sqrt(o.weight_eikonal);
  hp.eik_w.clear();
  hp.eik_h.clear();
  hp.eik_sw.clear();
  for (int h = 0; h + 1 < g.ny; ++h) {
    for (int w = 0; w + 1 < g.nx; ++w) {
      const int64_t node = static_cast<int64_t>(h) * g.nx + w;
      if (hp.node_to_active[static_cast<size_t>(node)] < 0 || hp.node_to_active[static_cast<size_t>(node) + 1] < 0 ||
          hp.node_to_active[static_cast<size_t>(node + g.nx)] < 0) {
        continue;
      }
      hp.eik_w.push_back(w);
      hp.eik_h.push_back(h);
      hp.eik_sw.push_back(sw_eik);
    }
  }
  hp.eik_nx.assign(hp.eik_w.size(), 0.0);
  hp.eik_ny.assign(hp.eik_w.size(), 0.0);
  hp.Refresh();
  // Margins: the earlier poses keep theirs, the new ones start wider.
  std::vector<int32_t> margins = c.aopt.pose_margin;
  if (static_cast<int>(margins.size()) != old_poses) margins.assign(static_cast<size_t>(old_poses), c.aopt.margin);
  margins.resize(static_cast<size_t>(num_poses), num_poses <= kEarlyFrames ? kEarlyMargin : kFreshPoseMargin);
  c.aopt.margin = AnalysisOptions{}.margin;
  c.aopt.pose_margin = std::move(margins);
  ComputeReach(c, static_cast<size_t>(old_rows));
  struct Deferring {
    sdfslam_hip_context& c;
    explicit Deferring(sdfslam_hip_context& ctx) : c(ctx) { c.deferring = true; }
    ~Deferring() { c.deferring = false; }
  } deferring(c);
  AllocateBase(c, map_values, poses);
  sdfslam_normals_request nr = o.normals;
  nr.frame_cos_sin = cs.data();
  StageDeviceNormals(c, nr, poses);
  BuildEpoch(c, map_values, poses);
  c.init_pending = true;
  c.init_lambda = c.opts.lambda_init;
  c.lm_ready = true;
}

void FillResult(sdfslam_hip_context& c, const LmState& st, sdfslam_solve_result* result) {
  result->converged = st.converged;
  result->iterations = st.iter;
  result->initial_cost = st.initial_cost;
  result->final_cost = st.iter == 0 ? st.initial_cost : st.final_cost;
  const int len = st.iter + 1;
  result->cost_history_length = len;
  if (result->cost_history != nullptr && result->cost_history_capacity > 0) {
    const int n = std::min({len, result->cost_history_capacity, c.history_capacity});
    HIPC(hipMemcpyAsync(result->cost_history, c.dw.cost_history, static_cast<size_t>(n) * sizeof(double),
                        hipMemcpyDeviceToHost, c.stream));
    HIPC(hipStreamSynchronize(c.stream));
  }
}

int SolveContext(sdfslam_hip_context& c, const sdfslam_solver_options& options,
                 sdfslam_solve_result* result) {
  if (options.max_iterations + 1 > c.history_capacity) {
    c.history_capacity = options.max_iterations + 1;
    c.dw.cost_history = c.hist.Alloc<double>(static_cast<size_t>(c.history_capacity));
    SyncSlot(c);
    ++c.generation;
  }
  const LmOptions o = ToLm(options, c.history_capacity);
  InitLm(c, options.lambda_init);
  const LmState st = RunIterations(c, o, std::max(options.max_iterations, 0), false);
  if (st.not_spd != 0) return StateStatus(st);
  if (result != nullptr) FillResult(c, st, result);
  return SDFSLAM_OK;
}

}  // namespace

extern "C" {

const char* sdfslam_hip_status_string(int status) {
  switch (status) {
    case SDFSLAM_OK: return "ok";
    case SDFSLAM_ERR_INVALID_ARGUMENT: return "invalid argument";
    case SDFSLAM_ERR_NOT_SPD: return "sparse Cholesky factorization failed";
    case SDFSLAM_ERR_HIP: return "HIP runtime error";
    case SDFSLAM_ERR_OOM: return "out of device memory";
    case SDFSLAM_ERR_NO_DEVICE: return "no HIP device";
    case SDFSLAM_ERR_INTERNAL: return "internal error";
    default: return "unknown status";
  }
}

const char* sdfslam_hip_last_error(void) { return g_last_error.c_str(); }

int sdfslam_hip_context_create(const sdfslam_problem_view* problem,
                               const sdfslam_solver_options* options, const double* map_values,
                               const double* poses, sdfslam_hip_context** out) {
  return Guard([&]() -> int {
    CreateContext(problem, options, map_values, poses, out);
    return SDFSLAM_OK;
  });
}

int sdfslam_hip_device_memory(int64_t* free_bytes, int64_t* total_bytes) {
  return Guard([&]() -> int {
    if (free_bytes == nullptr || total_bytes == nullptr) {
      throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "null argument");
    }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
      throw Failure(SDFSLAM_ERR_NO_DEVICE, "no HIP device");
    }
    size_t f = 0;
    size_t t = 0;
    HIPC(hipMemGetInfo(&f, &t));
    *free_bytes = static_cast<int64_t>(f);
    *total_bytes = static_cast<int64_t>(t);
    return SDFSLAM_OK;
  });
}

int sdfslam_hip_context_update(sdfslam_hip_context* ctx, const sdfslam_problem_view* problem,
                               const double* map_values, const double* poses) {
  return Guard([&]() -> int {
    if (ctx == nullptr) throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "null argument");
    UpdateContext(*ctx, problem, map_values, poses);
    return SDFSLAM_OK;
  });
}

int sdfslam_hip_context_update_device_normals(sdfslam_hip_context* ctx, const sdfslam_problem_view* problem,
                                              const double* map_values, const double* poses,
                                              const sdfslam_normals_request* normals) {
  return Guard([&]() -> int {
    if (ctx == nullptr || normals == nullptr) throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "null argument");
    UpdateContext(*ctx, problem, map_values, poses, normals);
    return SDFSLAM_OK;
  });
}

int sdfslam_hip_context_append_frames(sdfslam_hip_context* ctx, int32_t num_frames,
                                      const int64_t* points_per_frame, const double* scan_xy,
                                      const double* odometry, const sdfslam_frame_options* options,
                                      const double* map_values, const double* poses) {
  return Guard([&]() -> int {
    if (ctx == nullptr || options == nullptr) throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "null argument");
    AppendFrames(*ctx, num_frames, points_per_frame, scan_xy, odometry, *options, map_values, poses);
    return SDFSLAM_OK;
  });
}

int sdfslam_hip_context_point_rows(sdfslam_hip_context* ctx, int64_t capacity, int64_t* num_points,
                                   int32_t* frame, double* x, double* y, double* expected_sdf,
                                   double* sqrt_weight) {
  return Guard([&]() -> int {
    if (ctx == nullptr || capacity < 0) throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "bad argument");
    sdfslam_hip_context& c = *ctx;
    Bind(c);
    if (num_points != nullptr) *num_points = c.hp->view.num_points;
    const auto n = static_cast<size_t>(std::min<int64_t>(capacity, c.hp->view.num_points));
    if (frame != nullptr) {
      HIPC(hipMemcpyAsync(frame, c.pt_frame_d, n * sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
    }
    if (x != nullptr) HIPC(hipMemcpyAsync(x, c.pt_x_d, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    if (y != nullptr) HIPC(hipMemcpyAsync(y, c.pt_y_d, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    if (expected_sdf != nullptr) {
      HIPC(hipMemcpyAsync(expected_sdf, c.pt_delta_d, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    }
    if (sqrt_weight != nullptr) {
      HIPC(hipMemcpyAsync(sqrt_weight, c.pt_sw_d, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    }
    HIPC(hipStreamSynchronize(c.stream));
    return SDFSLAM_OK;
  });
}

int sdfslam_hip_set_graphs(int32_t enabled) {
  g_graphs.store(enabled != 0);
  return SDFSLAM_OK;
}

int sdfslam_hip_set_lds_forms(int32_t enabled) {
  SetLdsForms(enabled != 0);
  return SDFSLAM_OK;
}

int sdfslam_hip_pose_margins(sdfslam_hip_context* ctx, int32_t* margins, int32_t capacity,
                             int32_t* num_poses) {
  return Guard([&]() -> int {
    if (ctx == nullptr || capacity < 0) throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "bad argument");
    const int n = ctx->hp->view.num_poses;
    if (num_poses != nullptr) *num_poses = n;
    if (margins != nullptr) {
      const std::vector<int32_t>& pm = ctx->aopt.pose_margin;
      for (int P = 0; P < std::min(n, capacity); ++P) {
        margins[P] = static_cast<int>(pm.size()) == n ? pm[static_cast<size_t>(P)] : ctx->aopt.margin;
      }
    }
    return SDFSLAM_OK;
  });
}

int sdfslam_hip_context_trim(sdfslam_hip_context* ctx) {
  return Guard([&]() -> int {
    if (ctx == nullptr) throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "null context");
    Bind(*ctx);
    HIPC(hipStreamSynchronize(ctx->stream));
    ctx->epoch.Trim();
    ctx->base.Trim();
    ctx->normals.Trim();
    ctx->rows.Trim();
    ctx->frames.Trim();
    ctx->batch.Trim();
    return SDFSLAM_OK;
  });
}

int sdfslam_hip_context_destroy(sdfslam_hip_context* ctx) {
  if (ctx != nullptr) (void)hipSetDevice(ctx->device);
  delete ctx;
  return SDFSLAM_OK;
}

int sdfslam_hip_set_state(sdfslam_hip_context* ctx, const double* map_values, const double* poses) {
  return Guard([&]() -> int {
    if (ctx == nullptr) throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "null context");
    Bind(*ctx);
    DropSpeculation(*ctx);
    const sdfslam_problem_view& v = ctx->hp->view;
    if (map_values != nullptr) {
      HIPC(hipMemcpyAsync(ctx->dw.map, map_values, sizeof(double) * v.grid.nx * v.grid.ny,
                          hipMemcpyHostToDevice, ctx->stream));
    }
    if (poses != nullptr) {
      HIPC(hipMemcpyAsync(ctx->dw.poses, poses, sizeof(double) * 3 * v.num_poses, hipMemcpyHostToDevice,
                          ctx->stream));
    }
    BuildEpoch(*ctx);
    HIPC(hipStreamSynchronize(ctx->stream));
    return SDFSLAM_OK;
  });
}

int sdfslam_hip_get_state(sdfslam_hip_context* ctx, double* map_values, double* poses) {
  return Guard([&]() -> int {
    if (ctx == nullptr) throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "null context");
    Bind(*ctx);
    std::vector<double> m, p;
    DownloadState(*ctx, m, p);
    if (map_values != nullptr) std::copy(m.begin(), m.end(), map_values);
    if (poses != nullptr) std::copy(p.begin(), p.end(), poses);
    return SDFSLAM_OK;
  });
}

int sdfslam_hip_analysis(sdfslam_hip_context* ctx, sdfslam_hip_analysis_info* info) {
  return Guard([&]() -> int {
    if (ctx == nullptr || info == nullptr) throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "null argument");
    const Analysis& a = ctx->an;
    info->num_parameters = a.n;
    info->num_residuals = ctx->hp->rows();
    info->pattern_nnz_lower = a.pattern_nnz_lower;
    info->num_supernodes = a.num_sn;
    info->factor_nnz = a.factor_nnz;
    info->factor_flops = a.factor_flops;
    info->update_doubles = a.upd_total;
    info->tree_levels = a.num_levels;
    info->pattern_epochs = ctx->pattern_epochs;
    info->analysis_seconds = ctx->analysis_seconds;
    info->device_bytes =
        static_cast<int64_t>(ctx->base.bytes() + ctx->epoch.bytes() + ctx->normals.bytes() + ctx->rows.bytes() +
                             ctx->frames.bytes() + ctx->hist.bytes());
    info->panel_step_launches = a.panel_step_launches;
    info->panel_step_flops = a.panel_step_flops;
    info->speculative_epochs = ctx->speculative_epochs;
    info->gemm_launches = a.gemm_launches;
    info->gemm_flops = a.gemm_flops;
    info->diag_blocks = 0;
    for (int s = 0; s < a.num_sn; ++s) {
      info->diag_blocks += (a.sn_first[static_cast<size_t>(s) + 1] - a.sn_first[static_cast<size_t>(s)] + 63) / 64;
    }
    return SDFSLAM_OK;
  });
}

int sdfslam_hip_assemble(sdfslam_hip_context* ctx, double* residuals, int32_t* j_cols,
                         double* j_vals) {
  return Guard([&]() -> int {
    if (ctx == nullptr || residuals == nullptr || j_cols == nullptr || j_vals == nullptr) {
      throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "null argument");
    }
    sdfslam_hip_context& c = *ctx;
    Bind(c);
    PrepareStage(c, nullptr);
    LaunchAssemble(SlotArgs{c.slot_dev, &c.slot}, c.shape, false, c.stream);
    LaunchAssembleEikonalJ(SlotArgs{c.slot_dev, &c.slot}, c.shape, c.j_eik, c.stream);
    const sdfslam_problem_view& v = c.hp->view;
    const auto np = static_cast<size_t>(v.num_points);
    const auto ne = static_cast<size_t>(v.num_eikonal);
    const auto no = static_cast<size_t>(v.num_odom);
    std::vector<double> r(np), jw(4 * np), jp(3 * np), re(ne), je(3 * ne), odr(3 * no), oji(9 * no),
        ojj(9 * no);
    std::vector<uint32_t> key(np);
    auto d2h = [&](void* dst, const void* src, size_t bytes) {
      if (bytes > 0) HIPC(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c.stream));
    };
    d2h(r.data(), c.dw.r_pt, np * 8);
    d2h(jw.data(), c.dw.jw, 4 * np * 8);
    d2h(jp.data(), c.dw.jp, 3 * np * 8);
    d2h(key.data(), c.dw.cellkey, np * 4);
    d2h(re.data(), c.dw.r_eik, ne * 8);
    d2h(je.data(), c.j_eik, 3 * ne * 8);
    d2h(odr.data(), c.dw.od_r, 3 * no * 8);
    d2h(oji.data(), c.dw.od_ji, 9 * no * 8);
    d2h(ojj.data(), c.dw.od_jj, 9 * no * 8);
    HIPC(hipStreamSynchronize(c.stream));
    // Row-ELL in the reference's triplet order within each row.
    const int nx = v.grid.nx;
    const int ncx = nx - 1;
    const int na = v.num_active_nodes;
    const uint32_t invalid = static_cast<uint32_t>(ncx * (v.grid.ny - 1));
    for (size_t s = 0; s < np; ++s) {
      residuals[s] = r[s];
      int32_t* cols = j_cols + 7 * s;
      double* vals = j_vals + 7 * s;
      int k = 0;
      if (key[s] < invalid) {
        const int cw = static_cast<int>(key[s]) % ncx;
        const int ch = static_cast<int>(key[s]) / ncx;
        const int ids[4] = {ch * nx + cw, ch * nx + cw + 1, (ch + 1) * nx + cw, (ch + 1) * nx + cw + 1};
        for (int t = 0; t < 4; ++t) {
          const int col = v.node_to_active[ids[t]];
          if (col >= 0) {
            cols[k] = col;
            vals[k++] = jw[t * np + s];
          }
        }
        const int f = v.point_frame[s];
        if (f >= 1) {
          for (int t = 0; t < 3; ++t) {
            cols[k] = na + 3 * (f - 1) + t;
            vals[k++] = jp[t * np + s];
          }
        }
      }
      for (; k < 7; ++k) {
        cols[k] = -1;
        vals[k] = 0.0;
      }
    }
    for (size_t e = 0; e < ne; ++e) {
      const size_t row = np + e;
      residuals[row] = re[e];
      int32_t* cols = j_cols + 7 * row;
      double* vals = j_vals + 7 * row;
      const int g = v.eik_h[e] * nx + v.eik_w[e];
      const int ids[3] = {g, g + 1, g + nx};
      int k = 0;
      for (int t = 0; t < 3; ++t) {
        const int col = v.node_to_active[ids[t]];
        if (col >= 0) {
          cols[k] = col;
          vals[k++] = je[3 * e + t];
        }
      }
      for (; k < 7; ++k) {
        cols[k] = -1;
        vals[k] = 0.0;
      }
    }
    for (size_t o = 0; o < no; ++o) {
      for (int rr = 0; rr < 3; ++rr) {
        const size_t row = np + ne + 3 * o + rr;
        residuals[row] = odr[3 * o + rr];
        int32_t* cols = j_cols + 7 * row;
        double* vals = j_vals + 7 * row;
        int k = 0;
        const int fi = v.odom_i[o];
        const int fj = v.odom_j[o];
        if (fi >= 1) {
          for (int t = 0; t < 3; ++t) {
            cols[k] = na + 3 * (fi - 1) + t;
            vals[k++] = oji[9 * o + 3 * rr + t];
          }
        }
        for (int t = 0; t < 3; ++t) {
          cols[k] = na + 3 * (fj - 1) + t;
          vals[k++] = ojj[9 * o + 3 * rr + t];
        }
        for (; k < 7; ++k) {
          cols[k] = -1;
          vals[k] = 0.0;
        }
      }
    }
    return SDFSLAM_OK;
  });
}

int sdfslam_hip_normal_equations(sdfslam_hip_context* ctx, double lambda, int64_t capacity,
                                 int64_t* nnz, int32_t* col_ptr, int32_t* row_idx, double* values,
                                 uint8_t* structural, double* rhs) {
  return Guard([&]() -> int {
    if (ctx == nullptr || nnz == nullptr) throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "null argument");
    sdfslam_hip_context& c = *ctx;
    Bind(c);
    const LmOptions o = ToLm(c.opts, c.history_capacity);
    for (int attempt = 0; attempt < 2; ++attempt) {
      PrepareStage(c, &lambda);
      EnqueueAssemblyAndNormal(c, o, false);
      LmState st{};
      ReadState(c, st);
      if (st.overflow == 0) break;
      if (attempt == 1) throw Failure(SDFSLAM_ERR_INTERNAL, "pattern re-analysis failed");
      BuildEpoch(c);
    }
    const Analysis& an = c.an;
    *nnz = an.pattern_nnz_lower;
    if (an.pattern_nnz_lower > capacity) {
      g_last_error = "normal-equation buffers hold fewer entries than the pattern";
      return SDFSLAM_ERR_INVALID_ARGUMENT;
    }
    const std::vector<PatternEntry> pat = LowerPattern(an, c.hp->view);
    if (static_cast<int64_t>(pat.size()) != an.pattern_nnz_lower) {
      throw Failure(SDFSLAM_ERR_INTERNAL, "pattern size mismatch");
    }
    std::vector<double> panel(static_cast<size_t>(an.panel_total));
    HIPC(hipMemcpyAsync(panel.data(), c.dw.panel, panel.size() * sizeof(double), hipMemcpyDeviceToHost,
                        c.stream));
    HIPC(hipStreamSynchronize(c.stream));
    if (rhs != nullptr) {
      HIPC(hipMemcpyAsync(rhs, c.dw.rhs, sizeof(double) * an.n, hipMemcpyDeviceToHost, c.stream));
      HIPC(hipStreamSynchronize(c.stream));
    }
    if (col_ptr != nullptr) {
      std::fill(col_ptr, col_ptr + an.n + 1, 0);
      for (const PatternEntry& e : pat) ++col_ptr[e.col + 1];
      for (int j = 0; j < an.n; ++j) col_ptr[j + 1] += col_ptr[j];
    }
    for (size_t k = 0; k < pat.size(); ++k) {
      if (row_idx != nullptr) row_idx[k] = pat[k].row;
      if (values != nullptr) values[k] = panel[static_cast<size_t>(pat[k].off)];
    }
    if (structural != nullptr) {
      // Pairs of columns coupled by some Jacobian row of the device assembly.
      const HostProblem& hp = *c.hp;
      const int64_t rows = hp.rows();
      std::vector<double> r(static_cast<size_t>(rows));
      std::vector<int32_t> jc(7 * static_cast<size_t>(rows));
      std::vector<double> jv(7 * static_cast<size_t>(rows));
      const int st = sdfslam_hip_assemble(ctx, r.data(), jc.data(), jv.data());
      if (st != SDFSLAM_OK) throw Failure(st, g_last_error);
      std::vector<int64_t> keys;
      for (int64_t row = 0; row < rows; ++row) {
        const int32_t* cs = jc.data() + 7 * row;
        for (int a = 0; a < 7 && cs[a] >= 0; ++a) {
          for (int b = 0; b < 7 && cs[b] >= 0; ++b) {
            if (cs[a] >= cs[b]) keys.push_back(static_cast<int64_t>(cs[a]) * an.n + cs[b]);
          }
        }
      }
      std::sort(keys.begin(), keys.end());
      keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
      for (size_t k = 0; k < pat.size(); ++k) {
        const int64_t key = static_cast<int64_t>(pat[k].row) * an.n + pat[k].col;
        structural[k] = (pat[k].row == pat[k].col || std::binary_search(keys.begin(), keys.end(), key)) ? 1 : 0;
      }
    }
    return SDFSLAM_OK;
  });
}

int sdfslam_hip_elimination_order(sdfslam_hip_context* ctx, int32_t* order) {
  return Guard([&]() -> int {
    if (ctx == nullptr || order == nullptr) throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "null argument");
    std::copy(ctx->an.iperm.begin(), ctx->an.iperm.end(), order);
    return SDFSLAM_OK;
  });
}

int sdfslam_hip_reference_factor_counts(sdfslam_hip_context* ctx, int64_t* factor_nnz, double* factor_flops) {
  return Guard([&]() -> int {
    if (ctx == nullptr || factor_nnz == nullptr || factor_flops == nullptr) {
      throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "null argument");
    }
    const int64_t rows = ctx->hp->rows();
    std::vector<double> r(static_cast<size_t>(rows));
    std::vector<int32_t> jc(7 * static_cast<size_t>(rows));
    std::vector<double> jv(7 * static_cast<size_t>(rows));
    const int st = sdfslam_hip_assemble(ctx, r.data(), jc.data(), jv.data());
    if (st != SDFSLAM_OK) throw Failure(st, g_last_error);
    ReferenceFactorCounts(ctx->an, rows, jc.data(), factor_nnz, factor_flops);
    return SDFSLAM_OK;
  });
}

int sdfslam_hip_linear_step(sdfslam_hip_context* ctx, double lambda, double* step) {
  return Guard([&]() -> int {
    if (ctx == nullptr || step == nullptr) throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "null argument");
    sdfslam_hip_context& c = *ctx;
    Bind(c);
    const LmOptions o = ToLm(c.opts, c.history_capacity);
    LmState st{};
    for (int attempt = 0; attempt < 2; ++attempt) {
      PrepareStage(c, &lambda);
      EnqueueAssemblyAndNormal(c, o, false);
      EnqueueFactorAndSolve(c, false);
      ReadState(c, st);
      if (st.overflow == 0) break;
      if (attempt == 1) throw Failure(SDFSLAM_ERR_INTERNAL, "pattern re-analysis failed");
      BuildEpoch(c);
    }
    if (st.not_spd != 0) return StateStatus(st);
    HIPC(hipMemcpyAsync(step, c.dw.step, sizeof(double) * c.an.n, hipMemcpyDeviceToHost, c.stream));
    HIPC(hipStreamSynchronize(c.stream));
    return SDFSLAM_OK;
  });
}

int sdfslam_hip_cost(sdfslam_hip_context* ctx, double* cost) {
  return Guard([&]() -> int {
    if (ctx == nullptr || cost == nullptr) throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "null argument");
    sdfslam_hip_context& c = *ctx;
    Bind(c);
    LaunchCost(SlotArgs{c.slot_dev, &c.slot}, c.shape, c.stream);
    LaunchSumPartials(SlotArgs{c.slot_dev, &c.slot}, c.scalar, c.stream);
    HIPC(hipMemcpyAsync(cost, c.scalar, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    HIPC(hipStreamSynchronize(c.stream));
    return SDFSLAM_OK;
  });
}

int sdfslam_hip_context_solve(sdfslam_hip_context* ctx, const sdfslam_solver_options* options,
                              sdfslam_solve_result* result) {
  return Guard([&]() -> int {
    if (ctx == nullptr || options == nullptr) throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "null argument");
    const auto t0 = std::chrono::steady_clock::now();
    Bind(*ctx);
    ctx->opts = *options;
    const int st = SolveContext(*ctx, *options, result);
    if (result != nullptr) {
      result->duration_seconds =
          std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return st;
  });
}

int sdfslam_hip_batch_solve(sdfslam_hip_context* const* contexts, int32_t count,
                            const sdfslam_solver_options* options, sdfslam_solve_result* results,
                            int32_t* statuses, sdfslam_hip_batch_info* info) {
  return Guard([&]() -> int {
    if (contexts == nullptr || options == nullptr || count < 1) {
      throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "batch needs contexts and options");
    }
    for (int b = 0; b < count; ++b) {
      if (contexts[b] == nullptr) throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "null context in batch");
      if (contexts[b]->device != contexts[0]->device) {
        throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "batch contexts on different devices");
      }
      for (int q = 0; q < b; ++q) {
        if (contexts[q] == contexts[b]) throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "context twice in batch");
      }
    }
    const auto t0 = std::chrono::steady_clock::now();
    RunBatchPending(contexts, count);
    BatchRun run(contexts, count);
    std::vector<int> status;
    std::vector<LmState> states;
    run.Run(*options, status, states);
    const double seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (info != nullptr) {
      *info = run.info();
      info->run_seconds = seconds;
    }
    for (int b = 0; b < count; ++b) {
      if (statuses != nullptr) statuses[b] = status[static_cast<size_t>(b)];
      if (results != nullptr) {
        FillResult(*contexts[b], states[static_cast<size_t>(b)], &results[b]);
        results[b].duration_seconds = seconds;
      }
    }
    return SDFSLAM_OK;
  });
}

int sdfslam_hip_init_from_scans(const double* points_xy, const double* origins_xy, int64_t num_points,
                                const sdfslam_grid* grid, int32_t signed_distance, double* map_values) {
  return Guard([&]() -> int {
    if (grid == nullptr || map_values == nullptr || num_points < 0 ||
        (num_points > 0 && (points_xy == nullptr || origins_xy == nullptr))) {
      throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "null argument");
    }
    if (grid->nx < 2 || grid->ny < 2) throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "bad grid");
    const int st = InitFromScansOnDevice(points_xy, origins_xy, num_points, *grid, signed_distance, map_values);
    if (st != SDFSLAM_OK) throw Failure(st, NormalsLastError());
    return SDFSLAM_OK;
  });
}

int sdfslam_hip_init_from_scans_noalloc(const double* points_xy, const double* origins_xy,
                                        int64_t num_points, const sdfslam_grid* grid,
                                        int32_t signed_distance, double* map_values, void* stream) {
  static_assert(kNoAllocMaxPoints == SDFSLAM_NO_ALLOC_MAX_POINTS, "header and kernels disagree");
  return Guard([&]() -> int {
    if (grid == nullptr || num_points < 0) throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "null argument");
    const int st = InitFromScansNoAlloc(points_xy, origins_xy, num_points, *grid, signed_distance, map_values,
                                        static_cast<hipStream_t>(stream));
    if (st != SDFSLAM_OK) throw Failure(st, NormalsLastError());
    return SDFSLAM_OK;
  });
}

int sdfslam_hip_node_normals(const double* points_xy, const double* origins_xy, int64_t num_points,
                             int32_t k_neighbors, const sdfslam_grid* grid, int32_t method,
                             double corner_threshold, double* node_normal_xy, double* node_scale) {
  return Guard([&]() -> int {
    if (grid == nullptr || node_normal_xy == nullptr || node_scale == nullptr || num_points < 0 ||
        (num_points > 0 && (points_xy == nullptr || origins_xy == nullptr))) {
      throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "null argument");
    }
    if (grid->nx < 2 || grid->ny < 2 || method < SDFSLAM_NORMALS_PCA || method > SDFSLAM_NORMALS_WEIGHTED) {
      throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "bad grid or normal method");
    }
    const int st = NodeNormalsOnDevice(points_xy, origins_xy, num_points, k_neighbors, *grid, method,
                                       corner_threshold, node_normal_xy, node_scale);
    if (st != SDFSLAM_OK) throw Failure(st, NormalsLastError());
    return SDFSLAM_OK;
  });
}

int sdfslam_hip_context_iterate(sdfslam_hip_context* ctx, int32_t iterations,
                                int32_t collect_phase_times) {
  return Guard([&]() -> int {
    if (ctx == nullptr || iterations < 0) throw Failure(SDFSLAM_ERR_INVALID_ARGUMENT, "bad argument");
    sdfslam_hip_context& c = *ctx;
    Bind(c);
    LmOptions o = ToLm(c.opts, c.history_capacity);
    o.max_iterations = INT_MAX;
    for (double& p : c.phase_ms) p = 0.0;
    c.timed_iterations = 0;
    c.panel_step_ms = 0.0;
    c.panel_steps_timed = 0;
    c.gemm_ms = 0.0;
    c.gemms_timed = 0;
    LmState st{};
    ReadState(c, st);
    if (st.halt != 0) {
      st.halt = 0;
      WriteState(c, st);
    }
    st = RunIterations(c, o, iterations, collect_phase_times != 0);
    if (st.not_spd != 0) return StateStatus(st);
    return SDFSLAM_OK;
  });
}

int sdfslam_hip_phase_times_get(sdfslam_hip_context* ctx, sdfslam_hip_phase_times* out) {
  if (ctx == nullptr || out == nullptr) return SDFSLAM_ERR_INVALID_ARGUMENT;
  const double n = std::max(1, ctx->timed_iterations);
  out->assembly_ms = ctx->phase_ms[kAssembly] / n;
  out->normal_ms = ctx->phase_ms[kNormal] / n;
  out->factor_ms = ctx->phase_ms[kFactor] / n;
  out->solve_ms = ctx->phase_ms[kSolve] / n;
  out->update_ms = ctx->phase_ms[kUpdate] / n;
  out->total_ms = (ctx->phase_ms[0] + ctx->phase_ms[1] + ctx->phase_ms[2] + ctx->phase_ms[3] +
                   ctx->phase_ms[4]) / n;
  out->point_assembly_kernel_ms = ctx->phase_ms[kPhases] / n;
  out->panel_step_kernel_ms =
      ctx->panel_steps_timed > 0 ? ctx->panel_step_ms / static_cast<double>(ctx->panel_steps_timed) : 0.0;
  out->gemm_kernel_ms = ctx->gemms_timed > 0 ? ctx->gemm_ms / static_cast<double>(ctx->gemms_timed) : 0.0;
  return SDFSLAM_OK;
}

int sdfslam_hip_solve(const sdfslam_problem_view* problem, const sdfslam_solver_options* options,
                      double* map_values, double* poses, sdfslam_solve_result* result) {
  return Guard([&]() -> int {
    const auto t0 = std::chrono::steady_clock::now();
    sdfslam_hip_context* raw = nullptr;
    CreateContext(problem, options, map_values, poses, &raw);
    std::unique_ptr<sdfslam_hip_context> ctx(raw);
    const int st = SolveContext(*ctx, *options, result);
    if (st != SDFSLAM_OK) return st;
    std::vector<double> m, p;
    DownloadState(*ctx, m, p);
    std::copy(m.begin(), m.end(), map_values);
    std::copy(p.begin(), p.end(), poses);
    if (result != nullptr) {
      result->duration_seconds =
          std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return SDFSLAM_OK;
  });
}

}  // extern "C"
