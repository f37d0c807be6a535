// Synthetic workload generators (see scene.hpp).
#include "scene.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace sdf_slam {

namespace {

uint64_t SplitMix64(uint64_t& x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

uint64_t Rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

constexpr double kPi = 3.14159265358979323846;

}  // namespace

Rng::Rng(uint64_t seed) {
  uint64_t x = seed;
  for (auto& s : s_) {
    s = SplitMix64(x);
  }
}

uint64_t Rng::Next() {
  const uint64_t result = Rotl(s_[1] * 5, 7) * 9;
  const uint64_t t = s_[1] << 17;
  s_[2] ^= s_[0];
  s_[3] ^= s_[1];
  s_[1] ^= s_[2];
  s_[0] ^= s_[3];
  s_[2] ^= t;
  s_[3] = Rotl(s_[3], 45);
  return result;
}

double Rng::Uniform() { return static_cast<double>(Next() >> 11) * 0x1.0p-53; }

double Rng::Uniform(double lo, double hi) { return lo + (hi - lo) * Uniform(); }

double Rng::Normal() {
  if (has_spare_) {
    has_spare_ = false;
    return spare_;
  }
  double u1 = Uniform();
  while (u1 <= 0.0) {
    u1 = Uniform();
  }
  const double u2 = Uniform();
  const double r = std::sqrt(-2.0 * std::log(u1));
  spare_ = r * std::sin(2.0 * kPi * u2);
  has_spare_ = true;
  return r * std::cos(2.0 * kPi * u2);
}

double CastRay(const std::vector<Segment>& walls, const Vec2& origin, double angle,
               double max_range) {
  const Vec2 d{std::cos(angle), std::sin(angle)};
  double best = max_range;
  bool hit = false;
  for (const Segment& s : walls) {
    const Vec2 e = s.b - s.a;
    const double denom = d.x * e.y - d.y * e.x;
    if (std::abs(denom) < 1e-14) {
      continue;
    }
    const Vec2 w = s.a - origin;
    const double t = (w.x * e.y - w.y * e.x) / denom;
    const double u = (w.x * d.y - w.y * d.x) / denom;
    if (t > 1e-9 && u >= 0.0 && u <= 1.0 && t <= best) {
      best = t;
      hit = true;
    }
  }
  return hit ? best : -1.0;
}

namespace {

/// Scans from ground-truth poses: `beams` beams evenly over `fov` centred
/// on the heading, Gaussian range noise, random dropouts; beams with no hit
/// within max_range are dropped (no return).
std::vector<Scan> SimulateScans(const std::vector<Segment>& walls,
                                const std::vector<Pose2>& poses, int beams, double fov,
                                double max_range, double sigma, double dropout, Rng& rng) {
  std::vector<Scan> scans(poses.size());
  const bool full_circle = fov >= 2.0 * kPi - 1e-12;
  for (size_t i = 0; i < poses.size(); ++i) {
    for (int k = 0; k < beams; ++k) {
      const double rel = full_circle ? 2.0 * kPi * k / beams
                                     : -0.5 * fov + (k + 0.5) * fov / beams;
      const double range = CastRay(walls, poses[i].Translation(), poses[i].theta + rel, max_range);
      const double noise = rng.Normal(sigma);
      const double drop = rng.Uniform();
      if (range <= 0.0 || drop < dropout) {
        continue;
      }
      const double r = range + noise;
      scans[i].points.push_back({r * std::cos(rel), r * std::sin(rel)});
    }
  }
  return scans;
}

/// Noisy odometry from ground truth; dataset poses integrate it from the
/// first ground-truth pose.
void MakeOdometry(Workload& wl, double sigma_t, double sigma_r, Rng& rng) {
  const auto& gt = wl.ground_truth;
  std::vector<Pose2> noisy;
  noisy.reserve(gt.size());
  for (size_t i = 0; i + 1 < gt.size(); ++i) {
    Pose2 rel = gt[i].RelativeTo(gt[i + 1]);
    rel.x += rng.Normal(sigma_t);
    rel.y += rng.Normal(sigma_t);
    rel.theta = WrapAngle(rel.theta + rng.Normal(sigma_r));
    noisy.push_back(rel);
  }
  wl.dataset_poses = {gt.front()};
  for (const Pose2& rel : noisy) {
    wl.dataset_poses.push_back(wl.dataset_poses.back().Compose(rel));
  }
  // RelativeOdometry(dataset poses), runner.cpp:52-59.
  wl.odometry.clear();
  for (size_t i = 0; i + 1 < wl.dataset_poses.size(); ++i) {
    wl.odometry.push_back(wl.dataset_poses[i].RelativeTo(wl.dataset_poses[i + 1]));
  }
}

void AddRect(std::vector<Segment>& walls, double x0, double y0, double x1, double y1) {
  walls.push_back({{x0, y0}, {x1, y0}});
  walls.push_back({{x1, y0}, {x1, y1}});
  walls.push_back({{x1, y1}, {x0, y1}});
  walls.push_back({{x0, y1}, {x0, y0}});
}

double DistToBox(const Vec2& p, double x0, double y0, double x1, double y1) {
  const double dx = std::max({x0 - p.x, 0.0, p.x - x1});
  const double dy = std::max({y0 - p.y, 0.0, p.y - y1});
  return std::sqrt(dx * dx + dy * dy);
}

struct RoomParams {
  int num_poses{76};
  int beams{360};
  int grid{100};
  int iterations{100};
};

Workload MakeRoom(const std::string& name, uint64_t seed, const RoomParams& prm) {
  Workload wl;
  wl.name = name;
  Rng rng(seed);
  const double side = 10.0;
  AddRect(wl.walls, 0.0, 0.0, side, side);
  // Smooth closed loop around the room centre.
  const Vec2 centre{5.0 + rng.Uniform(-0.2, 0.2), 5.0 + rng.Uniform(-0.2, 0.2)};
  const double radius = 3.0;
  const double phase = rng.Uniform(0.0, 2.0 * kPi);
  for (int i = 0; i < prm.num_poses; ++i) {
    const double t = 2.0 * kPi * i / prm.num_poses;
    const double r = radius * (1.0 + 0.08 * std::sin(3.0 * t + phase));
    const double dr = radius * 0.24 * std::cos(3.0 * t + phase);
    const Vec2 p{centre.x + r * std::cos(t), centre.y + r * std::sin(t)};
    const Vec2 d{dr * std::cos(t) - r * std::sin(t), dr * std::sin(t) + r * std::cos(t)};
    wl.ground_truth.push_back({p.x, p.y, std::atan2(d.y, d.x)});
  }
  // Three axis-aligned boxes with sides in [0.5, 1.5] m, clear of the path.
  std::vector<std::array<double, 4>> boxes;
  int attempts = 0;
  while (boxes.size() < 3 && attempts < 100000) {
    ++attempts;
    const double w = rng.Uniform(0.5, 1.5);
    const double h = rng.Uniform(0.5, 1.5);
    const double cx = rng.Uniform(0.3 + 0.5 * w, side - 0.3 - 0.5 * w);
    const double cy = rng.Uniform(0.3 + 0.5 * h, side - 0.3 - 0.5 * h);
    const double x0 = cx - 0.5 * w;
    const double x1 = cx + 0.5 * w;
    const double y0 = cy - 0.5 * h;
    const double y1 = cy + 0.5 * h;
    bool ok = true;
    for (int s = 0; s < 4 * prm.num_poses && ok; ++s) {
      const double t = 2.0 * kPi * s / (4.0 * prm.num_poses);
      const double r = radius * (1.0 + 0.08 * std::sin(3.0 * t + phase));
      const Vec2 p{centre.x + r * std::cos(t), centre.y + r * std::sin(t)};
      ok = DistToBox(p, x0, y0, x1, y1) > 0.6;
    }
    for (const auto& b : boxes) {
      if (!(x1 + 0.3 < b[0] || b[2] + 0.3 < x0 || y1 + 0.3 < b[1] || b[3] + 0.3 < y0)) {
        ok = false;
      }
    }
    if (ok) {
      boxes.push_back({x0, y0, x1, y1});
    }
  }
  for (const auto& b : boxes) {
    AddRect(wl.walls, b[0], b[1], b[2], b[3]);
  }
  wl.scans = SimulateScans(wl.walls, wl.ground_truth, prm.beams, 2.0 * kPi, 8.0, 0.01, 0.0, rng);
  MakeOdometry(wl, 0.005, 0.002, rng);
  wl.initial = wl.dataset_poses;
  // 100x100 nodes over the room plus a 0.15 m rim, so walls on the domain
  // edge keep map Jacobians.
  wl.map = GridMap(prm.grid, prm.grid, {-0.15, -0.15}, {side + 0.15, side + 0.15}, 0.0);
  wl.problem.hallucination = {4, 0.05, true};
  wl.problem.normals.method = NormalMethod::kPca;
  wl.problem.normals.k_neighbors = 10;
  wl.problem.smooth_gradient = true;
  wl.solver.backend = SolverBackend::kHip;
  wl.solver.max_iterations = prm.iterations;
  wl.solver.lambda_init = 1.0;
  wl.solver.lambda_factor = 1.0;
  return wl;
}

struct OfficeParams {
  double side{50.0};
  int blocks{3};
  int num_poses{910};
  int beams{180};
  double max_range{30.0};
  int grid{150};
};

/// Random cuts of [lo, hi] into pieces of length in [3, 8] m.
std::vector<double> RoomCuts(double lo, double hi, Rng& rng) {
  const double len = hi - lo;
  const int n_min = std::max(1, static_cast<int>(std::ceil(len / 8.0)));
  const int n_max = std::max(n_min, static_cast<int>(std::floor(len / 3.0)));
  const int n = n_min + static_cast<int>(rng.Uniform() * (n_max - n_min + 1));
  std::vector<double> cuts{lo};
  const double piece = len / n;
  for (int k = 1; k < n; ++k) {
    const double jitter = std::min(0.4 * (piece - 3.0), 1.0);
    cuts.push_back(lo + k * piece + rng.Uniform(-jitter, jitter));
  }
  cuts.push_back(hi);
  return cuts;
}

/// Wall segments along [a, b] at a fixed coordinate, minus door gaps.
void AddWallWithGaps(std::vector<Segment>& walls, bool horizontal, double fixed, double a,
                     double b, std::vector<std::pair<double, double>> gaps) {
  std::sort(gaps.begin(), gaps.end());
  double cur = a;
  auto emit = [&](double u, double v) {
    if (v - u < 1e-6) {
      return;
    }
    if (horizontal) {
      walls.push_back({{u, fixed}, {v, fixed}});
    } else {
      walls.push_back({{fixed, u}, {fixed, v}});
    }
  };
  for (const auto& g : gaps) {
    emit(cur, g.first);
    cur = g.second;
  }
  emit(cur, b);
}

Workload MakeOffice(const std::string& name, uint64_t seed, const OfficeParams& prm) {
  Workload wl;
  wl.name = name;
  Rng rng(seed);
  const double side = prm.side;
  const int nb = prm.blocks;
  AddRect(wl.walls, 0.0, 0.0, side, side);
  // Corridor centre lines: one along each outer wall, nb-1 inside.
  std::vector<double> cx(static_cast<size_t>(nb) + 1);
  std::vector<double> cy(static_cast<size_t>(nb) + 1);
  std::vector<double> wx(static_cast<size_t>(nb) + 1);
  std::vector<double> wy(static_cast<size_t>(nb) + 1);
  for (int i = 0; i <= nb; ++i) {
    wx[static_cast<size_t>(i)] = rng.Uniform(1.5, 2.0);
    wy[static_cast<size_t>(i)] = rng.Uniform(1.5, 2.0);
  }
  for (int i = 0; i <= nb; ++i) {
    const auto k = static_cast<size_t>(i);
    if (i == 0) {
      cx[k] = 0.5 * wx[k];
      cy[k] = 0.5 * wy[k];
    } else if (i == nb) {
      cx[k] = side - 0.5 * wx[k];
      cy[k] = side - 0.5 * wy[k];
    } else {
      cx[k] = side * i / nb + rng.Uniform(-1.0, 1.0);
      cy[k] = side * i / nb + rng.Uniform(-1.0, 1.0);
    }
  }
  // Blocks between corridors, divided into rooms of 3-8 m with door gaps
  // onto the corridors.
  for (int bi = 0; bi < nb; ++bi) {
    for (int bj = 0; bj < nb; ++bj) {
      const double x0 = cx[static_cast<size_t>(bi)] + 0.5 * wx[static_cast<size_t>(bi)];
      const double x1 = cx[static_cast<size_t>(bi) + 1] - 0.5 * wx[static_cast<size_t>(bi) + 1];
      const double y0 = cy[static_cast<size_t>(bj)] + 0.5 * wy[static_cast<size_t>(bj)];
      const double y1 = cy[static_cast<size_t>(bj) + 1] - 0.5 * wy[static_cast<size_t>(bj) + 1];
      const std::vector<double> xs = RoomCuts(x0, x1, rng);
      const std::vector<double> ys = RoomCuts(y0, y1, rng);
      const int rx = static_cast<int>(xs.size()) - 1;
      const int ry = static_cast<int>(ys.size()) - 1;
      std::vector<std::pair<double, double>> gaps_bottom, gaps_top, gaps_left, gaps_right;
      for (int i = 0; i < rx; ++i) {
        for (int j = 0; j < ry; ++j) {
          std::vector<int> sides;  // 0 bottom, 1 top, 2 left, 3 right
          if (j == 0) sides.push_back(0);
          if (j == ry - 1) sides.push_back(1);
          if (i == 0) sides.push_back(2);
          if (i == rx - 1) sides.push_back(3);
          if (sides.empty()) {
            continue;
          }
          const int pick = sides[static_cast<size_t>(rng.Uniform() * sides.size()) % sides.size()];
          const double width = rng.Uniform(0.8, 1.2);
          const bool along_x = pick <= 1;
          const double lo = along_x ? xs[static_cast<size_t>(i)] : ys[static_cast<size_t>(j)];
          const double hi = along_x ? xs[static_cast<size_t>(i) + 1] : ys[static_cast<size_t>(j) + 1];
          const double start = rng.Uniform(lo + 0.3, std::max(lo + 0.3, hi - 0.3 - width));
          const std::pair<double, double> gap{start, std::min(start + width, hi - 0.1)};
          (pick == 0 ? gaps_bottom : pick == 1 ? gaps_top : pick == 2 ? gaps_left : gaps_right)
              .push_back(gap);
        }
      }
      AddWallWithGaps(wl.walls, true, y0, x0, x1, gaps_bottom);
      AddWallWithGaps(wl.walls, true, y1, x0, x1, gaps_top);
      AddWallWithGaps(wl.walls, false, x0, y0, y1, gaps_left);
      AddWallWithGaps(wl.walls, false, x1, y0, y1, gaps_right);
      for (int i = 1; i < rx; ++i) {
        wl.walls.push_back({{xs[static_cast<size_t>(i)], y0}, {xs[static_cast<size_t>(i)], y1}});
      }
      for (int j = 1; j < ry; ++j) {
        wl.walls.push_back({{x0, ys[static_cast<size_t>(j)]}, {x1, ys[static_cast<size_t>(j)]}});
      }
    }
  }
  // Multi-lap route through corridor intersections: outer ring, a vertical
  // serpentine, a horizontal serpentine, then the outer ring again.
  std::vector<std::pair<int, int>> route{{0, 0}};
  auto go = [&route](int i, int j) {
    if (route.back() != std::make_pair(i, j)) {
      route.emplace_back(i, j);
    }
  };
  auto outer_lap = [&]() {
    go(nb, 0);
    go(nb, nb);
    go(0, nb);
    go(0, 0);
  };
  outer_lap();
  {
    int j = 0;
    for (int i = 1; i < nb; ++i) {
      go(i, j);
      j = nb - j;
      go(i, j);
    }
    go(nb, j);
    if (j == 0) go(nb, nb);
    go(0, nb);
    go(0, 0);
  }
  {
    int i = 0;
    for (int j = 1; j < nb; ++j) {
      go(i, j);
      i = nb - i;
      go(i, j);
    }
    go(i, nb);
    if (i == 0) go(nb, nb);
    go(nb, 0);
    go(0, 0);
  }
  outer_lap();
  std::vector<Vec2> pts;
  std::vector<double> arc{0.0};
  for (const auto& [i, j] : route) {
    pts.push_back({cx[static_cast<size_t>(i)], cy[static_cast<size_t>(j)]});
    if (pts.size() > 1) {
      arc.push_back(arc.back() + (pts.back() - pts[pts.size() - 2]).norm());
    }
  }
  const double total = arc.back();
  auto at = [&](double s) {
    s = std::clamp(s, 0.0, total);
    size_t k = static_cast<size_t>(std::upper_bound(arc.begin(), arc.end(), s) - arc.begin());
    k = std::clamp<size_t>(k, 1, arc.size() - 1);
    const double t = (s - arc[k - 1]) / std::max(arc[k] - arc[k - 1], 1e-12);
    return pts[k - 1] + (pts[k] - pts[k - 1]) * t;
  };
  const double ph1 = rng.Uniform(0.0, 2.0 * kPi);
  const double ph2 = rng.Uniform(0.0, 2.0 * kPi);
  for (int n = 0; n < prm.num_poses; ++n) {
    const double s = total * n / prm.num_poses;
    const Vec2 a = at(s - 0.6);
    const Vec2 b = at(s + 0.6);
    const double heading = std::atan2(b.y - a.y, b.x - a.x);
    const double off =
        0.25 * std::sin(2.0 * kPi * s / 37.0 + ph1) + 0.1 * std::sin(2.0 * kPi * s / 11.0 + ph2);
    const Vec2 p = at(s) + Vec2{-std::sin(heading), std::cos(heading)} * off;
    wl.ground_truth.push_back({p.x, p.y, heading});
  }
  wl.scans = SimulateScans(wl.walls, wl.ground_truth, prm.beams, kPi, prm.max_range, 0.02, 0.05,
                           rng);
  MakeOdometry(wl, 0.02, 0.01, rng);
  const double margin = 1.85 * side / 50.0;
  wl.map = GridMap(prm.grid, prm.grid, {-margin, -margin}, {side + margin, side + margin}, 0.0);
  wl.problem.normals.method = NormalMethod::kPca;
  wl.problem.normals.k_neighbors = 7;
  wl.problem.smooth_gradient = true;
  wl.solver.backend = SolverBackend::kHip;
  wl.solver.lambda_init = 5.0;
  wl.solver.lambda_factor = 0.9;
  return wl;
}

void WarmStart(Workload& wl, double sigma_t, double sigma_r, Rng& rng) {
  wl.initial = wl.ground_truth;
  for (size_t i = 1; i < wl.initial.size(); ++i) {
    wl.initial[i].x += rng.Normal(sigma_t);
    wl.initial[i].y += rng.Normal(sigma_t);
    wl.initial[i].theta = WrapAngle(wl.initial[i].theta + rng.Normal(sigma_r));
  }
}

std::vector<Scan> CornerScans(const std::vector<Pose2>& poses) {
  std::vector<Scan> scans;
  for (const Pose2& pose : poses) {
    Scan scan;
    for (int k = 1; k <= 40; ++k) {
      const double s = 0.25 * k;
      scan.points.push_back(pose.RotateInverse(Vec2{0.0, s} - pose.Translation()));
      scan.points.push_back(pose.RotateInverse(Vec2{s, 0.0} - pose.Translation()));
    }
    scans.push_back(scan);
  }
  return scans;
}

}  // namespace

Workload MakeWorkload(const std::string& name, uint64_t seed) {
  if (name == "room") {
    return MakeRoom(name, seed, RoomParams{});
  }
  if (name == "room_small") {
    return MakeRoom(name, seed, RoomParams{20, 90, 40, 10});
  }
  if (name == "office150" || name == "office200" || name == "office_small") {
    OfficeParams prm;
    if (name == "office_small") {
      prm.side = 20.0;
      prm.blocks = 1;
      prm.num_poses = 60;
      prm.beams = 90;
      prm.max_range = 15.0;
      prm.grid = 60;
    } else if (name == "office200") {
      prm.grid = 200;
    }
    Workload wl = MakeOffice(name, seed, prm);
    Rng warm(seed ^ 0x5EEDF00DULL);
    if (name == "office150") {
      wl.problem.hallucination = {6, 0.1, true};
      wl.problem.weights.eikonal = 0.018;
      wl.solver.step_tolerance = 1e-3;
      wl.solver.max_iterations = 5;
      wl.incremental = true;
      wl.iterations_per_increment = 5;
      wl.initial = wl.dataset_poses;
    } else {
      wl.problem.hallucination = {4, 0.05, true};
      wl.problem.weights.eikonal = 0.01;
      wl.solver.step_tolerance = 0.0;
      wl.solver.max_iterations = name == "office200" ? 200 : 10;
      WarmStart(wl, 0.05, 0.01, warm);
    }
    return wl;
  }
  if (name.rfind("batch", 0) == 0 || name.rfind("served", 0) == 0) {
    // Uneven instances of the batched and served forms: the seed picks the
    // grid (48..160 nodes a side, cells 0.1..0.36 m) for "batch", or
    // (100..150 nodes, 20..50 m floorplans) for "served", the pose count,
    // and 180 or 360 beams; the suffix number offsets the seed.
    const bool batch = name[0] == 'b';
    const std::string digits = name.substr(batch ? 5 : 6);
    const uint64_t inst = digits.empty() ? 0 : std::stoull(digits);
    Rng pick(seed * 0x9E3779B97F4A7C15ULL + inst);
    OfficeParams prm;
    if (batch) {
      prm.grid = 48 + static_cast<int>(pick.Uniform() * 113.0);
      const double cell = pick.Uniform(0.1, 0.36);
      prm.side = cell * (prm.grid - 1) / 1.074;
      prm.num_poses = 20 + static_cast<int>(pick.Uniform() * 281.0);
      prm.beams = pick.Uniform() < 0.5 ? 180 : 360;
    } else {
      prm.grid = 100 + static_cast<int>(pick.Uniform() * 51.0);
      prm.side = pick.Uniform(20.0, 50.0);
      prm.num_poses = 400;
      prm.beams = 180;
    }
    prm.blocks = std::max(1, static_cast<int>(std::lround(prm.side / 16.0)));
    prm.max_range = std::min(30.0, prm.side);
    Workload wl = MakeOffice(name, seed + inst, prm);
    Rng warm((seed + inst) ^ 0x5EEDF00DULL);
    if (batch) {
      wl.problem.hallucination = {4, 0.05, true};
      wl.problem.weights.eikonal = 0.01;
      wl.solver.step_tolerance = 0.0;
      wl.solver.max_iterations = 30;
      WarmStart(wl, 0.05, 0.01, warm);
    } else {
      wl.problem.hallucination = {6, 0.1, true};
      wl.problem.weights.eikonal = 0.018;
      wl.solver.step_tolerance = 1e-3;
      wl.solver.max_iterations = 5;
      wl.incremental = true;
      wl.iterations_per_increment = 5;
      wl.initial = wl.dataset_poses;
    }
    return wl;
  }
  if (name == "corner") {
    // tests/test_e2e.cpp:15-61.
    Workload wl;
    wl.name = name;
    const Pose2 pose_0{4.0, 4.0, 0.2};
    const Pose2 gt_1{5.0, 5.0, -0.1};
    wl.ground_truth = {pose_0, gt_1};
    wl.scans = CornerScans(wl.ground_truth);
    wl.initial = {pose_0, {5.4, 4.7, 0.05}};
    wl.dataset_poses = wl.initial;
    wl.odometry = {pose_0.RelativeTo(gt_1)};
    wl.map = GridMap(40, 40, {-2.0, -2.0}, {12.0, 12.0}, 0.0);
    wl.problem.hallucination.points_per_scan_point = 4;
    wl.problem.hallucination.step_size = 0.2;
    wl.problem.normals.k_neighbors = 8;
    wl.problem.normals.method = NormalMethod::kPca;
    wl.solver.backend = SolverBackend::kHip;
    wl.solver.max_iterations = 40;
    return wl;
  }
  throw std::invalid_argument("unknown workload '" + name + "'");
}

}  // namespace sdf_slam
