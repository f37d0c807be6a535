// Seeded synthetic SLAM workloads: worlds of wall segments, trajectories,
// ray-cast scans with range noise and dropouts, noisy odometry, and the
// problem/solver options of each workload. A portable RNG (xoshiro256**
// seeded by splitmix64, Box-Muller normals) makes every workload identical
// on every platform.
//
// The dataset conventions follow the reference's runner: the dataset poses
// are the odometry-integrated poses, odometry residual measurements are
// RelativeOdometry(dataset poses) (src/pipeline/runner.cpp:52-59), and batch
// mode starts from the dataset poses unless an initial trajectory is given
// (runner.cpp:211-218).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "problem.hpp"
#include "solver.hpp"

namespace sdf_slam {

class Rng {
 public:
  explicit Rng(uint64_t seed);
  uint64_t Next();
  double Uniform();                      // [0, 1)
  double Uniform(double lo, double hi);  // [lo, hi)
  double Normal();                       // N(0, 1)
  double Normal(double sigma) { return sigma * Normal(); }

 private:
  uint64_t s_[4]{};
  bool has_spare_{false};
  double spare_{0.0};
};

struct Segment {
  Vec2 a;
  Vec2 b;
};

struct Workload {
  std::string name;
  GridMap map;
  std::vector<Scan> scans;
  std::vector<Pose2> ground_truth;
  std::vector<Pose2> dataset_poses;
  std::vector<Pose2> odometry;
  std::vector<Pose2> initial;
  std::vector<RelationMeasurement> relations;
  ProblemOptions problem;
  SolverOptions solver;
  bool incremental{false};
  int iterations_per_increment{5};
  std::vector<Segment> walls;
};

/// Workloads by name:
///   "room"        10x10 m room, 3 boxes, 76 poses, 360 beams/360 deg, 8 m,
///                 100x100 grid, 4 hallucination samples @0.05, 100 iterations
///   "office150"   50x50 m office, 910 poses, 180 beams/180 deg, 30 m, 5 %
///                 dropouts, 150x150 grid, incremental, 5 iterations/increment
///   "office200"   the same office and trajectory, 200x200 grid, 4
///                 hallucination samples @0.05, warm start from ground truth
///                 + N(0, 0.05 m)/N(0, 0.01 rad), 200 iterations
///   "room_small", "office_small"  reduced versions for quick parity checks
///   "corner"      two corner scans with a perturbed second frame
///                 (the reference's e2e setup, tests/test_e2e.cpp:15-61)
///   "batch<k>"    uneven batch instance k: 48..160 nodes a side at
///                 0.1..0.36 m, 20..300 poses, 180 or 360 beams, warm start,
///                 30 iterations
///   "served<k>"   incremental session k: 20..50 m office, 100..150 nodes a
///                 side, 400 frames, 5 iterations per increment
Workload MakeWorkload(const std::string& name, uint64_t seed);

/// Ray-casts one beam against the walls; returns the hit distance or a
/// negative value when nothing is hit within max_range.
double CastRay(const std::vector<Segment>& walls, const Vec2& origin, double angle,
               double max_range);

}  // namespace sdf_slam
