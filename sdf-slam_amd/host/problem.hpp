// Host mirror of the reference's Problem: residual specs, state layout and
// the builder that derives hallucinated points, normals, the active region,
// eikonal and odometry/relation specs from scans and poses.
//   types            - src/problem/problem.hpp:14-87
//   Problem          - src/problem/problem.hpp:92-161, src/problem/problem.cpp:11-115
//   hallucination    - src/core/hallucination.hpp:9-29, hallucination.cpp:5-28
//   normals          - src/core/normals.hpp:13-69, normals.cpp:70-220
// Residual evaluation, ApplyStep and Cost are hot-path work and run on the
// device through the C ABI (include/sdfslam_hip.h).
#pragma once

#include <cstdint>
#include <functional>
#include <vector>

#include "geometry.hpp"
#include "sdfslam_hip.h"

namespace sdf_slam {

/// A LiDAR scan; points in the sensor frame (src/core/scan.hpp:12-14).
struct Scan {
  std::vector<Vec2> points;
};

struct Weights {
  double scan{1.0};
  double hallucination{1.0};
  double eikonal{1.0};
  double odometry{1.0};
  double relation{1.0};
};

struct PointSpec {
  int frame{0};
  Vec2 point_sensor{};
  double expected_sdf{0.0};
  double sqrt_weight{1.0};
};

struct EikonalSpec {
  int w{0};
  int h{0};
  Vec2 normal{};
  double sqrt_weight{1.0};
};

struct OdomSpec {
  int frame_i{0};
  int frame_j{0};
  Pose2 measurement;
  double sqrt_weight{1.0};
};

struct RelationMeasurement {
  int frame_i{0};
  int frame_j{0};
  Pose2 measurement;
  double sqrt_weight{1.0};
};

struct HallucinationOptions {
  int points_per_scan_point{6};
  double step_size{0.1};
  bool both_directions{true};
};

struct HallucinatedPoint {
  Vec2 point_sensor;
  double expected_sdf{0.0};
};

enum class NormalMethod : std::uint8_t { kPca, kHybrid, kWeighted };

struct NormalOptions {
  NormalMethod method{NormalMethod::kWeighted};
  int k_neighbors{10};
  bool per_scan{false};
  double corner_threshold{0.2};
};

struct ProblemOptions {
  Weights weights;
  HallucinationOptions hallucination;
  NormalOptions normals;
  double huber_delta{0.0};
  bool smooth_gradient{true};
  double smooth_gradient_step{0.0};
  bool active_region{false};
  int active_margin{2};
};

struct ScanNormals {
  std::vector<Vec2> points;
  std::vector<Vec2> normals;
  std::vector<double> cornerness;
};

struct NodeNormal {
  Vec2 normal{};
  double eikonal_scale{1.0};
};

/// Beam samples at +-j*step around each scan point, alternating free-space
/// (+) and occluded (-) sides (hallucination.cpp:5-28).
std::vector<HallucinatedPoint> GenerateHallucinatedPoints(const Scan& scan,
                                                          const HallucinationOptions& options);

/// kNN-PCA normals oriented toward the scan origin (normals.cpp:70-160).
ScanNormals ComputeScanNormals(const std::vector<Scan>& scans, const std::vector<Pose2>& poses,
                               int k_neighbors, bool per_scan = false);

/// The all-scans branch of ComputeScanNormals on points already in world
/// coordinates, each with its sensor origin.
ScanNormals ScanNormalsOf(std::vector<Vec2> points, const std::vector<Vec2>& origins, int k_neighbors);

/// Nearest-scan-point normal per grid node (normals.cpp:162-220).
std::vector<NodeNormal> AssignNodeNormals(const GridMap& map, const ScanNormals& scan_normals,
                                          const NormalOptions& options);

/// Node normals computed elsewhere (the device backend): from every scan
/// point in world coordinates and its sensor origin (frame by frame), the
/// values AssignNodeNormals(ComputeScanNormals(...)) gives with the
/// all-scans neighbourhood.
using NodeNormalsFn = std::function<std::vector<NodeNormal>(
    const std::vector<Vec2>& points, const std::vector<Vec2>& origins, const GridMap& map,
    const NormalOptions& options)>;

/// A NodeNormalsFn marker: the node normals are computed later, on the
/// device, for a context update (sdfslam_hip_context_update_device_normals).
/// The problem gets a zero normal and eikonal scale 1 at every node, which
/// for the pca and hybrid methods (scale 1 everywhere) gives the specs the
/// real normals give, with placeholder normal values.
struct DeferredNodeNormals {
  std::vector<NodeNormal> operator()(const std::vector<Vec2>& points, const std::vector<Vec2>& origins,
                                     const GridMap& map, const NormalOptions& options) const;
};

class Problem {
 public:
  /// Builds the specs exactly as problem.cpp:11-115: point rows per frame
  /// (scan points, then hallucinated points), the active node set, eikonal
  /// rows h-major, consecutive odometry, then relations with frame_j below
  /// the frame count.
  /// With `node_normals` (and the all-scans neighbourhood) the node normals
  /// come from that function instead of the host computation.
  Problem(GridMap map, std::vector<Pose2> poses, const std::vector<Scan>& scans,
          const std::vector<Pose2>& odometry, const ProblemOptions& options,
          const std::vector<RelationMeasurement>& relations = {},
          const NodeNormalsFn& node_normals = nullptr);

  [[nodiscard]] const GridMap& map() const { return map_; }
  GridMap& map() { return map_; }
  [[nodiscard]] const std::vector<Pose2>& poses() const { return poses_; }
  std::vector<Pose2>& poses() { return poses_; }
  [[nodiscard]] const std::vector<PointSpec>& point_specs() const { return point_specs_; }
  [[nodiscard]] const std::vector<EikonalSpec>& eikonal_specs() const { return eikonal_specs_; }
  [[nodiscard]] const std::vector<OdomSpec>& odom_specs() const { return odom_specs_; }
  [[nodiscard]] double huber_delta() const { return huber_delta_; }
  [[nodiscard]] bool smooth_gradient() const { return smooth_gradient_; }
  [[nodiscard]] double smooth_gradient_step() const { return smooth_gradient_step_; }

  [[nodiscard]] int num_active_nodes() const { return static_cast<int>(active_to_node_.size()); }
  [[nodiscard]] int num_parameters() const {
    return num_active_nodes() + 3 * (static_cast<int>(poses_.size()) - 1);
  }
  [[nodiscard]] int num_residuals() const {
    return static_cast<int>(point_specs_.size() + eikonal_specs_.size() + 3 * odom_specs_.size());
  }
  [[nodiscard]] int NodeColumn(int node_id) const {
    return node_to_active_[static_cast<size_t>(node_id)];
  }
  [[nodiscard]] int PoseColumn(int frame) const { return num_active_nodes() + 3 * (frame - 1); }
  [[nodiscard]] int ActiveNode(int active_index) const {
    return active_to_node_[static_cast<size_t>(active_index)];
  }
  [[nodiscard]] const std::vector<int>& node_to_active() const { return node_to_active_; }

  /// Structure-of-arrays copy of the specs behind a sdfslam_problem_view.
  /// The view's pointers stay valid while this Problem lives and its specs
  /// are unchanged.
  [[nodiscard]] sdfslam_problem_view View() const;
  /// Flat state arrays: nx*ny map values and 3*N pose components.
  void GetState(std::vector<double>& map_values, std::vector<double>& poses) const;
  void SetState(const double* map_values, const double* poses);

 private:
  void BuildSoa();

  GridMap map_;
  std::vector<Pose2> poses_;
  double huber_delta_{0.0};
  bool smooth_gradient_{false};
  double smooth_gradient_step_{0.0};
  std::vector<PointSpec> point_specs_;
  std::vector<EikonalSpec> eikonal_specs_;
  std::vector<OdomSpec> odom_specs_;
  std::vector<int> node_to_active_;
  std::vector<int> active_to_node_;

  struct Soa {
    std::vector<int32_t> point_frame;
    std::vector<double> point_x, point_y, point_delta, point_sw;
    std::vector<int32_t> eik_w, eik_h;
    std::vector<double> eik_nx, eik_ny, eik_sw;
    std::vector<int32_t> odom_i, odom_j;
    std::vector<double> odom_dx, odom_dy, odom_dth, odom_sw;
  } soa_;
};

}  // namespace sdf_slam
