// Problem builder (host setup that produces the specs the device path
// consumes). See problem.hpp for the reference files each part follows.
#include "problem.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>

#include "knn.hpp"

namespace sdf_slam {

std::vector<HallucinatedPoint> GenerateHallucinatedPoints(const Scan& scan,
                                                          const HallucinationOptions& options) {
  // hallucination.cpp:5-28: per_side samples per direction, the free-space
  // sample (+j*step toward the sensor) before the occluded one.
  std::vector<HallucinatedPoint> result;
  if (options.points_per_scan_point <= 0) {
    return result;
  }
  const int per_side = options.points_per_scan_point / (options.both_directions ? 2 : 1);
  result.reserve(scan.points.size() * static_cast<size_t>(options.points_per_scan_point));
  for (const Vec2& point : scan.points) {
    const double norm = point.norm();
    if (norm < 1e-9) {
      continue;
    }
    const Vec2 toward_origin = -point / norm * options.step_size;
    for (int j = 1; j <= per_side; ++j) {
      result.push_back({point + toward_origin * j, options.step_size * j});
      if (options.both_directions) {
        result.push_back({point - toward_origin * j, -options.step_size * j});
      }
    }
  }
  return result;
}

namespace {

/// PCA of the k nearest neighbours of points[i]: the eigenvector of the
/// smaller covariance eigenvalue, oriented toward `origin`, and the
/// eigenvalue ratio (normals.cpp:70-103). The 2x2 symmetric eigenproblem is
/// solved in closed form.
void PcaNormal(const std::vector<Vec2>& points, const PointGrid& grid, size_t i, int k,
               const Vec2& origin, Vec2& normal_out, double& cornerness_out,
               std::vector<int>& idx, std::vector<double>& d2) {
  const int found = grid.Knn(points[i], k, idx.data(), d2.data());
  Vec2 mean{0.0, 0.0};
  for (int j = 0; j < found; ++j) {
    mean = mean + points[static_cast<size_t>(idx[static_cast<size_t>(j)])];
  }
  mean = mean / static_cast<double>(found);
  double a = 0.0;
  double b = 0.0;
  double c = 0.0;
  for (int j = 0; j < found; ++j) {
    const Vec2 d = points[static_cast<size_t>(idx[static_cast<size_t>(j)])] - mean;
    a += d.x * d.x;
    b += d.x * d.y;
    c += d.y * d.y;
  }
  const double half_tr = 0.5 * (a + c);
  const double half_diff = 0.5 * (a - c);
  const double rad = std::sqrt(half_diff * half_diff + b * b);
  const double l_min = half_tr - rad;
  const double l_max = half_tr + rad;
  Vec2 normal;
  const Vec2 v1{b, l_min - a};
  const Vec2 v2{l_min - c, b};
  const double n1 = v1.norm();
  const double n2 = v2.norm();
  if (n1 < 1e-300 && n2 < 1e-300) {
    normal = a <= c ? Vec2{1.0, 0.0} : Vec2{0.0, 1.0};
  } else if (n1 >= n2) {
    normal = v1 / n1;
  } else {
    normal = v2 / n2;
  }
  if (normal.dot(origin - points[i]) < 0.0) {
    normal = -normal;
  }
  normal_out = normal;
  cornerness_out = std::max(l_min, 0.0) / std::max(l_max, 1e-12);
}

}  // namespace

ScanNormals ComputeScanNormals(const std::vector<Scan>& scans, const std::vector<Pose2>& poses,
                               int k_neighbors, bool per_scan) {
  if (scans.size() != poses.size()) {
    throw std::invalid_argument("scan count != pose count");
  }
  const int k = std::max(2, k_neighbors);
  ScanNormals result;
  std::vector<int> idx(static_cast<size_t>(k));
  std::vector<double> d2(static_cast<size_t>(k));
  if (per_scan) {
    // normals.cpp:104-124: neighbourhoods inside one scan, sensor frame.
    for (size_t s = 0; s < scans.size(); ++s) {
      const std::vector<Vec2>& local = scans[s].points;
      if (local.empty()) {
        continue;
      }
      const PointGrid grid(local);
      const int kk = std::min<int>(k, static_cast<int>(local.size()));
      for (size_t i = 0; i < local.size(); ++i) {
        Vec2 normal;
        double corner = 0.0;
        PcaNormal(local, grid, i, kk, Vec2{0.0, 0.0}, normal, corner, idx, d2);
        result.points.push_back(poses[s].Apply(local[i]));
        result.normals.push_back(poses[s].Rotate(normal));
        result.cornerness.push_back(corner);
      }
    }
    return result;
  }
  // normals.cpp:128-160: one neighbourhood structure over all scans.
  std::vector<Vec2> points;
  std::vector<Vec2> origins;
  for (size_t i = 0; i < scans.size(); ++i) {
    for (const Vec2& p : scans[i].points) {
      points.push_back(poses[i].Apply(p));
      origins.push_back(poses[i].Translation());
    }
  }
  return ScanNormalsOf(std::move(points), origins, k_neighbors);
}

ScanNormals ScanNormalsOf(std::vector<Vec2> points, const std::vector<Vec2>& origins, int k_neighbors) {
  if (points.size() != origins.size()) {
    throw std::invalid_argument("point count != origin count");
  }
  const int k = std::max(2, k_neighbors);
  ScanNormals result;
  result.points = std::move(points);
  const size_t n = result.points.size();
  result.normals.resize(n);
  result.cornerness.resize(n);
  if (n == 0) {
    return result;
  }
  std::vector<int> idx(static_cast<size_t>(k));
  std::vector<double> d2(static_cast<size_t>(k));
  const PointGrid grid(result.points);
  const int kk = static_cast<int>(std::min<size_t>(static_cast<size_t>(k), n));
  for (size_t i = 0; i < n; ++i) {
    PcaNormal(result.points, grid, i, kk, origins[i], result.normals[i], result.cornerness[i], idx,
              d2);
  }
  return result;
}

std::vector<NodeNormal> AssignNodeNormals(const GridMap& map, const ScanNormals& scan_normals,
                                          const NormalOptions& options) {
  std::vector<NodeNormal> node_normals(static_cast<size_t>(map.num_nodes()));
  if (scan_normals.points.empty()) {
    return node_normals;
  }
  const PointGrid grid(scan_normals.points);
  int nearest = 0;
  double dist_sq = 0.0;
  for (int h = 0; h < map.ny(); ++h) {
    for (int w = 0; w < map.nx(); ++w) {
      const Vec2 node = map.NodePosition(w, h);
      grid.Knn(node, 1, &nearest, &dist_sq);
      const auto ni = static_cast<size_t>(nearest);
      const Vec2 n_pca = scan_normals.normals[ni];
      const double corner =
          std::clamp(scan_normals.cornerness[ni] / options.corner_threshold, 0.0, 1.0);
      NodeNormal out;
      switch (options.method) {
        case NormalMethod::kPca:
          out.normal = n_pca;
          break;
        case NormalMethod::kHybrid: {
          // normals.cpp:186-205: blend with the sign-matched closest-point cue.
          const Vec2 to_surface = scan_normals.points[ni] - node;
          const double norm = to_surface.norm();
          if (norm < 1e-9) {
            out.normal = n_pca;
            break;
          }
          Vec2 cue = to_surface / norm;
          if (cue.dot(n_pca) < 0.0) {
            cue = -cue;
          }
          const Vec2 blended = n_pca * (1.0 - corner) + cue * corner;
          const double bn = blended.norm();
          out.normal = bn > 1e-9 ? blended / bn : n_pca;
          break;
        }
        case NormalMethod::kWeighted:
          out.normal = n_pca;
          out.eikonal_scale = 1.0 - corner;
          break;
      }
      node_normals[static_cast<size_t>(map.NodeId(w, h))] = out;
    }
  }
  return node_normals;
}

Problem::Problem(GridMap map, std::vector<Pose2> poses, const std::vector<Scan>& scans,
                 const std::vector<Pose2>& odometry, const ProblemOptions& options,
                 const std::vector<RelationMeasurement>& relations, const NodeNormalsFn& node_normals_fn)
    : map_(std::move(map)),
      poses_(std::move(poses)),
      huber_delta_(options.huber_delta),
      smooth_gradient_(options.smooth_gradient),
      smooth_gradient_step_(options.smooth_gradient_step) {
  if (scans.size() != poses_.size()) {
    throw std::invalid_argument("scan count != pose count");
  }
  if (!odometry.empty() && odometry.size() != scans.size() - 1) {
    throw std::invalid_argument("odometry count must be frames - 1");
  }
  // Point rows (problem.cpp:26-37).
  const double sqrt_ws = std::sqrt(options.weights.scan);
  const double sqrt_wh = std::sqrt(options.weights.hallucination);
  for (size_t i = 0; i < scans.size(); ++i) {
    const int frame = static_cast<int>(i);
    for (const Vec2& p : scans[i].points) {
      point_specs_.push_back({frame, p, 0.0, sqrt_ws});
    }
    for (const HallucinatedPoint& hall : GenerateHallucinatedPoints(scans[i], options.hallucination)) {
      point_specs_.push_back({frame, hall.point_sensor, hall.expected_sdf, sqrt_wh});
    }
  }
  // Active nodes (problem.cpp:39-70).
  const int num_nodes = map_.num_nodes();
  node_to_active_.assign(static_cast<size_t>(num_nodes), -1);
  std::vector<char> active(static_cast<size_t>(num_nodes), 0);
  if (options.active_region) {
    const int margin = options.active_margin;
    for (const PointSpec& spec : point_specs_) {
      const Vec2 global = poses_[static_cast<size_t>(spec.frame)].Apply(spec.point_sensor);
      GridMap::CellRef cell;
      if (!map_.Locate(global, cell)) {
        continue;
      }
      for (int h = std::max(0, cell.h - margin); h <= std::min(map_.ny() - 1, cell.h + 1 + margin);
           ++h) {
        for (int w = std::max(0, cell.w - margin);
             w <= std::min(map_.nx() - 1, cell.w + 1 + margin); ++w) {
          active[static_cast<size_t>(map_.NodeId(w, h))] = 1;
        }
      }
    }
  } else {
    std::fill(active.begin(), active.end(), 1);
  }
  for (int node = 0; node < num_nodes; ++node) {
    if (active[static_cast<size_t>(node)] != 0) {
      node_to_active_[static_cast<size_t>(node)] = static_cast<int>(active_to_node_.size());
      active_to_node_.push_back(node);
    }
  }
  // Eikonal rows (problem.cpp:72-94).
  std::vector<NodeNormal> node_normals;
  if (node_normals_fn && node_normals_fn.target<DeferredNodeNormals>() != nullptr) {
    if (options.normals.per_scan || options.normals.method == NormalMethod::kWeighted) {
      throw std::invalid_argument("deferred node normals need the all-scans pca or hybrid method");
    }
    node_normals = DeferredNodeNormals{}({}, {}, map_, options.normals);
  } else if (node_normals_fn && !options.normals.per_scan) {
    // The all-scans neighbourhood's inputs (ComputeScanNormals' order).
    std::vector<Vec2> points;
    std::vector<Vec2> origins;
    for (size_t i = 0; i < scans.size(); ++i) {
      for (const Vec2& p : scans[i].points) {
        points.push_back(poses_[i].Apply(p));
        origins.push_back(poses_[i].Translation());
      }
    }
    node_normals = node_normals_fn(points, origins, map_, options.normals);
  } else {
    const ScanNormals scan_normals =
        ComputeScanNormals(scans, poses_, options.normals.k_neighbors, options.normals.per_scan);
    node_normals = AssignNodeNormals(map_, scan_normals, options.normals);
  }
  const double sqrt_we = std::sqrt(options.weights.eikonal);
  for (int h = 0; h + 1 < map_.ny(); ++h) {
    for (int w = 0; w + 1 < map_.nx(); ++w) {
      const int node = map_.NodeId(w, h);
      if (node_to_active_[static_cast<size_t>(node)] < 0 ||
          node_to_active_[static_cast<size_t>(map_.NodeId(w + 1, h))] < 0 ||
          node_to_active_[static_cast<size_t>(map_.NodeId(w, h + 1))] < 0) {
        continue;
      }
      const NodeNormal& nn = node_normals[static_cast<size_t>(node)];
      const double scale = nn.eikonal_scale;
      if (scale <= 0.0) {
        continue;
      }
      eikonal_specs_.push_back({w, h, nn.normal, sqrt_we * std::sqrt(scale)});
    }
  }
  // Odometry (problem.cpp:96-100) and relations (problem.cpp:102-114).
  const double sqrt_wo = std::sqrt(options.weights.odometry);
  for (size_t i = 0; i + 1 < scans.size() && i < odometry.size(); ++i) {
    odom_specs_.push_back({static_cast<int>(i), static_cast<int>(i) + 1, odometry[i], sqrt_wo});
  }
  const double sqrt_wr = std::sqrt(options.weights.relation);
  const int frames = static_cast<int>(scans.size());
  for (const RelationMeasurement& rel : relations) {
    if (rel.frame_i < 0 || rel.frame_j <= rel.frame_i) {
      throw std::invalid_argument("relation frames must satisfy 0 <= i < j");
    }
    if (rel.frame_j < frames) {
      odom_specs_.push_back({rel.frame_i, rel.frame_j, rel.measurement, sqrt_wr * rel.sqrt_weight});
    }
  }
  BuildSoa();
}

std::vector<NodeNormal> DeferredNodeNormals::operator()(const std::vector<Vec2>&, const std::vector<Vec2>&,
                                                        const GridMap& map, const NormalOptions&) const {
  return std::vector<NodeNormal>(static_cast<size_t>(map.num_nodes()), NodeNormal{Vec2{0.0, 0.0}, 1.0});
}

void Problem::BuildSoa() {
  Soa& s = soa_;
  const size_t np = point_specs_.size();
  s.point_frame.resize(np);
  s.point_x.resize(np);
  s.point_y.resize(np);
  s.point_delta.resize(np);
  s.point_sw.resize(np);
  for (size_t i = 0; i < np; ++i) {
    const PointSpec& p = point_specs_[i];
    s.point_frame[i] = p.frame;
    s.point_x[i] = p.point_sensor.x;
    s.point_y[i] = p.point_sensor.y;
    s.point_delta[i] = p.expected_sdf;
    s.point_sw[i] = p.sqrt_weight;
  }
  const size_t ne = eikonal_specs_.size();
  s.eik_w.resize(ne);
  s.eik_h.resize(ne);
  s.eik_nx.resize(ne);
  s.eik_ny.resize(ne);
  s.eik_sw.resize(ne);
  for (size_t i = 0; i < ne; ++i) {
    const EikonalSpec& e = eikonal_specs_[i];
    s.eik_w[i] = e.w;
    s.eik_h[i] = e.h;
    s.eik_nx[i] = e.normal.x;
    s.eik_ny[i] = e.normal.y;
    s.eik_sw[i] = e.sqrt_weight;
  }
  const size_t no = odom_specs_.size();
  s.odom_i.resize(no);
  s.odom_j.resize(no);
  s.odom_dx.resize(no);
  s.odom_dy.resize(no);
  s.odom_dth.resize(no);
  s.odom_sw.resize(no);
  for (size_t i = 0; i < no; ++i) {
    const OdomSpec& o = odom_specs_[i];
    s.odom_i[i] = o.frame_i;
    s.odom_j[i] = o.frame_j;
    s.odom_dx[i] = o.measurement.x;
    s.odom_dy[i] = o.measurement.y;
    s.odom_dth[i] = o.measurement.theta;
    s.odom_sw[i] = o.sqrt_weight;
  }
}

sdfslam_problem_view Problem::View() const {
  sdfslam_problem_view v{};
  v.grid.nx = map_.nx();
  v.grid.ny = map_.ny();
  v.grid.min_x = map_.min_corner().x;
  v.grid.min_y = map_.min_corner().y;
  v.grid.dx = map_.dx();
  v.grid.dy = map_.dy();
  v.num_poses = static_cast<int32_t>(poses_.size());
  v.num_active_nodes = num_active_nodes();
  v.node_to_active = node_to_active_.data();
  v.num_points = static_cast<int64_t>(point_specs_.size());
  v.point_frame = soa_.point_frame.data();
  v.point_x = soa_.point_x.data();
  v.point_y = soa_.point_y.data();
  v.point_expected_sdf = soa_.point_delta.data();
  v.point_sqrt_weight = soa_.point_sw.data();
  v.num_eikonal = static_cast<int64_t>(eikonal_specs_.size());
  v.eik_w = soa_.eik_w.data();
  v.eik_h = soa_.eik_h.data();
  v.eik_nx = soa_.eik_nx.data();
  v.eik_ny = soa_.eik_ny.data();
  v.eik_sqrt_weight = soa_.eik_sw.data();
  v.num_odom = static_cast<int64_t>(odom_specs_.size());
  v.odom_i = soa_.odom_i.data();
  v.odom_j = soa_.odom_j.data();
  v.odom_dx = soa_.odom_dx.data();
  v.odom_dy = soa_.odom_dy.data();
  v.odom_dtheta = soa_.odom_dth.data();
  v.odom_sqrt_weight = soa_.odom_sw.data();
  v.huber_delta = huber_delta_;
  v.smooth_gradient = smooth_gradient_ ? 1 : 0;
  v.smooth_gradient_step = smooth_gradient_step_;
  return v;
}

void Problem::GetState(std::vector<double>& map_values, std::vector<double>& poses) const {
  map_values = map_.values();
  poses.resize(3 * poses_.size());
  for (size_t i = 0; i < poses_.size(); ++i) {
    poses[3 * i] = poses_[i].x;
    poses[3 * i + 1] = poses_[i].y;
    poses[3 * i + 2] = poses_[i].theta;
  }
}

void Problem::SetState(const double* map_values, const double* poses) {
  if (map_values != nullptr) {
    std::copy(map_values, map_values + map_.values().size(), map_.values().begin());
  }
  if (poses != nullptr) {
    for (size_t i = 0; i < poses_.size(); ++i) {
      poses_[i] = {poses[3 * i], poses[3 * i + 1], poses[3 * i + 2]};
    }
  }
}

}  // namespace sdf_slam
