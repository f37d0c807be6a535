// Host-side geometry of the problem builder: 2D vectors, rigid poses and the
// SDF grid. Mirrors the reference's value types so the builder reads like
// the reference:
//   Pose2 / WrapAngle  - src/core/pose.hpp:9-44
//   GridMap / CellRef  - src/core/grid_map.hpp:11-123
// Only what problem construction needs lives here; residual evaluation runs
// on the device (csrc/).
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <stdexcept>
#include <vector>

namespace sdf_slam {

struct Vec2 {
  double x{0.0};
  double y{0.0};
  Vec2() = default;
  Vec2(double x_, double y_) : x(x_), y(y_) {}
  Vec2 operator+(const Vec2& o) const { return {x + o.x, y + o.y}; }
  Vec2 operator-(const Vec2& o) const { return {x - o.x, y - o.y}; }
  Vec2 operator-() const { return {-x, -y}; }
  Vec2 operator*(double s) const { return {x * s, y * s}; }
  Vec2 operator/(double s) const { return {x / s, y / s}; }
  [[nodiscard]] double dot(const Vec2& o) const { return x * o.x + y * o.y; }
  [[nodiscard]] double norm() const { return std::sqrt(x * x + y * y); }
  [[nodiscard]] double squaredNorm() const { return x * x + y * y; }
};

/// cos and sin of an angle from one ::sincos call. The compiler turns a
/// std::cos and a std::sin of the same angle into that call where it can, and
/// ::sincos can differ from the separate functions in the last bit, so every
/// rotation calls it explicitly: the device, fed these values, places points
/// bit for bit where the host does.
inline void CosSin(double angle, double& c, double& s) { ::sincos(angle, &s, &c); }

/// atan2(sin a, cos a), pose.hpp:9.
inline double WrapAngle(double angle) { return std::atan2(std::sin(angle), std::cos(angle)); }

/// Rigid 2D pose (pose.hpp:12-44).
struct Pose2 {
  double x{0.0};
  double y{0.0};
  double theta{0.0};

  [[nodiscard]] Vec2 Translation() const { return {x, y}; }
  /// R(theta) p (pose.hpp:19-25).
  [[nodiscard]] Vec2 Rotate(const Vec2& p) const {
    double c, s;
    CosSin(theta, c, s);
    return {c * p.x - s * p.y, s * p.x + c * p.y};
  }
  /// R(theta)^T p.
  [[nodiscard]] Vec2 RotateInverse(const Vec2& p) const {
    double c, s;
    CosSin(theta, c, s);
    return {c * p.x + s * p.y, -s * p.x + c * p.y};
  }
  /// R p + t (pose.hpp:27-30).
  [[nodiscard]] Vec2 Apply(const Vec2& p) const { return Rotate(p) + Translation(); }
  /// this o rel (pose.hpp:32-36).
  [[nodiscard]] Pose2 Compose(const Pose2& rel) const {
    const Vec2 t = Apply(rel.Translation());
    return {t.x, t.y, WrapAngle(theta + rel.theta)};
  }
  /// dt = R(-theta)(t_other - t), dtheta wrapped (pose.hpp:38-43).
  [[nodiscard]] Pose2 RelativeTo(const Pose2& other) const {
    const Vec2 dt = RotateInverse(other.Translation() - Translation());
    return {dt.x, dt.y, WrapAngle(other.theta - theta)};
  }
};

/// SDF grid with row-major node values (grid_map.hpp:11-33).
class GridMap {
 public:
  GridMap() = default;
  GridMap(int nx, int ny, const Vec2& min_corner, const Vec2& max_corner,
          double initial_value = 0.0)
      : nx_(nx),
        ny_(ny),
        min_(min_corner),
        dx_((max_corner.x - min_corner.x) / (nx - 1)),
        dy_((max_corner.y - min_corner.y) / (ny - 1)),
        values_(static_cast<size_t>(nx) * static_cast<size_t>(ny), initial_value) {
    if (nx < 2 || ny < 2 || !(dx_ > 0.0) || !(dy_ > 0.0)) {
      throw std::invalid_argument("grid needs nx, ny >= 2 and max > min");
    }
  }

  /// A grid given by its spacing (as a sdfslam_grid carries it) rather than
  /// its corners, so node positions are min + w * dx with this exact dx.
  static GridMap FromSpacing(int nx, int ny, const Vec2& min_corner, double dx, double dy) {
    GridMap g(nx, ny, min_corner, Vec2{min_corner.x + 1.0, min_corner.y + 1.0});
    g.dx_ = dx;
    g.dy_ = dy;
    if (!(dx > 0.0) || !(dy > 0.0)) throw std::invalid_argument("grid needs dx, dy > 0");
    return g;
  }

  [[nodiscard]] int nx() const { return nx_; }
  [[nodiscard]] int ny() const { return ny_; }
  [[nodiscard]] double dx() const { return dx_; }
  [[nodiscard]] double dy() const { return dy_; }
  [[nodiscard]] Vec2 min_corner() const { return min_; }
  [[nodiscard]] int num_nodes() const { return nx_ * ny_; }
  [[nodiscard]] int NodeId(int w, int h) const { return h * nx_ + w; }
  [[nodiscard]] Vec2 NodePosition(int w, int h) const {
    return {min_.x + w * dx_, min_.y + h * dy_};
  }
  [[nodiscard]] const std::vector<double>& values() const { return values_; }
  std::vector<double>& values() { return values_; }

  struct CellRef {
    int w{0};
    int h{0};
  };

  /// Cell of p with the reference's boundary rules: reject outside
  /// [0, n-1] in grid units, clamp the last row/column into the cell below or
  /// left (grid_map.hpp:58-70).
  [[nodiscard]] bool Locate(const Vec2& p, CellRef& cell) const {
    const double fx = (p.x - min_.x) / dx_;
    const double fy = (p.y - min_.y) / dy_;
    if (fx < 0.0 || fy < 0.0 || fx > nx_ - 1 || fy > ny_ - 1) {
      return false;
    }
    cell.w = std::min(static_cast<int>(fx), nx_ - 2);
    cell.h = std::min(static_cast<int>(fy), ny_ - 2);
    return true;
  }

 private:
  int nx_{2};
  int ny_{2};
  Vec2 min_{};
  double dx_{1.0};
  double dy_{1.0};
  std::vector<double> values_;
};

}  // namespace sdf_slam
