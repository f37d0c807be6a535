// Exact k-nearest-neighbour search over 2D points with a uniform bucket grid.
// Stands in for the reference's nanoflann kd-tree (src/core/normals.cpp:13-31,
// 70-78): same result set (the k closest points, the query point itself
// included when it is in the set); ties are broken by point index.
#pragma once

#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

#include "geometry.hpp"

namespace sdf_slam {

class PointGrid {
 public:
  explicit PointGrid(const std::vector<Vec2>& points) : points_(points) {
    const size_t n = points.size();
    double x0 = std::numeric_limits<double>::max();
    double y0 = x0;
    double x1 = -x0;
    double y1 = -x0;
    for (const Vec2& p : points) {
      x0 = std::min(x0, p.x);
      y0 = std::min(y0, p.y);
      x1 = std::max(x1, p.x);
      y1 = std::max(y1, p.y);
    }
    if (n == 0) {
      x0 = y0 = 0.0;
      x1 = y1 = 1.0;
    }
    const double area = std::max((x1 - x0) * (y1 - y0), 1e-12);
    h_ = std::max(std::sqrt(area / std::max<double>(1.0, static_cast<double>(n))) * 1.5, 1e-9);
    gx_ = std::clamp(static_cast<int>((x1 - x0) / h_) + 1, 1, 4096);
    gy_ = std::clamp(static_cast<int>((y1 - y0) / h_) + 1, 1, 4096);
    h_ = std::max((x1 - x0) / gx_, (y1 - y0) / gy_) * (1.0 + 1e-12) + 1e-12;
    x0_ = x0;
    y0_ = y0;
    start_.assign(static_cast<size_t>(gx_) * static_cast<size_t>(gy_) + 1, 0);
    std::vector<int> cell_of(n);
    for (size_t i = 0; i < n; ++i) {
      const int c = CellIndex(CellX(points[i].x), CellY(points[i].y));
      cell_of[i] = c;
      ++start_[static_cast<size_t>(c) + 1];
    }
    for (size_t c = 0; c + 1 < start_.size(); ++c) {
      start_[c + 1] += start_[c];
    }
    items_.resize(n);
    std::vector<int> fill(start_.begin(), start_.end() - 1);
    for (size_t i = 0; i < n; ++i) {
      items_[static_cast<size_t>(fill[static_cast<size_t>(cell_of[i])]++)] = static_cast<int>(i);
    }
  }

  /// Fills idx/d2 with the k nearest points sorted by (distance, index);
  /// returns the number found (min(k, n)).
  int Knn(const Vec2& q, int k, int* idx, double* d2) const {
    const int n = static_cast<int>(points_.size());
    k = std::min(k, n);
    if (k <= 0) {
      return 0;
    }
    int found = 0;
    const int cx = std::clamp(CellX(q.x), 0, gx_ - 1);
    const int cy = std::clamp(CellY(q.y), 0, gy_ - 1);
    const int max_r = std::max(gx_, gy_);
    for (int r = 0; r <= max_r; ++r) {
      const int xa = cx - r;
      const int xb = cx + r;
      const int ya = cy - r;
      const int yb = cy + r;
      for (int y = std::max(ya, 0); y <= std::min(yb, gy_ - 1); ++y) {
        const bool edge_row = (y == ya || y == yb);
        for (int x = std::max(xa, 0); x <= std::min(xb, gx_ - 1); ++x) {
          if (!edge_row && x != xa && x != xb) {
            continue;
          }
          const int c = CellIndex(x, y);
          for (int t = start_[static_cast<size_t>(c)]; t < start_[static_cast<size_t>(c) + 1]; ++t) {
            const int i = items_[static_cast<size_t>(t)];
            const Vec2 d = points_[static_cast<size_t>(i)] - q;
            const double dd = d.x * d.x + d.y * d.y;
            Insert(i, dd, k, found, idx, d2);
          }
        }
      }
      if (xa <= 0 && ya <= 0 && xb >= gx_ - 1 && yb >= gy_ - 1) {
        break;
      }
      if (found == k) {
        // Lower bound on the distance from q to any cell not yet searched:
        // only sides of the block that still have grid cells beyond them count.
        double gap = std::numeric_limits<double>::max();
        if (xa > 0) {
          gap = std::min(gap, q.x - (x0_ + xa * h_));
        }
        if (xb < gx_ - 1) {
          gap = std::min(gap, (x0_ + (xb + 1) * h_) - q.x);
        }
        if (ya > 0) {
          gap = std::min(gap, q.y - (y0_ + ya * h_));
        }
        if (yb < gy_ - 1) {
          gap = std::min(gap, (y0_ + (yb + 1) * h_) - q.y);
        }
        if (gap > 0.0 && d2[k - 1] <= gap * gap) {
          break;
        }
      }
    }
    return found;
  }

 private:
  static void Insert(int i, double dd, int k, int& found, int* idx, double* d2) {
    if (found == k) {
      if (dd > d2[k - 1] || (dd == d2[k - 1] && i > idx[k - 1])) {
        return;
      }
      --found;
    }
    int pos = found;
    while (pos > 0 && (d2[pos - 1] > dd || (d2[pos - 1] == dd && idx[pos - 1] > i))) {
      d2[pos] = d2[pos - 1];
      idx[pos] = idx[pos - 1];
      --pos;
    }
    d2[pos] = dd;
    idx[pos] = i;
    ++found;
  }
  [[nodiscard]] int CellX(double x) const {
    return std::clamp(static_cast<int>(std::floor((x - x0_) / h_)), -1, gx_);
  }
  [[nodiscard]] int CellY(double y) const {
    return std::clamp(static_cast<int>(std::floor((y - y0_) / h_)), -1, gy_);
  }
  [[nodiscard]] int CellIndex(int x, int y) const {
    x = std::clamp(x, 0, gx_ - 1);
    y = std::clamp(y, 0, gy_ - 1);
    return y * gx_ + x;
  }

  const std::vector<Vec2>& points_;
  double x0_{0.0};
  double y0_{0.0};
  double h_{1.0};
  int gx_{1};
  int gy_{1};
  std::vector<int> start_;
  std::vector<int> items_;
};

}  // namespace sdf_slam
