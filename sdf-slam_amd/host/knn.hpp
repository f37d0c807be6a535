// Exact k-nearest-neighbour search over 2D points with a kd-tree. Stands in
// for the reference's nanoflann kd-tree (src/core/normals.cpp:13-31, 70-78):
// same result set (the k closest points, the query point itself included
// when it is in the set); ties are broken by point index, and a subtree is
// skipped only when its box is strictly farther than the current k-th
// distance, so the result equals a brute-force search. Query cost does not
// grow with the distance from the query to the points (grid nodes far from
// any scan point are queried when node normals are assigned).
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
    const int n = static_cast<int>(points.size());
    order_.resize(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) order_[static_cast<size_t>(i)] = i;
    if (n > 0) {
      nodes_.reserve(static_cast<size_t>(2 * (n / kLeaf + 1)));
      Build(0, n);
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
    int stack[128];
    int top = 0;
    stack[top++] = 0;
    while (top > 0) {
      const Node& node = nodes_[static_cast<size_t>(stack[--top])];
      if (found == k && BoxDist2(node, q) > d2[k - 1]) {
        continue;
      }
      if (node.left < 0) {
        for (int t = node.begin; t < node.end; ++t) {
          const int i = order_[static_cast<size_t>(t)];
          const Vec2 d = points_[static_cast<size_t>(i)] - q;
          Insert(i, d.x * d.x + d.y * d.y, k, found, idx, d2);
        }
        continue;
      }
      // Visit the nearer child first (pushed last).
      const Node& l = nodes_[static_cast<size_t>(node.left)];
      const Node& r = nodes_[static_cast<size_t>(node.right)];
      if (BoxDist2(l, q) <= BoxDist2(r, q)) {
        stack[top++] = node.right;
        stack[top++] = node.left;
      } else {
        stack[top++] = node.left;
        stack[top++] = node.right;
      }
    }
    return found;
  }

 private:
  static constexpr int kLeaf = 8;
  struct Node {
    double x0, y0, x1, y1;  // bounding box of the node's points
    int begin, end;         // range in order_
    int left{-1}, right{-1};
  };

  int Build(int begin, int end) {
    const int id = static_cast<int>(nodes_.size());
    nodes_.push_back(Node{0.0, 0.0, 0.0, 0.0, begin, end});
    double x0 = std::numeric_limits<double>::max();
    double y0 = x0;
    double x1 = -x0;
    double y1 = -x0;
    for (int t = begin; t < end; ++t) {
      const Vec2& p = points_[static_cast<size_t>(order_[static_cast<size_t>(t)])];
      x0 = std::min(x0, p.x);
      y0 = std::min(y0, p.y);
      x1 = std::max(x1, p.x);
      y1 = std::max(y1, p.y);
    }
    nodes_[static_cast<size_t>(id)].x0 = x0;
    nodes_[static_cast<size_t>(id)].y0 = y0;
    nodes_[static_cast<size_t>(id)].x1 = x1;
    nodes_[static_cast<size_t>(id)].y1 = y1;
    if (end - begin <= kLeaf) {
      return id;
    }
    const bool split_x = (x1 - x0) >= (y1 - y0);
    const int mid = begin + (end - begin) / 2;
    std::nth_element(order_.begin() + begin, order_.begin() + mid, order_.begin() + end,
                     [&](int a, int b) {
                       const Vec2& pa = points_[static_cast<size_t>(a)];
                       const Vec2& pb = points_[static_cast<size_t>(b)];
                       const double ka = split_x ? pa.x : pa.y;
                       const double kb = split_x ? pb.x : pb.y;
                       return ka != kb ? ka < kb : a < b;
                     });
    const int left = Build(begin, mid);
    const int right = Build(mid, end);
    nodes_[static_cast<size_t>(id)].left = left;
    nodes_[static_cast<size_t>(id)].right = right;
    return id;
  }

  static double BoxDist2(const Node& b, const Vec2& q) {
    const double dx = q.x < b.x0 ? b.x0 - q.x : (q.x > b.x1 ? q.x - b.x1 : 0.0);
    const double dy = q.y < b.y0 ? b.y0 - q.y : (q.y > b.y1 ? q.y - b.y1 : 0.0);
    return dx * dx + dy * dy;
  }

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

  const std::vector<Vec2>& points_;
  std::vector<int> order_;
  std::vector<Node> nodes_;
};

}  // namespace sdf_slam
