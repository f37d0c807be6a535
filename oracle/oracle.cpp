// CPU oracle for the joint-LM hot path - TEST INFRASTRUCTURE ONLY (see
// oracle.h). A restatement of the reference's serial CPU path; every
// function names the reference file:line it follows.
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <limits>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

#include "ordering.hpp"

namespace {

// Host threads of the threaded phases (assembly chunks, J^T J column ranges,
// -J^T r, LDL^T subtrees); 1 runs everything serially. Every threaded phase
// gives the serial result bit for bit (lm_solver.cpp:217-219, 237-274,
// 84-105; parallel_ldlt.hpp:11-30).
int g_threads = 1;

// Runs fn(t) for t in [0, threads), t = 0 on the calling thread.
template <typename F>
void ForThreads(int threads, F&& fn) {
  if (threads <= 1) {
    fn(0);
    return;
  }
  std::vector<std::thread> pool;
  for (int t = 1; t < threads; ++t) pool.emplace_back([&fn, t] { fn(t); });
  fn(0);
  for (std::thread& th : pool) th.join();
}

int64_t ChunkBound(int64_t n, int chunks, int k) { return n * k / chunks; }

struct Grid {
  int nx, ny;
  double min_x, min_y, dx, dy;
  explicit Grid(const sdfslam_grid& g)
      : nx(g.nx), ny(g.ny), min_x(g.min_x), min_y(g.min_y), dx(g.dx), dy(g.dy) {}
  int NodeId(int w, int h) const { return h * nx + w; }
};

struct Cell {
  int w, h;
  double alpha, beta;
  int ids[4];
  double weights[4];
};

// grid_map.hpp:109-115
void FillCell(const Grid& g, Cell& c) {
  c.ids[0] = g.NodeId(c.w, c.h);
  c.ids[1] = g.NodeId(c.w + 1, c.h);
  c.ids[2] = g.NodeId(c.w, c.h + 1);
  c.ids[3] = g.NodeId(c.w + 1, c.h + 1);
  const double a = c.alpha;
  const double b = c.beta;
  c.weights[0] = (1.0 - a) * (1.0 - b);
  c.weights[1] = a * (1.0 - b);
  c.weights[2] = (1.0 - a) * b;
  c.weights[3] = a * b;
}

// grid_map.hpp:58-70
bool Locate(const Grid& g, double x, double y, Cell& c) {
  const double fx = (x - g.min_x) / g.dx;
  const double fy = (y - g.min_y) / g.dy;
  if (fx < 0.0 || fy < 0.0 || fx > g.nx - 1 || fy > g.ny - 1) {
    return false;
  }
  c.w = std::min(static_cast<int>(fx), g.nx - 2);
  c.h = std::min(static_cast<int>(fy), g.ny - 2);
  c.alpha = fx - c.w;
  c.beta = fy - c.h;
  FillCell(g, c);
  return true;
}

// grid_map.hpp:75-83
void LocateClamped(const Grid& g, double x, double y, Cell& c) {
  const double fx = (x - g.min_x) / g.dx;
  const double fy = (y - g.min_y) / g.dy;
  c.w = std::clamp(static_cast<int>(fx), 0, g.nx - 2);
  c.h = std::clamp(static_cast<int>(fy), 0, g.ny - 2);
  c.alpha = fx - c.w;
  c.beta = fy - c.h;
  FillCell(g, c);
}

// grid_map.hpp:86-93: sum in node order starting from 0.
double Interpolate(const double* v, const Cell& c) {
  double value = 0.0;
  for (int i = 0; i < 4; ++i) {
    value += c.weights[i] * v[c.ids[i]];
  }
  return value;
}

// grid_map.hpp:98-106
void BilinearGradient(const Grid& g, const double* v, const Cell& c, double& gx, double& gy) {
  const double d00 = v[c.ids[0]];
  const double d10 = v[c.ids[1]];
  const double d01 = v[c.ids[2]];
  const double d11 = v[c.ids[3]];
  gx = ((1.0 - c.beta) * (d10 - d00) + c.beta * (d11 - d01)) / g.dx;
  gy = ((1.0 - c.alpha) * (d01 - d00) + c.alpha * (d11 - d10)) / g.dy;
}

// residual_math.hpp:40-57
void SmoothGradient(const Grid& g, const double* v, const Cell& c, double& gx, double& gy) {
  auto node_gradient = [&](int w, int h, double& ngx, double& ngy) {
    const int w_lo = std::max(w - 1, 0);
    const int w_hi = std::min(w + 1, g.nx - 1);
    const int h_lo = std::max(h - 1, 0);
    const int h_hi = std::min(h + 1, g.ny - 1);
    ngx = (v[g.NodeId(w_hi, h)] - v[g.NodeId(w_lo, h)]) / ((w_hi - w_lo) * g.dx);
    ngy = (v[g.NodeId(w, h_hi)] - v[g.NodeId(w, h_lo)]) / ((h_hi - h_lo) * g.dy);
  };
  double g00x, g00y, g10x, g10y, g01x, g01y, g11x, g11y;
  node_gradient(c.w, c.h, g00x, g00y);
  node_gradient(c.w + 1, c.h, g10x, g10y);
  node_gradient(c.w, c.h + 1, g01x, g01y);
  node_gradient(c.w + 1, c.h + 1, g11x, g11y);
  const double a = c.alpha;
  const double b = c.beta;
  const double s00 = (1.0 - a) * (1.0 - b);
  const double s10 = a * (1.0 - b);
  const double s01 = (1.0 - a) * b;
  const double s11 = a * b;
  gx = s00 * g00x + s10 * g10x + s01 * g01x + s11 * g11x;
  gy = s00 * g00y + s10 * g10y + s01 * g01y + s11 * g11y;
}

// residual_math.hpp:63-73
void SmoothGradientMeters(const Grid& g, const double* v, double x, double y, double step,
                          double& gx, double& gy) {
  Cell c;
  auto value = [&](double qx, double qy) {
    LocateClamped(g, qx, qy, c);
    return Interpolate(v, c);
  };
  gx = (value(x + step, y) - value(x - step, y)) / (2.0 * step);
  gy = (value(x, y + step) - value(x, y - step)) / (2.0 * step);
}

struct PointEval {
  bool valid = false;
  double residual = 0.0;
  int ids[4] = {0, 0, 0, 0};
  double d_nodes[4] = {0, 0, 0, 0};
  double d_pose[3] = {0, 0, 0};
};

// residual_math.hpp:75-106 with Pose2::Apply (pose.hpp:27-30).
PointEval EvalPoint(const Grid& g, const double* v, const double* pose, double px, double py,
                    double delta, bool smooth, double smooth_step) {
  PointEval out;
  const double c = std::cos(pose[2]);
  const double s = std::sin(pose[2]);
  const double gxp = c * px + (-s) * py + pose[0];
  const double gyp = s * px + c * py + pose[1];
  Cell cell;
  if (!Locate(g, gxp, gyp, cell)) {
    out.residual = -delta;
    return out;
  }
  out.valid = true;
  out.residual = Interpolate(v, cell) - delta;
  for (int i = 0; i < 4; ++i) {
    out.ids[i] = cell.ids[i];
    out.d_nodes[i] = cell.weights[i];
  }
  double gx, gy;
  if (smooth && smooth_step > 0.0) {
    SmoothGradientMeters(g, v, gxp, gyp, smooth_step, gx, gy);
  } else if (smooth) {
    SmoothGradient(g, v, cell, gx, gy);
  } else {
    BilinearGradient(g, v, cell, gx, gy);
  }
  const double dx_dtheta = -s * px - c * py;
  const double dy_dtheta = c * px - s * py;
  out.d_pose[0] = gx;
  out.d_pose[1] = gy;
  out.d_pose[2] = gx * dx_dtheta + gy * dy_dtheta;
  return out;
}

// residual_math.hpp:158-171
void EvalEikonal(const Grid& g, const double* v, int w, int h, double nx, double ny, double& r,
                 int ids[3], double d[3]) {
  const double d00 = v[g.NodeId(w, h)];
  const double d10 = v[g.NodeId(w + 1, h)];
  const double d01 = v[g.NodeId(w, h + 1)];
  const double gx = (d10 - d00) / g.dx;
  const double gy = (d01 - d00) / g.dy;
  r = 1.0 - (nx * gx + ny * gy);
  ids[0] = g.NodeId(w, h);
  ids[1] = g.NodeId(w + 1, h);
  ids[2] = g.NodeId(w, h + 1);
  d[0] = nx / g.dx + ny / g.dy;
  d[1] = -nx / g.dx;
  d[2] = -ny / g.dy;
}

double Wrap(double a) { return std::atan2(std::sin(a), std::cos(a)); }

// residual_math.hpp:182-205 (Eigen's 2x2 products expanded term by term).
void EvalOdom(const double* pi, const double* pj, const double* m, double r[3], double ji[9],
              double jj[9]) {
  const double c = std::cos(pi[2]);
  const double s = std::sin(pi[2]);
  const double dtx = pj[0] - pi[0];
  const double dty = pj[1] - pi[1];
  const double relx = c * dtx + s * dty;
  const double rely = (-s) * dtx + c * dty;
  r[0] = relx - m[0];
  r[1] = rely - m[1];
  r[2] = Wrap(pj[2] - pi[2] - m[2]);
  // -rot_neg * j_hat = [[-s, c], [-c, -s]]
  for (int k = 0; k < 9; ++k) {
    ji[k] = 0.0;
    jj[k] = 0.0;
  }
  ji[0] = -c;
  ji[1] = -s;
  ji[3] = s;
  ji[4] = -c;
  ji[2] = (-s) * dtx + c * dty;
  ji[5] = (-c) * dtx + (-s) * dty;
  ji[8] = -1.0;
  jj[0] = c;
  jj[1] = s;
  jj[3] = -s;
  jj[4] = c;
  jj[8] = 1.0;
}

struct Csc {
  int64_t rows = 0;
  int64_t cols = 0;
  std::vector<int64_t> col_ptr;
  std::vector<int32_t> row_idx;
  std::vector<double> values;
};

struct Triplet {
  int32_t row;
  int32_t col;
  double value;
};

// setFromTriplets without duplicates (lm_solver.cpp:64-66, 312-313): values
// placed by (column, row), rows ascending within a column.
Csc FromTriplets(int64_t rows, int64_t cols, const std::vector<Triplet>& t) {
  Csc m;
  m.rows = rows;
  m.cols = cols;
  m.col_ptr.assign(static_cast<size_t>(cols) + 1, 0);
  for (const Triplet& x : t) {
    ++m.col_ptr[static_cast<size_t>(x.col) + 1];
  }
  for (int64_t j = 0; j < cols; ++j) {
    m.col_ptr[static_cast<size_t>(j) + 1] += m.col_ptr[static_cast<size_t>(j)];
  }
  m.row_idx.resize(t.size());
  m.values.resize(t.size());
  std::vector<int64_t> fill(m.col_ptr.begin(), m.col_ptr.end() - 1);
  for (const Triplet& x : t) {
    const int64_t p = fill[static_cast<size_t>(x.col)]++;
    m.row_idx[static_cast<size_t>(p)] = x.row;
    m.values[static_cast<size_t>(p)] = x.value;
  }
  for (int64_t j = 0; j < cols; ++j) {
    const int64_t a = m.col_ptr[static_cast<size_t>(j)];
    const int64_t b = m.col_ptr[static_cast<size_t>(j) + 1];
    std::vector<std::pair<int32_t, double>> tmp;
    tmp.reserve(static_cast<size_t>(b - a));
    for (int64_t p = a; p < b; ++p) {
      tmp.emplace_back(m.row_idx[static_cast<size_t>(p)], m.values[static_cast<size_t>(p)]);
    }
    std::sort(tmp.begin(), tmp.end(),
              [](const auto& u, const auto& v) { return u.first < v.first; });
    for (int64_t p = a; p < b; ++p) {
      m.row_idx[static_cast<size_t>(p)] = tmp[static_cast<size_t>(p - a)].first;
      m.values[static_cast<size_t>(p)] = tmp[static_cast<size_t>(p - a)].second;
    }
  }
  return m;
}

Csc Transpose(const Csc& a) {
  Csc t;
  t.rows = a.cols;
  t.cols = a.rows;
  t.col_ptr.assign(static_cast<size_t>(t.cols) + 1, 0);
  for (const int32_t r : a.row_idx) {
    ++t.col_ptr[static_cast<size_t>(r) + 1];
  }
  for (int64_t j = 0; j < t.cols; ++j) {
    t.col_ptr[static_cast<size_t>(j) + 1] += t.col_ptr[static_cast<size_t>(j)];
  }
  t.row_idx.resize(a.row_idx.size());
  t.values.resize(a.values.size());
  std::vector<int64_t> fill(t.col_ptr.begin(), t.col_ptr.end() - 1);
  for (int64_t j = 0; j < a.cols; ++j) {
    for (int64_t p = a.col_ptr[static_cast<size_t>(j)]; p < a.col_ptr[static_cast<size_t>(j) + 1];
         ++p) {
      const int32_t r = a.row_idx[static_cast<size_t>(p)];
      const int64_t q = fill[static_cast<size_t>(r)]++;
      t.row_idx[static_cast<size_t>(q)] = static_cast<int32_t>(j);
      t.values[static_cast<size_t>(q)] = a.values[static_cast<size_t>(p)];
    }
  }
  return t;
}

int64_t NumParams(const sdfslam_problem_view* p) {
  return static_cast<int64_t>(p->num_active_nodes) + 3 * (static_cast<int64_t>(p->num_poses) - 1);
}
int64_t NumRows(const sdfslam_problem_view* p) {
  return p->num_points + p->num_eikonal + 3 * p->num_odom;
}
int PoseColumn(const sdfslam_problem_view* p, int frame) {
  return p->num_active_nodes + 3 * (frame - 1);
}

// problem.hpp:115-121
double HuberWeight(double huber_delta, double weighted) {
  if (huber_delta <= 0.0) {
    return 1.0;
  }
  const double abs_r = std::abs(weighted);
  return abs_r <= huber_delta ? 1.0 : huber_delta / abs_r;
}

// AssemblePointRows (lm_solver.cpp:30-61) for point rows [begin, end).
void AssemblePointRows(const sdfslam_problem_view* p, const Grid& g, const double* map, const double* poses,
                       int64_t begin, int64_t end, double* residuals, std::vector<Triplet>& trip) {
  for (int64_t s = begin; s < end; ++s) {
    const int frame = p->point_frame[s];
    const PointEval e = EvalPoint(g, map, poses + 3 * frame, p->point_x[s], p->point_y[s],
                                  p->point_expected_sdf[s], p->smooth_gradient != 0,
                                  p->smooth_gradient_step);
    const auto row = static_cast<int32_t>(s);
    const double weighted = p->point_sqrt_weight[s] * e.residual;
    const double sw = std::sqrt(HuberWeight(p->huber_delta, weighted));
    residuals[s] = sw * weighted;
    if (!e.valid) {
      continue;
    }
    const double scale = sw * p->point_sqrt_weight[s];
    for (int k = 0; k < 4; ++k) {
      const int col = p->node_to_active[e.ids[k]];
      if (col >= 0) {
        trip.push_back({row, col, scale * e.d_nodes[k]});
      }
    }
    if (frame >= 1) {
      const int pc = PoseColumn(p, frame);
      for (int k = 0; k < 3; ++k) {
        trip.push_back({row, pc + k, scale * e.d_pose[k]});
      }
    }
  }
}

// AssembleEikonalRows (lm_solver.cpp:67-82) for eikonal rows [begin, end).
void AssembleEikonalRows(const sdfslam_problem_view* p, const Grid& g, const double* map, int64_t begin,
                         int64_t end, double* residuals, std::vector<Triplet>& trip) {
  for (int64_t e = begin; e < end; ++e) {
    double r;
    int ids[3];
    double d[3];
    EvalEikonal(g, map, p->eik_w[e], p->eik_h[e], p->eik_nx[e], p->eik_ny[e], r, ids, d);
    const auto row = static_cast<int32_t>(p->num_points + e);
    const double sw = p->eik_sqrt_weight[e];
    residuals[row] = sw * r;
    for (int k = 0; k < 3; ++k) {
      const int col = p->node_to_active[ids[k]];
      if (col >= 0) {
        trip.push_back({row, col, sw * d[k]});
      }
    }
  }
}

// Residuals and J (lm_solver.cpp:232-313): point and eikonal rows in
// g_threads contiguous chunks each (lm_solver.cpp:237-274), their triplets
// concatenated in chunk order, then the odometry tail, serially.
void Assemble(const sdfslam_problem_view* p, const double* map, const double* poses,
              double* residuals, Csc& jac) {
  const Grid g(p->grid);
  const int chunks = std::max(1, g_threads);
  std::vector<std::vector<Triplet>> parts(2 * static_cast<size_t>(chunks));
  ForThreads(chunks, [&](int t) {
    std::vector<Triplet>& pt = parts[static_cast<size_t>(t)];
    const int64_t b = ChunkBound(p->num_points, chunks, t);
    const int64_t e = ChunkBound(p->num_points, chunks, t + 1);
    pt.reserve(static_cast<size_t>(e - b) * 7);
    AssemblePointRows(p, g, map, poses, b, e, residuals, pt);
    std::vector<Triplet>& et = parts[static_cast<size_t>(chunks + t)];
    AssembleEikonalRows(p, g, map, ChunkBound(p->num_eikonal, chunks, t),
                        ChunkBound(p->num_eikonal, chunks, t + 1), residuals, et);
  });
  size_t total = static_cast<size_t>(p->num_odom) * 18;
  for (const auto& part : parts) total += part.size();
  std::vector<Triplet> trip;
  trip.reserve(total);
  for (const auto& part : parts) trip.insert(trip.end(), part.begin(), part.end());
  // Odometry tail (lm_solver.cpp:275-296).
  for (int64_t o = 0; o < p->num_odom; ++o) {
    const int fi = p->odom_i[o];
    const int fj = p->odom_j[o];
    const double m[3] = {p->odom_dx[o], p->odom_dy[o], p->odom_dtheta[o]};
    double r[3], ji[9], jj[9];
    EvalOdom(poses + 3 * fi, poses + 3 * fj, m, r, ji, jj);
    const auto row = static_cast<int32_t>(p->num_points + p->num_eikonal + 3 * o);
    const double sw = p->odom_sqrt_weight[o];
    for (int rr = 0; rr < 3; ++rr) {
      residuals[row + rr] = sw * r[rr];
      if (fi >= 1) {
        const int ci = PoseColumn(p, fi);
        for (int c = 0; c < 3; ++c) {
          trip.push_back({row + rr, ci + c, sw * ji[3 * rr + c]});
        }
      }
      const int cj = PoseColumn(p, fj);
      for (int c = 0; c < 3; ++c) {
        trip.push_back({row + rr, cj + c, sw * jj[3 * rr + c]});
      }
    }
  }
  jac = FromTriplets(NumRows(p), NumParams(p), trip);
}

void NormalColumns(const Csc& j, const Csc& jt, int64_t begin, int64_t end, Csc& out);

// ParallelAtA (lm_solver.cpp:113-172): column j accumulates rows k of
// column j ascending, inner i ascending within row k; first product assigned.
// Column ranges run on g_threads threads (lm_solver.cpp:84-105); each column
// keeps the serial accumulation order.
Csc NormalMatrix(const Csc& j) {
  const Csc jt = Transpose(j);
  const int chunks = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(g_threads, j.cols)));
  std::vector<Csc> parts(static_cast<size_t>(chunks));
  ForThreads(chunks, [&](int t) {
    NormalColumns(j, jt, ChunkBound(j.cols, chunks, t), ChunkBound(j.cols, chunks, t + 1),
                  parts[static_cast<size_t>(t)]);
  });
  Csc out;
  out.rows = j.cols;
  out.cols = j.cols;
  out.col_ptr.assign(static_cast<size_t>(j.cols) + 1, 0);
  size_t nnz = 0;
  for (const Csc& part : parts) nnz += part.row_idx.size();
  out.row_idx.reserve(nnz);
  out.values.reserve(nnz);
  int64_t col = 0;
  for (const Csc& part : parts) {
    const int64_t base = static_cast<int64_t>(out.row_idx.size());
    for (size_t c = 0; c + 1 < part.col_ptr.size(); ++c) {
      out.col_ptr[static_cast<size_t>(++col)] = base + part.col_ptr[c + 1];
    }
    out.row_idx.insert(out.row_idx.end(), part.row_idx.begin(), part.row_idx.end());
    out.values.insert(out.values.end(), part.values.begin(), part.values.end());
  }
  return out;
}

// Columns [begin, end) of J^T J into `out` (col_ptr relative to the range).
void NormalColumns(const Csc& j, const Csc& jt, int64_t begin, int64_t end, Csc& out) {
  out.col_ptr.assign(static_cast<size_t>(end - begin) + 1, 0);
  std::vector<char> mask(static_cast<size_t>(j.cols), 0);
  std::vector<double> values(static_cast<size_t>(j.cols), 0.0);
  std::vector<int32_t> indices;
  for (int64_t col = begin; col < end; ++col) {
    indices.clear();
    for (int64_t q = j.col_ptr[static_cast<size_t>(col)];
         q < j.col_ptr[static_cast<size_t>(col) + 1]; ++q) {
      const double y = j.values[static_cast<size_t>(q)];
      const int32_t k = j.row_idx[static_cast<size_t>(q)];
      for (int64_t t = jt.col_ptr[static_cast<size_t>(k)];
           t < jt.col_ptr[static_cast<size_t>(k) + 1]; ++t) {
        const auto i = static_cast<size_t>(jt.row_idx[static_cast<size_t>(t)]);
        const double x = jt.values[static_cast<size_t>(t)];
        if (mask[i] == 0) {
          mask[i] = 1;
          values[i] = x * y;
          indices.push_back(static_cast<int32_t>(i));
        } else {
          values[i] += x * y;
        }
      }
    }
    std::sort(indices.begin(), indices.end());
    for (const int32_t i : indices) {
      out.row_idx.push_back(i);
      out.values.push_back(values[static_cast<size_t>(i)]);
      mask[static_cast<size_t>(i)] = 0;
    }
    out.col_ptr[static_cast<size_t>(col - begin) + 1] = static_cast<int64_t>(out.row_idx.size());
  }
}

// lm_solver.cpp:321-330: coeffRef inserts a missing diagonal entry.
void Damp(Csc& a, double lambda, bool marquardt) {
  Csc out;
  out.rows = a.rows;
  out.cols = a.cols;
  out.col_ptr.assign(static_cast<size_t>(a.cols) + 1, 0);
  out.row_idx.reserve(a.row_idx.size() + static_cast<size_t>(a.cols));
  out.values.reserve(a.values.size() + static_cast<size_t>(a.cols));
  for (int64_t j = 0; j < a.cols; ++j) {
    bool placed = false;
    auto damp_value = [&](double diag) {
      if (marquardt) {
        return diag + lambda * std::max(diag, 1e-12);
      }
      return diag + lambda;
    };
    for (int64_t p = a.col_ptr[static_cast<size_t>(j)]; p < a.col_ptr[static_cast<size_t>(j) + 1];
         ++p) {
      const int32_t i = a.row_idx[static_cast<size_t>(p)];
      if (!placed && i > j) {
        out.row_idx.push_back(static_cast<int32_t>(j));
        out.values.push_back(damp_value(0.0));
        placed = true;
      }
      if (i == j) {
        out.row_idx.push_back(i);
        out.values.push_back(damp_value(a.values[static_cast<size_t>(p)]));
        placed = true;
      } else {
        out.row_idx.push_back(i);
        out.values.push_back(a.values[static_cast<size_t>(p)]);
      }
    }
    if (!placed) {
      out.row_idx.push_back(static_cast<int32_t>(j));
      out.values.push_back(damp_value(0.0));
    }
    out.col_ptr[static_cast<size_t>(j) + 1] = static_cast<int64_t>(out.row_idx.size());
  }
  a = std::move(out);
}

void NegAtbColumns(const Csc& j, const double* r, double* rhs, int64_t begin, int64_t end);

// ParallelNegAtb (lm_solver.cpp:178-197): two alternating accumulators.
void NegAtb(const Csc& j, const double* r, double* rhs) {
  const int chunks = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(g_threads, j.cols)));
  ForThreads(chunks, [&](int t) {
    NegAtbColumns(j, r, rhs, ChunkBound(j.cols, chunks, t), ChunkBound(j.cols, chunks, t + 1));
  });
}

void NegAtbColumns(const Csc& j, const double* r, double* rhs, int64_t begin, int64_t end) {
  for (int64_t col = begin; col < end; ++col) {
    double tmp_a = 0.0;
    double tmp_b = 0.0;
    int64_t p = j.col_ptr[static_cast<size_t>(col)];
    const int64_t end = j.col_ptr[static_cast<size_t>(col) + 1];
    while (p < end) {
      tmp_a += j.values[static_cast<size_t>(p)] * r[j.row_idx[static_cast<size_t>(p)]];
      ++p;
      if (p < end) {
        tmp_b += j.values[static_cast<size_t>(p)] * r[j.row_idx[static_cast<size_t>(p)]];
        ++p;
      }
    }
    rhs[col] = -(tmp_a + tmp_b);
  }
}

// Recursive coordinate bisection of the active grid nodes, poses last.
void Order(const sdfslam_problem_view* p, int32_t* order) {
  const int nx = p->grid.nx;
  const int ny = p->grid.ny;
  int64_t k = 0;
  std::function<void(int, int, int, int)> rec = [&](int w0, int w1, int h0, int h1) {
    if (w1 <= w0 || h1 <= h0) {
      return;
    }
    if ((w1 - w0) * (h1 - h0) <= 64) {
      for (int h = h0; h < h1; ++h) {
        for (int w = w0; w < w1; ++w) {
          const int col = p->node_to_active[h * nx + w];
          if (col >= 0) {
            order[k++] = col;
          }
        }
      }
      return;
    }
    if (w1 - w0 >= h1 - h0) {
      const int m = (w0 + w1) / 2;
      rec(w0, m, h0, h1);
      rec(m + 1, w1, h0, h1);
      rec(m, m + 1, h0, h1);
    } else {
      const int m = (h0 + h1) / 2;
      rec(w0, w1, h0, m);
      rec(w0, w1, m + 1, h1);
      rec(w0, w1, m, m + 1);
    }
  };
  rec(0, nx, 0, ny);
  for (int64_t c = p->num_active_nodes; c < NumParams(p); ++c) {
    order[k++] = static_cast<int32_t>(c);
  }
}

bool LdltRow(int k, int n, const std::vector<int64_t>& cp, const std::vector<int>& ci,
             const std::vector<double>& cx, const std::vector<int>& parent, const std::vector<int64_t>& lp,
             std::vector<int>& li, std::vector<double>& lx, std::vector<double>& diag, std::vector<int64_t>& nzc,
             std::vector<double>& y, std::vector<int>& pattern, std::vector<int>& tags);
void SubtreeSchedule(const std::vector<int>& parent, const std::vector<int64_t>& lnz, int threads,
                     std::vector<std::vector<int>>& rows_of, std::vector<int>& trunk);

// Statistics of the last factorization: nonzeros of L below the diagonal
// and sum over columns of c^2 + 3c (c = a column's count below the diagonal).
double g_last_factor_flops = 0.0;
int64_t g_last_factor_nnz = 0;

// LDL^T of P A P^T with the up-looking row kernel of parallel_ldlt.hpp:165-206
// (Eigen's factorize_preordered), then Eigen's SimplicialLDLT::solve order.
int LdltSolve(const Csc& a, const int32_t* order, const double* b, double* x, int64_t* lnz_out) {
  const int n = static_cast<int>(a.cols);
  std::vector<int> perm(static_cast<size_t>(n));
  std::vector<int> pinv(static_cast<size_t>(n));
  for (int k = 0; k < n; ++k) {
    perm[static_cast<size_t>(k)] = order != nullptr ? order[k] : k;
  }
  for (int k = 0; k < n; ++k) {
    pinv[static_cast<size_t>(perm[static_cast<size_t>(k)])] = k;
  }
  // Upper triangle of C = P A P^T by columns (permute_symm_to_symm<Lower, Upper>).
  std::vector<int64_t> cp(static_cast<size_t>(n) + 1, 0);
  for (int j = 0; j < n; ++j) {
    for (int64_t q = a.col_ptr[static_cast<size_t>(j)]; q < a.col_ptr[static_cast<size_t>(j) + 1];
         ++q) {
      const int i = a.row_idx[static_cast<size_t>(q)];
      if (i < j) {
        continue;  // lower triangle of a is the source
      }
      const int pi = pinv[static_cast<size_t>(i)];
      const int pj = pinv[static_cast<size_t>(j)];
      ++cp[static_cast<size_t>(std::max(pi, pj)) + 1];
    }
  }
  for (int j = 0; j < n; ++j) {
    cp[static_cast<size_t>(j) + 1] += cp[static_cast<size_t>(j)];
  }
  std::vector<int> ci(static_cast<size_t>(cp[static_cast<size_t>(n)]));
  std::vector<double> cx(ci.size());
  {
    std::vector<int64_t> fill(cp.begin(), cp.end() - 1);
    for (int j = 0; j < n; ++j) {
      for (int64_t q = a.col_ptr[static_cast<size_t>(j)];
           q < a.col_ptr[static_cast<size_t>(j) + 1]; ++q) {
        const int i = a.row_idx[static_cast<size_t>(q)];
        if (i < j) {
          continue;
        }
        const int pi = pinv[static_cast<size_t>(i)];
        const int pj = pinv[static_cast<size_t>(j)];
        const int col = std::max(pi, pj);
        const int64_t t = fill[static_cast<size_t>(col)]++;
        ci[static_cast<size_t>(t)] = std::min(pi, pj);
        cx[static_cast<size_t>(t)] = a.values[static_cast<size_t>(q)];
      }
    }
  }
  // Symbolic: elimination tree and column counts (LDL's ldl_symbolic).
  std::vector<int> parent(static_cast<size_t>(n), -1);
  std::vector<int> flag(static_cast<size_t>(n), -1);
  std::vector<int64_t> lnz(static_cast<size_t>(n), 0);
  for (int k = 0; k < n; ++k) {
    flag[static_cast<size_t>(k)] = k;
    for (int64_t q = cp[static_cast<size_t>(k)]; q < cp[static_cast<size_t>(k) + 1]; ++q) {
      int i = ci[static_cast<size_t>(q)];
      if (i < k) {
        for (; flag[static_cast<size_t>(i)] != k; i = parent[static_cast<size_t>(i)]) {
          if (parent[static_cast<size_t>(i)] == -1) {
            parent[static_cast<size_t>(i)] = k;
          }
          ++lnz[static_cast<size_t>(i)];
          flag[static_cast<size_t>(i)] = k;
        }
      }
    }
  }
  std::vector<int64_t> lp(static_cast<size_t>(n) + 1, 0);
  for (int k = 0; k < n; ++k) {
    lp[static_cast<size_t>(k) + 1] = lp[static_cast<size_t>(k)] + lnz[static_cast<size_t>(k)];
  }
  if (lnz_out != nullptr) {
    *lnz_out = lp[static_cast<size_t>(n)];
  }
  {
    double flops = 0.0;
    for (int k = 0; k < n; ++k) {
      const double c = static_cast<double>(lnz[static_cast<size_t>(k)]);
      flops += c * c + 3.0 * c;
    }
    g_last_factor_flops = flops;
    g_last_factor_nnz = lp[static_cast<size_t>(n)];
  }
  std::vector<int> li(static_cast<size_t>(lp[static_cast<size_t>(n)]));
  std::vector<double> lx(li.size());
  std::vector<double> diag(static_cast<size_t>(n));
  std::vector<int64_t> nzc(static_cast<size_t>(n), 0);
  // Rows over elimination-tree subtrees (parallel_ldlt.hpp:11-30, 223-239):
  // row k reads and appends to columns that are its descendants, so disjoint
  // subtrees run on separate threads and their common ancestors (the trunk)
  // after them; every column receives its entries in ascending row order,
  // as serially, so the factor is the serial one bit for bit.
  std::vector<std::vector<int>> rows_of;
  std::vector<int> trunk;
  SubtreeSchedule(parent, lnz, g_threads, rows_of, trunk);
  std::atomic<bool> ok{true};
  auto run_rows = [&](const std::vector<int>& rows) {
    std::vector<double> y(static_cast<size_t>(n), 0.0);
    std::vector<int> pattern(static_cast<size_t>(n), 0);
    std::vector<int> tags(static_cast<size_t>(n), -1);
    for (const int k : rows) {
      if (!ok.load(std::memory_order_relaxed)) return;
      if (!LdltRow(k, n, cp, ci, cx, parent, lp, li, lx, diag, nzc, y, pattern, tags)) {
        ok.store(false);
        return;
      }
    }
  };
  ForThreads(static_cast<int>(rows_of.size()), [&](int t) { run_rows(rows_of[static_cast<size_t>(t)]); });
  if (ok.load()) run_rows(trunk);
  if (!ok.load()) return SDFSLAM_ERR_NOT_SPD;
  // x = P^T L^-T D^-1 L^-1 P b
  std::vector<double> t(static_cast<size_t>(n));
  for (int k = 0; k < n; ++k) {
    t[static_cast<size_t>(k)] = b[perm[static_cast<size_t>(k)]];
  }
  for (int j = 0; j < n; ++j) {
    const double tj = t[static_cast<size_t>(j)];
    for (int64_t q = lp[static_cast<size_t>(j)]; q < lp[static_cast<size_t>(j) + 1]; ++q) {
      t[static_cast<size_t>(li[static_cast<size_t>(q)])] -= lx[static_cast<size_t>(q)] * tj;
    }
  }
  for (int j = 0; j < n; ++j) {
    t[static_cast<size_t>(j)] /= diag[static_cast<size_t>(j)];
  }
  for (int j = n - 1; j >= 0; --j) {
    double s = t[static_cast<size_t>(j)];
    for (int64_t q = lp[static_cast<size_t>(j)]; q < lp[static_cast<size_t>(j) + 1]; ++q) {
      s -= lx[static_cast<size_t>(q)] * t[static_cast<size_t>(li[static_cast<size_t>(q)])];
    }
    t[static_cast<size_t>(j)] = s;
  }
  for (int k = 0; k < n; ++k) {
    x[perm[static_cast<size_t>(k)]] = t[static_cast<size_t>(k)];
  }
  return SDFSLAM_OK;
}

// One row of the up-looking LDL^T (parallel_ldlt.hpp:165-206, Eigen's
// factorize_preordered): the pattern of row k by elimination-tree reach,
// d = fma(y_k, 1, 0), l_ki = y_i / d_i, fused y updates, unfused d -= l_ki y_i.
// Returns false at a zero pivot.
bool LdltRow(int k, int n, const std::vector<int64_t>& cp, const std::vector<int>& ci,
             const std::vector<double>& cx, const std::vector<int>& parent, const std::vector<int64_t>& lp,
             std::vector<int>& li, std::vector<double>& lx, std::vector<double>& diag, std::vector<int64_t>& nzc,
             std::vector<double>& y, std::vector<int>& pattern, std::vector<int>& tags) {
  {
    y[static_cast<size_t>(k)] = 0.0;
    int top = n;
    tags[static_cast<size_t>(k)] = k;
    nzc[static_cast<size_t>(k)] = 0;
    for (int64_t q = cp[static_cast<size_t>(k)]; q < cp[static_cast<size_t>(k) + 1]; ++q) {
      int i = ci[static_cast<size_t>(q)];
      if (i <= k) {
        y[static_cast<size_t>(i)] += cx[static_cast<size_t>(q)];
        int len = 0;
        for (; tags[static_cast<size_t>(i)] != k; i = parent[static_cast<size_t>(i)]) {
          pattern[static_cast<size_t>(len++)] = i;
          tags[static_cast<size_t>(i)] = k;
        }
        while (len > 0) {
          pattern[static_cast<size_t>(--top)] = pattern[static_cast<size_t>(--len)];
        }
      }
    }
    double d = std::fma(y[static_cast<size_t>(k)], 1.0, 0.0);
    y[static_cast<size_t>(k)] = 0.0;
    for (; top < n; ++top) {
      const int i = pattern[static_cast<size_t>(top)];
      const double yi = y[static_cast<size_t>(i)];
      y[static_cast<size_t>(i)] = 0.0;
      const double l_ki = yi / diag[static_cast<size_t>(i)];
      const int64_t p2 = lp[static_cast<size_t>(i)] + nzc[static_cast<size_t>(i)];
      int64_t q = lp[static_cast<size_t>(i)];
      for (; q < p2; ++q) {
        y[static_cast<size_t>(li[static_cast<size_t>(q)])] =
            std::fma(-lx[static_cast<size_t>(q)], yi, y[static_cast<size_t>(li[static_cast<size_t>(q)])]);
      }
      const double diag_update = l_ki * yi;
      d -= diag_update;
      li[static_cast<size_t>(q)] = k;
      lx[static_cast<size_t>(q)] = l_ki;
      ++nzc[static_cast<size_t>(i)];
    }
    diag[static_cast<size_t>(k)] = d;
    return d != 0.0;
  }
}

// Thread schedule of the numeric LDL^T (the role of BuildSchedule,
// parallel_ldlt.hpp:59-148, with a balance rule of its own): starting from
// the roots, the subtree of largest work (sum over its rows of 1 + the
// column count) is split, its root moving to the serial trunk, until no
// subtree holds more than a quarter of a thread's share; the subtrees then
// go to threads largest first, each to the least loaded. Rows ascend within
// each list. Any such partition gives the same factor.
void SubtreeSchedule(const std::vector<int>& parent, const std::vector<int64_t>& lnz, int threads,
                     std::vector<std::vector<int>>& rows_of, std::vector<int>& trunk) {
  const int n = static_cast<int>(parent.size());
  rows_of.assign(1, {});
  trunk.clear();
  if (threads <= 1 || n < 256) {
    rows_of[0].resize(static_cast<size_t>(n));
    for (int k = 0; k < n; ++k) rows_of[0][static_cast<size_t>(k)] = k;
    return;
  }
  std::vector<std::vector<int>> children(static_cast<size_t>(n));
  std::vector<double> work(static_cast<size_t>(n), 0.0);
  std::vector<int> roots;
  for (int k = 0; k < n; ++k) {
    const double c = static_cast<double>(lnz[static_cast<size_t>(k)]);
    work[static_cast<size_t>(k)] += 1.0 + c;
    const int p = parent[static_cast<size_t>(k)];
    if (p >= 0) {
      children[static_cast<size_t>(p)].push_back(k);
      work[static_cast<size_t>(p)] += work[static_cast<size_t>(k)];  // parent > k: k's total is final
    } else {
      roots.push_back(k);
    }
  }
  double total = 0.0;
  for (const int r : roots) total += work[static_cast<size_t>(r)];
  const double limit = total / (4.0 * threads);
  std::vector<int> parts = roots;
  while (true) {
    size_t big = 0;
    for (size_t q = 1; q < parts.size(); ++q) {
      if (work[static_cast<size_t>(parts[q])] > work[static_cast<size_t>(parts[big])]) big = q;
    }
    const int r = parts[big];
    if (work[static_cast<size_t>(r)] <= limit || children[static_cast<size_t>(r)].empty() ||
        parts.size() > 64 * static_cast<size_t>(threads)) {
      break;
    }
    trunk.push_back(r);
    parts.erase(parts.begin() + static_cast<std::ptrdiff_t>(big));
    for (const int c : children[static_cast<size_t>(r)]) parts.push_back(c);
  }
  std::sort(parts.begin(), parts.end(), [&](int a, int b) {
    return work[static_cast<size_t>(a)] > work[static_cast<size_t>(b)] ||
           (work[static_cast<size_t>(a)] == work[static_cast<size_t>(b)] && a < b);
  });
  rows_of.assign(static_cast<size_t>(threads), {});
  std::vector<double> load(static_cast<size_t>(threads), 0.0);
  std::vector<int> stack;
  for (const int r : parts) {
    const size_t t = static_cast<size_t>(std::min_element(load.begin(), load.end()) - load.begin());
    load[t] += work[static_cast<size_t>(r)];
    stack.push_back(r);
    while (!stack.empty()) {
      const int v = stack.back();
      stack.pop_back();
      rows_of[t].push_back(v);
      for (const int c : children[static_cast<size_t>(v)]) stack.push_back(c);
    }
  }
  for (auto& rows : rows_of) std::sort(rows.begin(), rows.end());
  std::sort(trunk.begin(), trunk.end());
}

// The fill-reducing order of the reference's LDL^T: approximate minimum
// degree on the pattern of the damped normal matrix (AMDOrdering applied by
// SimplicialLDLT::analyzePattern, parallel_ldlt.hpp:39-42), order[k] = the
// column eliminated k-th.
std::vector<int32_t> AmdOfMatrix(const Csc& a) {
  return oracle::AmdOrder(static_cast<int>(a.cols), a.col_ptr, a.row_idx);
}

// Problem::Cost (problem.cpp:132-156)
double Cost(const sdfslam_problem_view* p, const double* map, const double* poses) {
  const Grid g(p->grid);
  double cost = 0.0;
  for (int64_t s = 0; s < p->num_points; ++s) {
    const PointEval e = EvalPoint(g, map, poses + 3 * p->point_frame[s], p->point_x[s],
                                  p->point_y[s], p->point_expected_sdf[s], false, 0.0);
    const double r = p->point_sqrt_weight[s] * e.residual;
    if (p->huber_delta > 0.0 && std::abs(r) > p->huber_delta) {
      cost += p->huber_delta * (2.0 * std::abs(r) - p->huber_delta);
    } else {
      cost += r * r;
    }
  }
  for (int64_t e = 0; e < p->num_eikonal; ++e) {
    double r;
    int ids[3];
    double d[3];
    EvalEikonal(g, map, p->eik_w[e], p->eik_h[e], p->eik_nx[e], p->eik_ny[e], r, ids, d);
    const double wr = p->eik_sqrt_weight[e] * r;
    cost += wr * wr;
  }
  for (int64_t o = 0; o < p->num_odom; ++o) {
    const double m[3] = {p->odom_dx[o], p->odom_dy[o], p->odom_dtheta[o]};
    double r[3], ji[9], jj[9];
    EvalOdom(poses + 3 * p->odom_i[o], poses + 3 * p->odom_j[o], m, r, ji, jj);
    const double sw = p->odom_sqrt_weight[o];
    const double a = sw * r[0];
    const double b = sw * r[1];
    const double c = sw * r[2];
    cost += a * a + b * b + c * c;
  }
  return cost;
}

// Problem::ApplyStep (problem.cpp:117-130)
void ApplyStep(const sdfslam_problem_view* p, const double* step, double* map, double* poses) {
  const int nn = p->grid.nx * p->grid.ny;
  for (int node = 0; node < nn; ++node) {
    const int a = p->node_to_active[node];
    if (a >= 0) {
      map[node] += step[a];
    }
  }
  for (int i = 1; i < p->num_poses; ++i) {
    const int col = PoseColumn(p, i);
    poses[3 * i] += step[col];
    poses[3 * i + 1] += step[col + 1];
    poses[3 * i + 2] = Wrap(poses[3 * i + 2] + step[col + 2]);
  }
}

thread_local double g_phase_seconds[4] = {0, 0, 0, 0};

}  // namespace

struct oracle_matrix {
  Csc m;
};

extern "C" {

double oracle_wrap_angle(double angle) { return Wrap(angle); }

void oracle_eval_points_frozen(const sdfslam_grid* grid, const double* values, const double* poses,
                               int64_t n, const int32_t* frame, const double* px, const double* py,
                               const double* delta, const int32_t* cell_w, const int32_t* cell_h,
                               double* residual, double* d_nodes, double* d_pose) {
  for (int64_t k = 0; k < n; ++k) {
    const double* q = poses + 3 * static_cast<int64_t>(frame[k]);
    const double c = std::cos(q[2]);
    const double s = std::sin(q[2]);
    const double x = px[k];
    const double y = py[k];
    const double gx = c * x + (-s) * y + q[0];
    const double gy = s * x + c * y + q[1];
    const int w = cell_w[k];
    const int h = cell_h[k];
    const double ox = grid->min_x + w * grid->dx;  // GridMap::NodePosition
    const double oy = grid->min_y + h * grid->dy;
    const double alpha = (gx - ox) / grid->dx;
    const double beta = (gy - oy) / grid->dy;
    const double wt[4] = {(1.0 - alpha) * (1.0 - beta), alpha * (1.0 - beta), (1.0 - alpha) * beta,
                          alpha * beta};
    const int64_t base = static_cast<int64_t>(h) * grid->nx + w;
    const double v[4] = {values[base], values[base + 1], values[base + grid->nx], values[base + grid->nx + 1]};
    double r = -delta[k];
    for (int i = 0; i < 4; ++i) r += wt[i] * v[i];
    residual[k] = r;
    for (int i = 0; i < 4; ++i) d_nodes[4 * k + i] = wt[i];
    const double sx = ((1.0 - beta) * (v[1] - v[0]) + beta * (v[3] - v[2])) / grid->dx;
    const double sy = ((1.0 - alpha) * (v[2] - v[0]) + alpha * (v[3] - v[1])) / grid->dy;
    const double dx_dtheta = -s * x - c * y;
    const double dy_dtheta = c * x - s * y;
    d_pose[3 * k] = sx;
    d_pose[3 * k + 1] = sy;
    d_pose[3 * k + 2] = sx * dx_dtheta + sy * dy_dtheta;
  }
}

void oracle_init_from_scans(const sdfslam_grid* grid, double* map_values, const double* scan_xy,
                            const int64_t* scan_ptr, int32_t num_scans, const double* poses,
                            int32_t num_poses, int signed_distance) {
  // map_init.cpp:34-43: the scans placed by their poses, each point with its
  // pose's translation as origin (Pose2::Apply, pose.hpp:27-30).
  std::vector<double> px, py, ox, oy;
  for (int32_t i = 0; i < num_scans && i < num_poses; ++i) {
    const double* q = poses + 3 * i;
    const double c = std::cos(q[2]);
    const double s = std::sin(q[2]);
    for (int64_t k = scan_ptr[i]; k < scan_ptr[i + 1]; ++k) {
      const double x = scan_xy[2 * k];
      const double y = scan_xy[2 * k + 1];
      px.push_back(c * x + (-s) * y + q[0]);
      py.push_back(s * x + c * y + q[1]);
      ox.push_back(q[0]);
      oy.push_back(q[1]);
    }
  }
  if (px.empty()) return;  // map_init.cpp:44-46
  // map_init.cpp:51-68, with an exhaustive search in place of the k-d tree.
  for (int h = 0; h < grid->ny; ++h) {
    for (int w = 0; w < grid->nx; ++w) {
      const double qx = grid->min_x + w * grid->dx;
      const double qy = grid->min_y + h * grid->dy;
      size_t index = 0;
      double squared = std::numeric_limits<double>::max();
      for (size_t i = 0; i < px.size(); ++i) {
        const double d0 = qx - px[i];
        const double d1 = qy - py[i];
        const double d = d0 * d0 + d1 * d1;
        if (d < squared) {
          squared = d;
          index = i;
        }
      }
      double sign = 1.0;
      if (signed_distance != 0) {
        const double ax = qx - ox[index];
        const double ay = qy - oy[index];
        const double bx = px[index] - ox[index];
        const double by = py[index] - oy[index];
        const double to_node = std::sqrt(ax * ax + ay * ay);
        const double to_surface = std::sqrt(bx * bx + by * by);
        sign = to_node <= to_surface ? 1.0 : -1.0;
      }
      map_values[static_cast<size_t>(h) * grid->nx + w] = sign * std::sqrt(squared);
    }
  }
}

void oracle_compose(const double a[3], const double rel[3], double out[3]) {
  const double c = std::cos(a[2]);
  const double s = std::sin(a[2]);
  out[0] = c * rel[0] + (-s) * rel[1] + a[0];
  out[1] = s * rel[0] + c * rel[1] + a[1];
  out[2] = Wrap(a[2] + rel[2]);
}

void oracle_relative(const double a[3], const double b[3], double out[3]) {
  const double c = std::cos(a[2]);
  const double s = std::sin(a[2]);
  const double dx = b[0] - a[0];
  const double dy = b[1] - a[1];
  out[0] = c * dx + s * dy;
  out[1] = (-s) * dx + c * dy;
  out[2] = Wrap(b[2] - a[2]);
}

int oracle_locate(const sdfslam_grid* grid, double x, double y, int clamped, int32_t* w,
                  int32_t* h, double* alpha, double* beta) {
  const Grid g(*grid);
  Cell c{};
  int ok = 1;
  if (clamped != 0) {
    LocateClamped(g, x, y, c);
  } else {
    ok = Locate(g, x, y, c) ? 1 : 0;
  }
  *w = c.w;
  *h = c.h;
  *alpha = c.alpha;
  *beta = c.beta;
  return ok;
}

double oracle_interpolate(const sdfslam_grid* grid, const double* values, double x, double y,
                          int clamped) {
  const Grid g(*grid);
  Cell c{};
  if (clamped != 0) {
    LocateClamped(g, x, y, c);
  } else if (!Locate(g, x, y, c)) {
    return 0.0;
  }
  return Interpolate(values, c);
}

int oracle_eval_point(const sdfslam_grid* grid, const double* values, const double pose[3],
                      double px, double py, double delta, int smooth_gradient,
                      double smooth_gradient_step, double* residual, int32_t node_ids[4],
                      double d_nodes[4], double d_pose[3]) {
  const Grid g(*grid);
  const PointEval e =
      EvalPoint(g, values, pose, px, py, delta, smooth_gradient != 0, smooth_gradient_step);
  *residual = e.residual;
  for (int i = 0; i < 4; ++i) {
    node_ids[i] = e.ids[i];
    d_nodes[i] = e.d_nodes[i];
  }
  for (int i = 0; i < 3; ++i) {
    d_pose[i] = e.d_pose[i];
  }
  return e.valid ? 1 : 0;
}

void oracle_eval_eikonal(const sdfslam_grid* grid, const double* values, int32_t w, int32_t h,
                         double nx, double ny, double* residual, int32_t node_ids[3],
                         double d_nodes[3]) {
  const Grid g(*grid);
  int ids[3];
  EvalEikonal(g, values, w, h, nx, ny, *residual, ids, d_nodes);
  for (int i = 0; i < 3; ++i) {
    node_ids[i] = ids[i];
  }
}

void oracle_eval_odom(const double pose_i[3], const double pose_j[3], const double meas[3],
                      double residual[3], double d_pose_i[9], double d_pose_j[9]) {
  EvalOdom(pose_i, pose_j, meas, residual, d_pose_i, d_pose_j);
}

oracle_matrix* oracle_matrix_from_csc(int64_t rows, int64_t cols, const int64_t* col_ptr,
                                      const int32_t* row_idx, const double* values) {
  auto* m = new oracle_matrix;
  m->m.rows = rows;
  m->m.cols = cols;
  m->m.col_ptr.assign(col_ptr, col_ptr + cols + 1);
  m->m.row_idx.assign(row_idx, row_idx + col_ptr[cols]);
  m->m.values.assign(values, values + col_ptr[cols]);
  return m;
}

void oracle_matrix_free(oracle_matrix* m) { delete m; }

void oracle_matrix_shape(const oracle_matrix* m, int64_t* rows, int64_t* cols, int64_t* nnz) {
  *rows = m->m.rows;
  *cols = m->m.cols;
  *nnz = static_cast<int64_t>(m->m.row_idx.size());
}

void oracle_matrix_copy(const oracle_matrix* m, int64_t* col_ptr, int32_t* row_idx,
                        double* values) {
  std::copy(m->m.col_ptr.begin(), m->m.col_ptr.end(), col_ptr);
  std::copy(m->m.row_idx.begin(), m->m.row_idx.end(), row_idx);
  std::copy(m->m.values.begin(), m->m.values.end(), values);
}

int oracle_assemble(const sdfslam_problem_view* problem, const double* map_values,
                    const double* poses, double* residuals, oracle_matrix** jacobian) {
  auto* m = new oracle_matrix;
  Assemble(problem, map_values, poses, residuals, m->m);
  *jacobian = m;
  return SDFSLAM_OK;
}

oracle_matrix* oracle_normal_matrix(const oracle_matrix* jacobian) {
  auto* m = new oracle_matrix;
  m->m = NormalMatrix(jacobian->m);
  return m;
}

void oracle_damp(oracle_matrix* a, double lambda, int marquardt_scaling) {
  Damp(a->m, lambda, marquardt_scaling != 0);
}

void oracle_neg_atb(const oracle_matrix* jacobian, const double* residuals, double* rhs) {
  NegAtb(jacobian->m, residuals, rhs);
}

void oracle_order(const sdfslam_problem_view* problem, int32_t* order) { Order(problem, order); }

int oracle_ldlt_solve(const oracle_matrix* a, const int32_t* order, const double* b, double* x,
                      int64_t* factor_nnz) {
  return LdltSolve(a->m, order, b, x, factor_nnz);
}

double oracle_cost(const sdfslam_problem_view* problem, const double* map_values,
                   const double* poses) {
  return Cost(problem, map_values, poses);
}

void oracle_apply_step(const sdfslam_problem_view* problem, const double* step,
                       double* map_values, double* poses) {
  ApplyStep(problem, step, map_values, poses);
}

int oracle_linear_step(const sdfslam_problem_view* problem, const double* map_values,
                       const double* poses, double lambda, int marquardt_scaling, double* step) {
  const int64_t rows = NumRows(problem);
  const int64_t n = NumParams(problem);
  std::vector<double> r(static_cast<size_t>(rows));
  Csc j;
  Assemble(problem, map_values, poses, r.data(), j);
  Csc a = NormalMatrix(j);
  Damp(a, lambda, marquardt_scaling != 0);
  std::vector<double> rhs(static_cast<size_t>(n));
  NegAtb(j, r.data(), rhs.data());
  const std::vector<int32_t> order = AmdOfMatrix(a);
  return LdltSolve(a, order.data(), rhs.data(), step, nullptr);
}

// SolveLm (lm_solver.cpp:201-434)
int oracle_solve(const sdfslam_problem_view* problem, const sdfslam_solver_options* options,
                 double* map_values, double* poses, sdfslam_solve_result* result) {
  return oracle_solve_recorded(problem, options, map_values, poses, result, 0, nullptr, nullptr,
                               nullptr);
}

int oracle_solve_recorded(const sdfslam_problem_view* problem, const sdfslam_solver_options* options,
                          double* map_values, double* poses, sdfslam_solve_result* result,
                          int32_t num_records, const int32_t* record_iterations, double* record_map,
                          double* record_poses) {
  using clock = std::chrono::steady_clock;
  const auto start = clock::now();
  for (double& s : g_phase_seconds) {
    s = 0.0;
  }
  const int64_t rows = NumRows(problem);
  const int64_t n = NumParams(problem);
  const size_t map_size = static_cast<size_t>(problem->grid.nx) * problem->grid.ny;
  const size_t pose_size = 3 * static_cast<size_t>(problem->num_poses);
  std::vector<double> history;
  result->converged = 0;
  result->iterations = 0;
  result->initial_cost = Cost(problem, map_values, poses);
  history.push_back(result->initial_cost);
  double lambda = options->lambda_init;
  double previous_cost = result->initial_cost;
  double nielsen_nu = 2.0;
  // The ordering is recomputed when the pattern of the damped normal matrix
  // changes (lm_solver.cpp:333-350).
  std::vector<int32_t> order;
  std::vector<int64_t> pattern_ptr;
  std::vector<int32_t> pattern_idx;
  std::vector<double> r(static_cast<size_t>(rows));
  std::vector<double> rhs(static_cast<size_t>(n));
  std::vector<double> step(static_cast<size_t>(n));
  std::vector<double> map_backup;
  std::vector<double> pose_backup;
  int status = SDFSLAM_OK;
  for (int iter = 0; iter < options->max_iterations; ++iter) {
    auto t0 = clock::now();
    Csc j;
    Assemble(problem, map_values, poses, r.data(), j);
    auto t1 = clock::now();
    Csc a = NormalMatrix(j);
    Damp(a, lambda, options->marquardt_scaling != 0);
    NegAtb(j, r.data(), rhs.data());
    if (a.col_ptr != pattern_ptr || a.row_idx != pattern_idx) {
      order = AmdOfMatrix(a);
      pattern_ptr = a.col_ptr;
      pattern_idx = a.row_idx;
    }
    auto t2 = clock::now();
    status = LdltSolve(a, order.data(), rhs.data(), step.data(), nullptr);
    auto t3 = clock::now();
    if (status != SDFSLAM_OK) {
      break;
    }
    if (options->reject_worse_steps != 0 || options->trust_region != 0) {
      map_backup.assign(map_values, map_values + map_size);
      pose_backup.assign(poses, poses + pose_size);
    }
    ApplyStep(problem, step.data(), map_values, poses);
    const double cost = Cost(problem, map_values, poses);
    auto t4 = clock::now();
    g_phase_seconds[0] += std::chrono::duration<double>(t1 - t0).count();
    g_phase_seconds[1] += std::chrono::duration<double>(t2 - t1).count();
    g_phase_seconds[2] += std::chrono::duration<double>(t3 - t2).count();
    g_phase_seconds[3] += std::chrono::duration<double>(t4 - t3).count();
    result->iterations = iter + 1;
    double norm_sq = 0.0;
    for (const double s : step) {
      norm_sq += s * s;
    }
    if (std::sqrt(norm_sq) < options->step_tolerance) {
      history.push_back(cost);
      result->final_cost = cost;
      result->converged = 1;
      break;
    }
    double current_cost = cost;
    if (options->trust_region != 0) {
      double predicted = 0.0;
      for (int64_t k = 0; k < n; ++k) {
        predicted += step[static_cast<size_t>(k)] *
                     (lambda * step[static_cast<size_t>(k)] + rhs[static_cast<size_t>(k)]);
      }
      const double rho = predicted > 0.0 ? (previous_cost - cost) / predicted : -1.0;
      if (rho > 0.0) {
        previous_cost = cost;
        const double shrink = 1.0 - std::pow(2.0 * rho - 1.0, 3);
        lambda = std::max(lambda * std::max(1.0 / 3.0, shrink), 1e-7);
        nielsen_nu = 2.0;
      } else {
        std::copy(map_backup.begin(), map_backup.end(), map_values);
        std::copy(pose_backup.begin(), pose_backup.end(), poses);
        current_cost = previous_cost;
        lambda *= nielsen_nu;
        nielsen_nu *= 2.0;
      }
    } else if (cost < previous_cost) {
      lambda /= options->lambda_factor;
      previous_cost = cost;
    } else {
      lambda *= options->lambda_factor;
      if (options->reject_worse_steps != 0) {
        std::copy(map_backup.begin(), map_backup.end(), map_values);
        std::copy(pose_backup.begin(), pose_backup.end(), poses);
        current_cost = previous_cost;
      } else {
        previous_cost = cost;
      }
    }
    history.push_back(current_cost);
    result->final_cost = current_cost;
    // The state the iteration ends in, for every requested iteration count.
    for (int32_t q = 0; q < num_records; ++q) {
      if (record_iterations[q] == iter + 1) {
        if (record_map != nullptr) {
          std::copy(map_values, map_values + map_size, record_map + static_cast<size_t>(q) * map_size);
        }
        if (record_poses != nullptr) {
          std::copy(poses, poses + pose_size, record_poses + static_cast<size_t>(q) * pose_size);
        }
      }
    }
  }
  if (result->iterations == 0) {
    result->final_cost = result->initial_cost;
  }
  result->cost_history_length = static_cast<int32_t>(history.size());
  if (result->cost_history != nullptr) {
    const size_t cnt = std::min(history.size(), static_cast<size_t>(std::max(0, result->cost_history_capacity)));
    std::copy(history.begin(), history.begin() + static_cast<std::ptrdiff_t>(cnt), result->cost_history);
  }
  result->duration_seconds = std::chrono::duration<double>(clock::now() - start).count();
  return status;
}

void oracle_set_threads(int threads) { g_threads = std::max(1, threads); }

int oracle_threads(void) { return g_threads; }

void oracle_last_factor_stats(int64_t* factor_nnz, double* factor_flops) {
  if (factor_nnz != nullptr) *factor_nnz = g_last_factor_nnz;
  if (factor_flops != nullptr) *factor_flops = g_last_factor_flops;
}

void oracle_amd_order(const oracle_matrix* a, int32_t* order) {
  const std::vector<int32_t> o = AmdOfMatrix(a->m);
  std::copy(o.begin(), o.end(), order);
}

void oracle_last_phase_seconds(double out[4]) {
  for (int i = 0; i < 4; ++i) {
    out[i] = g_phase_seconds[i];
  }
}

}  // extern "C"
