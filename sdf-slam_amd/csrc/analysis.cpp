// Symbolic analysis for the device backend (see analysis.hpp).
#include "analysis.hpp"

#include <algorithm>
#include <cmath>
#include <functional>
#include <stdexcept>
#include <string>

namespace sdfslam {

namespace {

struct NdNode {
  int w0, w1, h0, h1;
  int left{-1};
  int right{-1};
  int axis{-1};  // -1 leaf, 0 separator at w = m, 1 separator at h = m
  int m{0};
  std::vector<int> poses;
};

bool BoxContains(const NdNode& t, int bw0, int bw1, int bh0, int bh1) {
  return bw0 >= t.w0 && bw1 < t.w1 && bh0 >= t.h0 && bh1 < t.h1;
}

}  // namespace

int64_t PanelOffset(const Analysis& a, int i, int j) {
  const int s = a.snode_of[static_cast<size_t>(j)];
  const int32_t* begin = a.sn_idx.data() + a.sn_idx_ptr[static_cast<size_t>(s)];
  const int32_t* end = a.sn_idx.data() + a.sn_idx_ptr[static_cast<size_t>(s) + 1];
  const int32_t* it = std::lower_bound(begin, end, i);
  if (it == end || *it != i) {
    return -1;
  }
  const int64_t row = it - begin;
  const int64_t col = j - a.sn_first[static_cast<size_t>(s)];
  return a.panel_off[static_cast<size_t>(s)] + col * a.sn_f[static_cast<size_t>(s)] + row;
}

Analysis Analyze(const sdfslam_problem_view& v, const double* map_values, const double* poses,
                 const AnalysisOptions& opt) {
  (void)map_values;
  Analysis a;
  a.nx = v.grid.nx;
  a.ny = v.grid.ny;
  const int ncx = a.nx - 1;
  const int ncy = a.ny - 1;
  a.num_cells = ncx * ncy;
  a.num_active = v.num_active_nodes;
  a.num_poses = v.num_poses;
  a.n = a.num_active + 3 * (a.num_poses - 1);
  const int num_nodes = a.nx * a.ny;
  const int N = a.num_poses;
  const int margin = std::max(1, opt.margin);

  std::vector<int32_t> active_to_node(static_cast<size_t>(a.num_active), -1);
  for (int g = 0; g < num_nodes; ++g) {
    const int c = v.node_to_active[g];
    if (c >= 0) {
      if (c >= a.num_active || active_to_node[static_cast<size_t>(c)] >= 0) {
        throw std::invalid_argument("node_to_active is not a bijection onto the active columns");
      }
      active_to_node[static_cast<size_t>(c)] = g;
    }
  }

  // Point rows per pose.
  a.pr_ptr.assign(static_cast<size_t>(N) + 1, 0);
  for (int64_t s = 0; s < v.num_points; ++s) {
    const int f = v.point_frame[s];
    if (f < 0 || f >= N) {
      throw std::invalid_argument("point frame out of range");
    }
    ++a.pr_ptr[static_cast<size_t>(f) + 1];
  }
  for (int f = 0; f < N; ++f) {
    a.pr_ptr[static_cast<size_t>(f) + 1] += a.pr_ptr[static_cast<size_t>(f)];
    a.max_points_per_pose =
        std::max<int>(a.max_points_per_pose,
                      static_cast<int>(a.pr_ptr[static_cast<size_t>(f) + 1] - a.pr_ptr[static_cast<size_t>(f)]));
  }
  a.pr_row.resize(static_cast<size_t>(v.num_points));
  {
    std::vector<int64_t> fill(a.pr_ptr.begin(), a.pr_ptr.end() - 1);
    for (int64_t s = 0; s < v.num_points; ++s) {
      a.pr_row[static_cast<size_t>(fill[static_cast<size_t>(v.point_frame[s])]++)] = static_cast<int32_t>(s);
    }
  }

  // Footprints: dilated cells of each pose's points at the current poses.
  a.words_per_pose = (a.num_cells + 31) / 32;
  a.fp_cellbits.assign(static_cast<size_t>(N) * static_cast<size_t>(a.words_per_pose), 0U);
  std::vector<int> cell_stamp(static_cast<size_t>(a.num_cells), -1);
  std::vector<int> dil_stamp(static_cast<size_t>(a.num_cells), -1);
  std::vector<int> node_stamp(static_cast<size_t>(num_nodes), -1);
  std::vector<std::vector<int32_t>> footprint(static_cast<size_t>(N));
  std::vector<int> fp_bbox(4 * static_cast<size_t>(N), -1);
  std::vector<std::vector<int32_t>> dilated(static_cast<size_t>(N));
  for (int P = 0; P < N; ++P) {
    const double c = std::cos(poses[3 * P + 2]);
    const double sn = std::sin(poses[3 * P + 2]);
    std::vector<int> cells;
    for (int64_t k = a.pr_ptr[static_cast<size_t>(P)]; k < a.pr_ptr[static_cast<size_t>(P) + 1]; ++k) {
      const int32_t s = a.pr_row[static_cast<size_t>(k)];
      const double px = v.point_x[s];
      const double py = v.point_y[s];
      const double gx = c * px + (-sn) * py + poses[3 * P];
      const double gy = sn * px + c * py + poses[3 * P + 1];
      const double fx = (gx - v.grid.min_x) / v.grid.dx;
      const double fy = (gy - v.grid.min_y) / v.grid.dy;
      if (!(fx >= -1.0 && fy >= -1.0 && fx <= a.nx && fy <= a.ny)) {
        continue;  // well outside the domain: no Jacobian row
      }
      const int w = std::clamp(static_cast<int>(std::floor(fx)), 0, ncx - 1);
      const int h = std::clamp(static_cast<int>(std::floor(fy)), 0, ncy - 1);
      const int cell = h * ncx + w;
      if (cell_stamp[static_cast<size_t>(cell)] != P) {
        cell_stamp[static_cast<size_t>(cell)] = P;
        cells.push_back(cell);
      }
    }
    uint32_t* bits = a.fp_cellbits.data() + static_cast<size_t>(P) * static_cast<size_t>(a.words_per_pose);
    std::vector<int32_t>& fp = footprint[static_cast<size_t>(P)];
    for (const int cell : cells) {
      const int cw = cell % ncx;
      const int ch = cell / ncx;
      for (int dh = -margin; dh <= margin; ++dh) {
        for (int dw = -margin; dw <= margin; ++dw) {
          const int w = cw + dw;
          const int h = ch + dh;
          if (w < 0 || h < 0 || w >= ncx || h >= ncy) {
            continue;
          }
          const int d = h * ncx + w;
          if (dil_stamp[static_cast<size_t>(d)] == P) {
            continue;
          }
          dil_stamp[static_cast<size_t>(d)] = P;
          bits[d >> 5] |= 1U << (d & 31);
          dilated[static_cast<size_t>(P)].push_back(d);
          if (P == 0) {
            continue;  // frame 0 is fixed: no pose columns
          }
          const int corner[4] = {h * a.nx + w, h * a.nx + w + 1, (h + 1) * a.nx + w,
                                 (h + 1) * a.nx + w + 1};
          for (const int g : corner) {
            if (v.node_to_active[g] >= 0 && node_stamp[static_cast<size_t>(g)] != P) {
              node_stamp[static_cast<size_t>(g)] = P;
              fp.push_back(g);
            }
          }
        }
      }
    }
    std::sort(fp.begin(), fp.end());
    if (!fp.empty()) {
      int bw0 = a.nx, bw1 = -1, bh0 = a.ny, bh1 = -1;
      for (const int g : fp) {
        bw0 = std::min(bw0, g % a.nx);
        bw1 = std::max(bw1, g % a.nx);
        bh0 = std::min(bh0, g / a.nx);
        bh1 = std::max(bh1, g / a.nx);
      }
      fp_bbox[4 * static_cast<size_t>(P)] = bw0;
      fp_bbox[4 * static_cast<size_t>(P) + 1] = bw1;
      fp_bbox[4 * static_cast<size_t>(P) + 2] = bh0;
      fp_bbox[4 * static_cast<size_t>(P) + 3] = bh1;
    }
  }
  // Candidate frames per cell: every frame whose dilated cells contain it
  // (ascending), for the per-cell gather of the point-row partial sums.
  a.cp_ptr.assign(static_cast<size_t>(a.num_cells) + 1, 0);
  for (int P = 0; P < N; ++P) {
    for (const int d : dilated[static_cast<size_t>(P)]) ++a.cp_ptr[static_cast<size_t>(d) + 1];
  }
  for (int d = 0; d < a.num_cells; ++d) a.cp_ptr[static_cast<size_t>(d) + 1] += a.cp_ptr[static_cast<size_t>(d)];
  a.cp_pose.resize(static_cast<size_t>(a.cp_ptr[static_cast<size_t>(a.num_cells)]));
  {
    std::vector<int64_t> fill(a.cp_ptr.begin(), a.cp_ptr.end() - 1);
    for (int P = 0; P < N; ++P) {
      for (const int d : dilated[static_cast<size_t>(P)]) {
        a.cp_pose[static_cast<size_t>(fill[static_cast<size_t>(d)]++)] = P;
      }
    }
  }
  // Inverse footprint: poses per node (ascending).
  std::vector<int64_t> np_ptr(static_cast<size_t>(num_nodes) + 1, 0);
  for (int P = 1; P < N; ++P) {
    for (const int g : footprint[static_cast<size_t>(P)]) {
      ++np_ptr[static_cast<size_t>(g) + 1];
    }
  }
  for (int g = 0; g < num_nodes; ++g) {
    np_ptr[static_cast<size_t>(g) + 1] += np_ptr[static_cast<size_t>(g)];
  }
  std::vector<int32_t> np_pose(static_cast<size_t>(np_ptr[static_cast<size_t>(num_nodes)]));
  {
    std::vector<int64_t> fill(np_ptr.begin(), np_ptr.end() - 1);
    for (int P = 1; P < N; ++P) {
      for (const int g : footprint[static_cast<size_t>(P)]) {
        np_pose[static_cast<size_t>(fill[static_cast<size_t>(g)]++)] = P;
      }
    }
  }

  // Odometry/relation specs per pose and distinct pose pairs.
  a.po_ptr.assign(static_cast<size_t>(N) + 1, 0);
  for (int64_t o = 0; o < v.num_odom; ++o) {
    const int i = v.odom_i[o];
    const int j = v.odom_j[o];
    if (i < 0 || j < 0 || i >= N || j >= N || i == j) {
      throw std::invalid_argument("odometry/relation frames out of range");
    }
    if (i >= 1) ++a.po_ptr[static_cast<size_t>(i) + 1];
    ++a.po_ptr[static_cast<size_t>(j) + 1];
  }
  for (int f = 0; f < N; ++f) {
    a.po_ptr[static_cast<size_t>(f) + 1] += a.po_ptr[static_cast<size_t>(f)];
  }
  a.po_spec.resize(static_cast<size_t>(a.po_ptr[static_cast<size_t>(N)]));
  std::vector<std::vector<int32_t>> pose_partners(static_cast<size_t>(N));
  std::vector<std::pair<std::pair<int, int>, int>> pair_specs;
  {
    std::vector<int64_t> fill(a.po_ptr.begin(), a.po_ptr.end() - 1);
    for (int64_t o = 0; o < v.num_odom; ++o) {
      const int i = v.odom_i[o];
      const int j = v.odom_j[o];
      if (i >= 1) a.po_spec[static_cast<size_t>(fill[static_cast<size_t>(i)]++)] = static_cast<int32_t>(o);
      a.po_spec[static_cast<size_t>(fill[static_cast<size_t>(j)]++)] = static_cast<int32_t>(o);
      if (i >= 1) {
        pair_specs.push_back({{std::min(i, j), std::max(i, j)}, static_cast<int>(o)});
        pose_partners[static_cast<size_t>(i)].push_back(j);
        pose_partners[static_cast<size_t>(j)].push_back(i);
      }
    }
  }
  std::stable_sort(pair_specs.begin(), pair_specs.end(),
                   [](const auto& x, const auto& y) { return x.first < y.first; });

  // Nested dissection tree over the node grid.
  std::vector<NdNode> tree;
  std::function<int(int, int, int, int)> build = [&](int w0, int w1, int h0, int h1) -> int {
    const int idx = static_cast<int>(tree.size());
    tree.push_back(NdNode{w0, w1, h0, h1});
    const int ww = w1 - w0;
    const int hh = h1 - h0;
    if (ww <= 0 || hh <= 0 || ww * hh <= opt.leaf_nodes || (ww < 3 && hh < 3)) {
      return idx;
    }
    int left, right;
    if (ww >= hh) {
      const int m = (w0 + w1) / 2;
      tree[static_cast<size_t>(idx)].axis = 0;
      tree[static_cast<size_t>(idx)].m = m;
      left = build(w0, m, h0, h1);
      right = build(m + 1, w1, h0, h1);
    } else {
      const int m = (h0 + h1) / 2;
      tree[static_cast<size_t>(idx)].axis = 1;
      tree[static_cast<size_t>(idx)].m = m;
      left = build(w0, w1, h0, m);
      right = build(w0, w1, m + 1, h1);
    }
    tree[static_cast<size_t>(idx)].left = left;
    tree[static_cast<size_t>(idx)].right = right;
    return idx;
  };
  build(0, a.nx, 0, a.ny);
  // Pose placement.
  for (int P = 1; P < N; ++P) {
    int t = 0;
    const int* bb = &fp_bbox[4 * static_cast<size_t>(P)];
    if (bb[0] >= 0) {
      while (tree[static_cast<size_t>(t)].axis >= 0) {
        const NdNode& node = tree[static_cast<size_t>(t)];
        if (BoxContains(tree[static_cast<size_t>(node.left)], bb[0], bb[1], bb[2], bb[3])) {
          t = node.left;
        } else if (BoxContains(tree[static_cast<size_t>(node.right)], bb[0], bb[1], bb[2], bb[3])) {
          t = node.right;
        } else {
          break;
        }
      }
    }
    tree[static_cast<size_t>(t)].poses.push_back(P);
  }
  // Post-order elimination: each tree node's own columns form a supernode.
  a.perm.assign(static_cast<size_t>(a.n), -1);
  a.iperm.assign(static_cast<size_t>(a.n), -1);
  a.sn_first.push_back(0);
  int next = 0;
  std::function<void(int)> emit = [&](int t) {
    const NdNode& node = tree[static_cast<size_t>(t)];
    if (node.left >= 0) emit(node.left);
    if (node.right >= 0) emit(node.right);
    const int start = next;
    auto take_node = [&](int w, int h) {
      const int col = v.node_to_active[h * a.nx + w];
      if (col >= 0) {
        a.perm[static_cast<size_t>(col)] = next;
        a.iperm[static_cast<size_t>(next)] = col;
        ++next;
      }
    };
    if (node.axis < 0) {
      for (int h = node.h0; h < node.h1; ++h) {
        for (int w = node.w0; w < node.w1; ++w) take_node(w, h);
      }
    } else if (node.axis == 0) {
      for (int h = node.h0; h < node.h1; ++h) take_node(node.m, h);
    } else {
      for (int w = node.w0; w < node.w1; ++w) take_node(w, node.m);
    }
    for (const int P : node.poses) {
      const int pc = a.num_active + 3 * (P - 1);
      for (int k = 0; k < 3; ++k) {
        a.perm[static_cast<size_t>(pc + k)] = next;
        a.iperm[static_cast<size_t>(next)] = pc + k;
        ++next;
      }
    }
    if (next > start) {
      a.sn_first.push_back(next);
    }
  };
  emit(0);
  if (next != a.n) {
    throw std::logic_error("ordering did not cover every column");
  }
  a.num_sn = static_cast<int>(a.sn_first.size()) - 1;
  a.snode_of.assign(static_cast<size_t>(a.n), -1);
  for (int s = 0; s < a.num_sn; ++s) {
    for (int j = a.sn_first[static_cast<size_t>(s)]; j < a.sn_first[static_cast<size_t>(s) + 1]; ++j) {
      a.snode_of[static_cast<size_t>(j)] = s;
    }
  }

  // Neighbours of an original column in the A pattern (new indices).
  auto neighbors = [&](int col, const std::function<void(int)>& fn) {
    if (col < a.num_active) {
      const int g = active_to_node[static_cast<size_t>(col)];
      const int w = g % a.nx;
      const int h = g / a.nx;
      for (int dh = -1; dh <= 1; ++dh) {
        for (int dw = -1; dw <= 1; ++dw) {
          const int ww = w + dw;
          const int hh = h + dh;
          if ((dw == 0 && dh == 0) || ww < 0 || hh < 0 || ww >= a.nx || hh >= a.ny) continue;
          const int c2 = v.node_to_active[hh * a.nx + ww];
          if (c2 >= 0) fn(a.perm[static_cast<size_t>(c2)]);
        }
      }
      for (int64_t k = np_ptr[static_cast<size_t>(g)]; k < np_ptr[static_cast<size_t>(g) + 1]; ++k) {
        const int pc = a.num_active + 3 * (np_pose[static_cast<size_t>(k)] - 1);
        for (int t = 0; t < 3; ++t) fn(a.perm[static_cast<size_t>(pc + t)]);
      }
    } else {
      const int P = 1 + (col - a.num_active) / 3;
      const int pc = a.num_active + 3 * (P - 1);
      for (int t = 0; t < 3; ++t) {
        if (pc + t != col) fn(a.perm[static_cast<size_t>(pc + t)]);
      }
      for (const int g : footprint[static_cast<size_t>(P)]) {
        fn(a.perm[static_cast<size_t>(v.node_to_active[g])]);
      }
      for (const int Q : pose_partners[static_cast<size_t>(P)]) {
        if (Q < 1) continue;
        const int qc = a.num_active + 3 * (Q - 1);
        for (int t = 0; t < 3; ++t) fn(a.perm[static_cast<size_t>(qc + t)]);
      }
    }
  };

  // Block symbolic factorization.
  a.children.assign(static_cast<size_t>(a.num_sn), {});
  a.sn_parent.assign(static_cast<size_t>(a.num_sn), -1);
  a.sn_level.assign(static_cast<size_t>(a.num_sn), 0);
  a.sn_m.assign(static_cast<size_t>(a.num_sn), 0);
  a.sn_f.assign(static_cast<size_t>(a.num_sn), 0);
  a.sn_idx_ptr.assign(static_cast<size_t>(a.num_sn) + 1, 0);
  std::vector<std::vector<int32_t>> R(static_cast<size_t>(a.num_sn));
  std::vector<int> mark(static_cast<size_t>(a.n), -1);
  for (int s = 0; s < a.num_sn; ++s) {
    const int first = a.sn_first[static_cast<size_t>(s)];
    const int last = a.sn_first[static_cast<size_t>(s) + 1] - 1;
    std::vector<int32_t>& rs = R[static_cast<size_t>(s)];
    auto add = [&](int i) {
      if (i > last && mark[static_cast<size_t>(i)] != s) {
        mark[static_cast<size_t>(i)] = s;
        rs.push_back(i);
      }
    };
    for (int j = first; j <= last; ++j) {
      neighbors(a.iperm[static_cast<size_t>(j)], add);
    }
    for (const int c : a.children[static_cast<size_t>(s)]) {
      for (const int i : R[static_cast<size_t>(c)]) add(i);
    }
    std::sort(rs.begin(), rs.end());
    if (!rs.empty()) {
      const int p = a.snode_of[static_cast<size_t>(rs.front())];
      a.sn_parent[static_cast<size_t>(s)] = p;
      a.children[static_cast<size_t>(p)].push_back(s);
    }
    int lvl = 0;
    for (const int c : a.children[static_cast<size_t>(s)]) {
      lvl = std::max(lvl, a.sn_level[static_cast<size_t>(c)] + 1);
    }
    a.sn_level[static_cast<size_t>(s)] = lvl;
    a.num_levels = std::max(a.num_levels, lvl + 1);
  }
  a.panel_off.assign(static_cast<size_t>(a.num_sn), 0);
  a.upd_off.assign(static_cast<size_t>(a.num_sn), 0);
  for (int s = 0; s < a.num_sn; ++s) {
    const int ns = a.sn_first[static_cast<size_t>(s) + 1] - a.sn_first[static_cast<size_t>(s)];
    const int ms = static_cast<int>(R[static_cast<size_t>(s)].size());
    a.sn_m[static_cast<size_t>(s)] = ms;
    a.sn_f[static_cast<size_t>(s)] = ns + ms;
    a.sn_idx_ptr[static_cast<size_t>(s) + 1] = a.sn_idx_ptr[static_cast<size_t>(s)] + ns + ms;
    a.panel_off[static_cast<size_t>(s)] = a.panel_total;
    a.panel_total += static_cast<int64_t>(ns + ms) * ns;
    a.upd_off[static_cast<size_t>(s)] = a.upd_total;
    a.upd_total += static_cast<int64_t>(ms) * ms;
    a.factor_nnz += static_cast<int64_t>(ns) * (ns + 1) / 2 + static_cast<int64_t>(ms) * ns;
    for (int k = 0; k < ns; ++k) {
      const double cj = static_cast<double>(ns - 1 - k + ms);
      a.factor_flops += cj * cj + 3.0 * cj;
    }
  }
  a.sn_idx.resize(static_cast<size_t>(a.sn_idx_ptr[static_cast<size_t>(a.num_sn)]));
  for (int s = 0; s < a.num_sn; ++s) {
    int64_t p = a.sn_idx_ptr[static_cast<size_t>(s)];
    for (int j = a.sn_first[static_cast<size_t>(s)]; j < a.sn_first[static_cast<size_t>(s) + 1]; ++j) {
      a.sn_idx[static_cast<size_t>(p++)] = j;
    }
    for (const int i : R[static_cast<size_t>(s)]) {
      a.sn_idx[static_cast<size_t>(p++)] = i;
    }
  }
  // Child -> parent relative maps.
  a.rel_off.assign(static_cast<size_t>(a.num_sn) + 1, 0);
  a.rel_split.assign(static_cast<size_t>(a.num_sn), 0);
  for (int s = 0; s < a.num_sn; ++s) {
    a.rel_off[static_cast<size_t>(s) + 1] = a.rel_off[static_cast<size_t>(s)] + a.sn_m[static_cast<size_t>(s)];
  }
  a.rel.resize(static_cast<size_t>(a.rel_off[static_cast<size_t>(a.num_sn)]));
  for (int c = 0; c < a.num_sn; ++c) {
    const int p = a.sn_parent[static_cast<size_t>(c)];
    if (p < 0) continue;
    const int32_t* pb = a.sn_idx.data() + a.sn_idx_ptr[static_cast<size_t>(p)];
    const int32_t* pe = a.sn_idx.data() + a.sn_idx_ptr[static_cast<size_t>(p) + 1];
    const int pn = a.sn_first[static_cast<size_t>(p) + 1] - a.sn_first[static_cast<size_t>(p)];
    int split = 0;
    int64_t q = a.rel_off[static_cast<size_t>(c)];
    for (const int i : R[static_cast<size_t>(c)]) {
      const int32_t* it = std::lower_bound(pb, pe, i);
      if (it == pe || *it != i) {
        throw std::logic_error("child front row missing from parent front");
      }
      const auto pos = static_cast<int32_t>(it - pb);
      a.rel[static_cast<size_t>(q++)] = pos;
      if (pos < pn) ++split;
    }
    a.rel_split[static_cast<size_t>(c)] = split;
  }
  // Node-node slots and constant eikonal contributions.
  a.nn_off.assign(5 * static_cast<size_t>(num_nodes), -1);
  a.nn_eik.assign(5 * static_cast<size_t>(num_nodes), 0.0);
  const int dws[5] = {0, 1, -1, 0, 1};
  const int dhs[5] = {0, 0, 1, 1, 1};
  for (int g = 0; g < num_nodes; ++g) {
    const int cu = v.node_to_active[g];
    if (cu < 0) continue;
    const int w = g % a.nx;
    const int h = g / a.nx;
    for (int k = 0; k < 5; ++k) {
      const int ww = w + dws[k];
      const int hh = h + dhs[k];
      if (ww < 0 || hh < 0 || ww >= a.nx || hh >= a.ny) continue;
      const int cv = v.node_to_active[hh * a.nx + ww];
      if (cv < 0) continue;
      const int pu = a.perm[static_cast<size_t>(cu)];
      const int pv = a.perm[static_cast<size_t>(cv)];
      const int64_t off = PanelOffset(a, std::max(pu, pv), std::min(pu, pv));
      if (off < 0) throw std::logic_error("node-node slot missing from the factor");
      a.nn_off[5 * static_cast<size_t>(g) + k] = off;
    }
  }
  a.eik_at.assign(static_cast<size_t>(num_nodes), -1);
  for (int64_t e = 0; e < v.num_eikonal; ++e) {
    const int w = v.eik_w[e];
    const int h = v.eik_h[e];
    if (w < 0 || h < 0 || w + 1 >= a.nx || h + 1 >= a.ny) {
      throw std::invalid_argument("eikonal spec outside the grid");
    }
    const int g = h * a.nx + w;
    if (a.eik_at[static_cast<size_t>(g)] >= 0) {
      throw std::invalid_argument("more than one eikonal spec at a grid node");
    }
    a.eik_at[static_cast<size_t>(g)] = static_cast<int32_t>(e);
    const double sw = v.eik_sqrt_weight[e];
    const double j0 = sw * (v.eik_nx[e] / v.grid.dx + v.eik_ny[e] / v.grid.dy);
    const double j1 = sw * (-v.eik_nx[e] / v.grid.dx);
    const double j2 = sw * (-v.eik_ny[e] / v.grid.dy);
    const int g1 = g + 1;
    const int g2 = g + a.nx;
    const bool a0 = v.node_to_active[g] >= 0;
    const bool a1 = v.node_to_active[g1] >= 0;
    const bool a2 = v.node_to_active[g2] >= 0;
    if (a0) a.nn_eik[5 * static_cast<size_t>(g)] += j0 * j0;
    if (a1) a.nn_eik[5 * static_cast<size_t>(g1)] += j1 * j1;
    if (a2) a.nn_eik[5 * static_cast<size_t>(g2)] += j2 * j2;
    if (a0 && a1) a.nn_eik[5 * static_cast<size_t>(g) + 1] += j0 * j1;   // right
    if (a0 && a2) a.nn_eik[5 * static_cast<size_t>(g) + 3] += j0 * j2;   // up
    if (a1 && a2) a.nn_eik[5 * static_cast<size_t>(g1) + 2] += j1 * j2;  // up-left of g1
  }
  // Pose footprint slots.
  a.fp_ptr.assign(static_cast<size_t>(N) + 1, 0);
  for (int P = 1; P < N; ++P) {
    a.fp_ptr[static_cast<size_t>(P) + 1] = static_cast<int64_t>(footprint[static_cast<size_t>(P)].size());
  }
  for (int P = 0; P < N; ++P) {
    a.fp_ptr[static_cast<size_t>(P) + 1] += a.fp_ptr[static_cast<size_t>(P)];
  }
  a.fp_node.resize(static_cast<size_t>(a.fp_ptr[static_cast<size_t>(N)]));
  a.fp_off.resize(a.fp_node.size());
  a.fp_stride.resize(a.fp_node.size());
  for (int P = 1; P < N; ++P) {
    const int p0 = a.perm[static_cast<size_t>(a.num_active + 3 * (P - 1))];
    int64_t q = a.fp_ptr[static_cast<size_t>(P)];
    for (const int g : footprint[static_cast<size_t>(P)]) {
      const int pu = a.perm[static_cast<size_t>(v.node_to_active[g])];
      int64_t off;
      int stride;
      if (pu < p0) {
        off = PanelOffset(a, p0, pu);
        stride = 1;
        if (PanelOffset(a, p0 + 2, pu) != off + 2) throw std::logic_error("pose rows not adjacent");
      } else {
        off = PanelOffset(a, pu, p0);
        const int s = a.snode_of[static_cast<size_t>(p0)];
        stride = a.sn_f[static_cast<size_t>(s)];
        if (PanelOffset(a, pu, p0 + 2) != off + 2 * static_cast<int64_t>(stride)) {
          throw std::logic_error("pose columns not adjacent");
        }
      }
      if (off < 0) throw std::logic_error("footprint slot missing from the factor");
      a.fp_node[static_cast<size_t>(q)] = g;
      a.fp_off[static_cast<size_t>(q)] = off;
      a.fp_stride[static_cast<size_t>(q)] = stride;
      ++q;
    }
  }
  // Pose diagonal blocks.
  a.pd_off.assign(6 * static_cast<size_t>(N), -1);
  for (int P = 1; P < N; ++P) {
    const int p0 = a.perm[static_cast<size_t>(a.num_active + 3 * (P - 1))];
    const int ij[6][2] = {{0, 0}, {1, 0}, {1, 1}, {2, 0}, {2, 1}, {2, 2}};
    for (int k = 0; k < 6; ++k) {
      a.pd_off[6 * static_cast<size_t>(P) + k] = PanelOffset(a, p0 + ij[k][0], p0 + ij[k][1]);
    }
  }
  // Off-diagonal odometry/relation pairs.
  a.op_ptr.push_back(0);
  for (size_t k = 0; k < pair_specs.size();) {
    size_t e = k;
    while (e < pair_specs.size() && pair_specs[e].first == pair_specs[k].first) {
      a.op_spec.push_back(pair_specs[e].second);
      ++e;
    }
    const int P = pair_specs[k].first.first;
    const int Q = pair_specs[k].first.second;
    a.op_pose.push_back(P);
    a.op_pose.push_back(Q);
    a.op_ptr.push_back(static_cast<int64_t>(a.op_spec.size()));
    const int pP = a.perm[static_cast<size_t>(a.num_active + 3 * (P - 1))];
    const int pQ = a.perm[static_cast<size_t>(a.num_active + 3 * (Q - 1))];
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c) {
        const int i = pP + r;
        const int j = pQ + c;
        a.op_off.push_back(PanelOffset(a, std::max(i, j), std::min(i, j)));
      }
    }
    k = e;
  }
  // Level schedule.
  a.level_sn.assign(static_cast<size_t>(a.num_levels), {});
  a.level_max_nb.assign(static_cast<size_t>(a.num_levels), 0);
  a.level_max_f.assign(static_cast<size_t>(a.num_levels), 0);
  a.level_max_m.assign(static_cast<size_t>(a.num_levels), 0);
  a.level_max_children.assign(static_cast<size_t>(a.num_levels), 0);
  for (int s = 0; s < a.num_sn; ++s) {
    const auto L = static_cast<size_t>(a.sn_level[static_cast<size_t>(s)]);
    a.level_sn[L].push_back(s);
    const int ns = a.sn_first[static_cast<size_t>(s) + 1] - a.sn_first[static_cast<size_t>(s)];
    a.level_max_nb[L] = std::max(a.level_max_nb[L], ns);
    a.level_max_f[L] = std::max(a.level_max_f[L], a.sn_f[static_cast<size_t>(s)]);
    a.level_max_m[L] = std::max(a.level_max_m[L], a.sn_m[static_cast<size_t>(s)]);
    a.level_max_children[L] =
        std::max(a.level_max_children[L], static_cast<int32_t>(a.children[static_cast<size_t>(s)].size()));
  }
  // Pattern size (lower triangle): node-node slots, footprint triples, pose
  // diagonal blocks, odometry pair blocks.
  for (const int64_t off : a.nn_off) {
    if (off >= 0) ++a.pattern_nnz_lower;
  }
  a.pattern_nnz_lower += 3 * static_cast<int64_t>(a.fp_node.size()) + 6 * (N - 1) +
                         9 * static_cast<int64_t>(a.op_pose.size() / 2);
  return a;
}

std::vector<PatternEntry> LowerPattern(const Analysis& a, const sdfslam_problem_view& v) {
  std::vector<PatternEntry> out;
  out.reserve(static_cast<size_t>(a.pattern_nnz_lower));
  auto push = [&](int ni, int nj, int64_t off) {
    const int oi = a.iperm[static_cast<size_t>(ni)];
    const int oj = a.iperm[static_cast<size_t>(nj)];
    out.push_back({std::max(oi, oj), std::min(oi, oj), off});
  };
  const int num_nodes = a.nx * a.ny;
  const int dws[5] = {0, 1, -1, 0, 1};
  const int dhs[5] = {0, 0, 1, 1, 1};
  for (int g = 0; g < num_nodes; ++g) {
    for (int k = 0; k < 5; ++k) {
      const int64_t off = a.nn_off[5 * static_cast<size_t>(g) + k];
      if (off < 0) continue;
      const int g2 = (g / a.nx + dhs[k]) * a.nx + (g % a.nx + dws[k]);
      const int pu = a.perm[static_cast<size_t>(v.node_to_active[g])];
      const int pv = a.perm[static_cast<size_t>(v.node_to_active[g2])];
      push(std::max(pu, pv), std::min(pu, pv), off);
    }
  }
  for (int P = 1; P < a.num_poses; ++P) {
    const int p0 = a.perm[static_cast<size_t>(a.num_active + 3 * (P - 1))];
    for (int64_t q = a.fp_ptr[static_cast<size_t>(P)]; q < a.fp_ptr[static_cast<size_t>(P) + 1]; ++q) {
      const int pu = a.perm[static_cast<size_t>(v.node_to_active[a.fp_node[static_cast<size_t>(q)]])];
      for (int t = 0; t < 3; ++t) {
        push(std::max(pu, p0 + t), std::min(pu, p0 + t),
             a.fp_off[static_cast<size_t>(q)] + t * static_cast<int64_t>(a.fp_stride[static_cast<size_t>(q)]));
      }
    }
    const int ij[6][2] = {{0, 0}, {1, 0}, {1, 1}, {2, 0}, {2, 1}, {2, 2}};
    for (int k = 0; k < 6; ++k) {
      push(p0 + ij[k][0], p0 + ij[k][1], a.pd_off[6 * static_cast<size_t>(P) + k]);
    }
  }
  for (size_t k = 0; k < a.op_pose.size() / 2; ++k) {
    const int P = a.op_pose[2 * k];
    const int Q = a.op_pose[2 * k + 1];
    const int pP = a.perm[static_cast<size_t>(a.num_active + 3 * (P - 1))];
    const int pQ = a.perm[static_cast<size_t>(a.num_active + 3 * (Q - 1))];
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c) {
        const int i = pP + r;
        const int j = pQ + c;
        push(std::max(i, j), std::min(i, j), a.op_off[9 * k + 3 * static_cast<size_t>(r) + c]);
      }
    }
  }
  std::sort(out.begin(), out.end(), [](const PatternEntry& x, const PatternEntry& y) {
    return x.col != y.col ? x.col < y.col : x.row < y.row;
  });
  return out;
}

}  // namespace sdfslam
