// Symbolic analysis for the device backend (see analysis.hpp).
#include "analysis.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <exception>
#include <functional>
#include <stdexcept>
#include <string>
#include <thread>

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

/// Analyses running at once in this process (contexts analysed from
/// several host threads share the cores).
std::atomic<int> g_active_analyses{0};

struct ActiveAnalysis {
  ActiveAnalysis() { ++g_active_analyses; }
  ~ActiveAnalysis() { --g_active_analyses; }
  ActiveAnalysis(const ActiveAnalysis&) = delete;
  ActiveAnalysis& operator=(const ActiveAnalysis&) = delete;
};

/// Host threads for one analysis: the machine's cores (at most 16) divided
/// among the analyses running at once.
int AnalysisThreads() {
  const unsigned hc = std::thread::hardware_concurrency();
  const int cores = static_cast<int>(std::clamp(hc == 0 ? 1U : hc, 1U, 16U));
  return std::max(1, cores / std::max(1, g_active_analyses.load()));
}

/// Splits [0, n) into `chunks` contiguous ranges; chunk k is
/// [Bound(n, chunks, k), Bound(n, chunks, k + 1)).
int64_t Bound(int64_t n, int chunks, int k) { return n * k / chunks; }

/// Runs fn(k) for k in [0, chunks) on separate threads and rethrows the
/// first exception.
template <typename F>
void RunChunks(int chunks, F&& fn) {
  if (chunks <= 1) {
    if (chunks == 1) fn(0);
    return;
  }
  std::vector<std::thread> threads;
  std::vector<std::exception_ptr> errors(static_cast<size_t>(chunks));
  threads.reserve(static_cast<size_t>(chunks) - 1);
  for (int k = 1; k < chunks; ++k) {
    threads.emplace_back([&, k] {
      try {
        fn(k);
      } catch (...) {
        errors[static_cast<size_t>(k)] = std::current_exception();
      }
    });
  }
  try {
    fn(0);
  } catch (...) {
    errors[0] = std::current_exception();
  }
  for (std::thread& t : threads) t.join();
  for (const std::exception_ptr& e : errors) {
    if (e) std::rethrow_exception(e);
  }
}

/// Parallel loop over [0, n): fn(i) for every i, in contiguous chunks.
template <typename F>
void ParallelFor(int64_t n, int threads, int64_t grain, F&& fn) {
  const int chunks = static_cast<int>(std::min<int64_t>(threads, std::max<int64_t>(n / grain, 1)));
  RunChunks(chunks, [&](int k) {
    for (int64_t i = Bound(n, chunks, k); i < Bound(n, chunks, k + 1); ++i) fn(i);
  });
}

/// CSR inversion: for sources 0..num_src-1 (in chunks) with target lists
/// lists[src], builds ptr (num_tgt + 1) and out such that each target's
/// sources are ascending. Only sources in [src_begin, num_src) take part.
void InvertLists(const std::vector<std::vector<int32_t>>& lists, int src_begin, int num_tgt,
                 int threads, std::vector<int64_t>& ptr, std::vector<int32_t>& out,
                 std::vector<int32_t>* pos = nullptr) {
  const int num_src = static_cast<int>(lists.size());
  const int span = std::max(num_src - src_begin, 0);
  const int chunks = std::max(1, std::min(threads, span / 16));
  std::vector<int64_t> cnt(static_cast<size_t>(chunks) * static_cast<size_t>(num_tgt), 0);
  RunChunks(chunks, [&](int k) {
    int64_t* c = cnt.data() + static_cast<size_t>(k) * static_cast<size_t>(num_tgt);
    for (int64_t s = src_begin + Bound(span, chunks, k); s < src_begin + Bound(span, chunks, k + 1); ++s) {
      for (const int32_t t : lists[static_cast<size_t>(s)]) ++c[t];
    }
  });
  ptr.assign(static_cast<size_t>(num_tgt) + 1, 0);
  int64_t run = 0;
  for (int t = 0; t < num_tgt; ++t) {
    for (int k = 0; k < chunks; ++k) {
      int64_t& c = cnt[static_cast<size_t>(k) * static_cast<size_t>(num_tgt) + static_cast<size_t>(t)];
      const int64_t n = c;
      c = run;  // becomes chunk k's write cursor for target t
      run += n;
    }
    ptr[static_cast<size_t>(t) + 1] = run;
  }
  out.resize(static_cast<size_t>(run));
  if (pos != nullptr) pos->resize(static_cast<size_t>(run));
  RunChunks(chunks, [&](int k) {
    int64_t* c = cnt.data() + static_cast<size_t>(k) * static_cast<size_t>(num_tgt);
    for (int64_t s = src_begin + Bound(span, chunks, k); s < src_begin + Bound(span, chunks, k + 1); ++s) {
      const std::vector<int32_t>& list = lists[static_cast<size_t>(s)];
      for (size_t i = 0; i < list.size(); ++i) {
        const int64_t at = c[list[i]]++;
        out[static_cast<size_t>(at)] = static_cast<int32_t>(s);
        if (pos != nullptr) (*pos)[static_cast<size_t>(at)] = static_cast<int32_t>(i);
      }
    }
  });
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
  auto margin_of = [&](int P) {
    return opt.pose_margin.empty() ? margin
                                   : std::max(1, opt.pose_margin[static_cast<size_t>(P)]);
  };
  if (!opt.pose_margin.empty() && static_cast<int>(opt.pose_margin.size()) != N) {
    throw std::invalid_argument("per-pose margins do not match the pose count");
  }
  const ActiveAnalysis active;
  const int threads = AnalysisThreads();

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

  // Footprints: dilated cells of each pose's points at the current poses
  // (pose P dilates by its own margin), one chunk of poses per thread.
  a.words_per_pose = (a.num_cells + 31) / 32;
  a.fp_cellbits.assign(static_cast<size_t>(N) * static_cast<size_t>(a.words_per_pose), 0U);
  std::vector<std::vector<int32_t>> footprint(static_cast<size_t>(N));
  std::vector<int> fp_bbox(4 * static_cast<size_t>(N), -1);
  std::vector<std::vector<int32_t>> dilated(static_cast<size_t>(N));
  {
    const int chunks = std::max(1, std::min(threads, N / 8));
    RunChunks(chunks, [&](int k) {
      std::vector<int> cell_stamp(static_cast<size_t>(a.num_cells), -1);
      std::vector<int> cells;
      std::vector<uint64_t> occ, row, dil;
      for (int P = static_cast<int>(Bound(N, chunks, k)); P < Bound(N, chunks, k + 1); ++P) {
        const int pm = margin_of(P);
        const double c = std::cos(poses[3 * P + 2]);
        const double sn = std::sin(poses[3 * P + 2]);
        cells.clear();
        for (int64_t q = a.pr_ptr[static_cast<size_t>(P)]; q < a.pr_ptr[static_cast<size_t>(P) + 1]; ++q) {
          const int32_t s = a.pr_row[static_cast<size_t>(q)];
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
        std::vector<int32_t>& dl = dilated[static_cast<size_t>(P)];
        if (cells.empty()) continue;
        // Occupied cells as row bitmaps over the dilated bounding box, then
        // dilated by pm cells along w (word shifts) and along h (row ORs).
        int cw0 = ncx, cw1 = -1, ch0 = ncy, ch1 = -1;
        for (const int cell : cells) {
          cw0 = std::min(cw0, cell % ncx);
          cw1 = std::max(cw1, cell % ncx);
          ch0 = std::min(ch0, cell / ncx);
          ch1 = std::max(ch1, cell / ncx);
        }
        const int W0 = std::max(cw0 - pm, 0);
        const int W1 = std::min(cw1 + pm, ncx - 1);
        const int H0 = std::max(ch0 - pm, 0);
        const int H1 = std::min(ch1 + pm, ncy - 1);
        const int lw = W1 - W0 + 2;  // one spare column for the node corners
        const int lh = H1 - H0 + 2;  // one spare row for the node corners
        const int nw = (lw + 63) / 64;
        occ.assign(static_cast<size_t>(lh) * nw, 0ULL);
        row.assign(static_cast<size_t>(lh) * nw, 0ULL);
        for (const int cell : cells) {
          const int w = cell % ncx - W0;
          const int h = cell / ncx - H0;
          occ[static_cast<size_t>(h) * nw + (w >> 6)] |= 1ULL << (w & 63);
        }
        for (int h = 0; h < lh - 1; ++h) {
          const uint64_t* x = occ.data() + static_cast<size_t>(h) * nw;
          uint64_t* y = row.data() + static_cast<size_t>(h) * nw;
          for (int i = 0; i < nw; ++i) y[i] = x[i];
          for (int k = 1; k <= pm; ++k) {
            for (int i = 0; i < nw; ++i) {
              const uint64_t up = (x[i] << k) | (i > 0 ? x[i - 1] >> (64 - k) : 0ULL);
              const uint64_t dn = (x[i] >> k) | (i + 1 < nw ? x[i + 1] << (64 - k) : 0ULL);
              y[i] |= up | dn;
            }
          }
          // Cell columns are local 0 .. lw - 2: clear what the shifts moved
          // past the grid's last cell column (and the spare column).
          for (int i = 0; i < nw; ++i) {
            const int valid = lw - 1 - 64 * i;
            if (valid < 64) y[i] &= valid <= 0 ? 0ULL : (1ULL << valid) - 1ULL;
          }
        }
        dil.assign(static_cast<size_t>(lh) * nw, 0ULL);
        for (int h = 0; h < lh - 1; ++h) {
          uint64_t* y = dil.data() + static_cast<size_t>(h) * nw;
          for (int dh = -pm; dh <= pm; ++dh) {
            const int hh = h + dh;
            if (hh < 0 || hh >= lh - 1) continue;
            const uint64_t* x = row.data() + static_cast<size_t>(hh) * nw;
            for (int i = 0; i < nw; ++i) y[i] |= x[i];
          }
          for (int i = 0; i < nw; ++i) {
            uint64_t m = y[i];
            while (m != 0ULL) {
              const int b = __builtin_ctzll(m);
              m &= m - 1;
              const int d = (H0 + h) * ncx + W0 + 64 * i + b;
              bits[d >> 5] |= 1U << (d & 31);
              dl.push_back(d);
            }
          }
        }
        if (P == 0) continue;  // frame 0 is fixed: no pose columns
        // Nodes of the dilated cells: node (w, h) is a corner of cells
        // (w-1..w, h-1..h); node rows are ascending, so fp comes out sorted.
        int bw0 = a.nx, bw1 = -1, bh0 = a.ny, bh1 = -1;
        for (int h = 0; h < lh; ++h) {
          const uint64_t* y0 = dil.data() + static_cast<size_t>(h) * nw;
          const uint64_t* y1 = h > 0 ? dil.data() + static_cast<size_t>(h - 1) * nw : nullptr;
          for (int i = 0; i < nw; ++i) {
            const uint64_t c0 = y0[i] | (y1 != nullptr ? y1[i] : 0ULL);
            const uint64_t cm = i > 0 ? (y0[i - 1] | (y1 != nullptr ? y1[i - 1] : 0ULL)) : 0ULL;
            uint64_t m = c0 | (c0 << 1) | (cm >> 63);
            while (m != 0ULL) {
              const int b = __builtin_ctzll(m);
              m &= m - 1;
              const int w = W0 + 64 * i + b;
              const int hg = H0 + h;
              if (w >= a.nx || hg >= a.ny) continue;
              const int g = hg * a.nx + w;
              if (v.node_to_active[g] < 0) continue;
              fp.push_back(g);
              bw0 = std::min(bw0, w);
              bw1 = std::max(bw1, w);
              bh0 = std::min(bh0, hg);
              bh1 = std::max(bh1, hg);
            }
          }
        }
        if (!fp.empty()) {
          fp_bbox[4 * static_cast<size_t>(P)] = bw0;
          fp_bbox[4 * static_cast<size_t>(P) + 1] = bw1;
          fp_bbox[4 * static_cast<size_t>(P) + 2] = bh0;
          fp_bbox[4 * static_cast<size_t>(P) + 3] = bh1;
        }
      }
    });
  }
  // Candidate frames per cell: every frame whose dilated cells contain it
  // (ascending), for the per-cell gather of the point-row partial sums.
  InvertLists(dilated, 0, a.num_cells, threads, a.cp_ptr, a.cp_pose, &a.cp_rank);
  a.dl_ptr.assign(static_cast<size_t>(N) + 1, 0);
  for (int P = 0; P < N; ++P) {
    a.dl_ptr[static_cast<size_t>(P) + 1] = a.dl_ptr[static_cast<size_t>(P)] +
                                           static_cast<int64_t>(dilated[static_cast<size_t>(P)].size());
  }
  a.fp_cellrank.resize(a.fp_cellbits.size());
  ParallelFor(N, threads, 16, [&](int64_t P) {
    const size_t w0 = static_cast<size_t>(P) * static_cast<size_t>(a.words_per_pose);
    int32_t run = 0;
    for (int w = 0; w < a.words_per_pose; ++w) {
      a.fp_cellrank[w0 + static_cast<size_t>(w)] = run;
      run += __builtin_popcount(a.fp_cellbits[w0 + static_cast<size_t>(w)]);
    }
  });
  // Inverse footprint: poses >= 1 per node (ascending).
  std::vector<int64_t> np_ptr;
  std::vector<int32_t> np_pose;
  InvertLists(footprint, 1, num_nodes, threads, np_ptr, np_pose);
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
    tree.push_back(NdNode{w0, w1, h0, h1, -1, -1, -1, 0, {}});
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
  auto neighbors = [&](int col, auto&& fn) {
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
  // A supernode's parent holds the first row of its front beyond it. The
  // block elimination tree that this defines is the elimination tree of the
  // quotient graph (one vertex per supernode, an edge where A has an entry
  // between two supernodes), so it is known before the fronts are: Liu's
  // algorithm with path compression over the quotient graph's lower
  // adjacency. Odometry and relation partners of poses placed in disjoint
  // dissection subtrees make this tree differ from the dissection tree, so
  // the heights come from it. Supernodes of one height are independent and
  // run in parallel (a mark array per thread); the children lists are
  // appended after each height and sorted at the end, the order the serial
  // loop gives.
  std::vector<std::vector<int>> marks(static_cast<size_t>(threads));
  std::vector<std::vector<int32_t>> q_upper(static_cast<size_t>(a.num_sn));
  {
    const int chunks = std::max(1, std::min(threads, a.num_sn / 16));
    RunChunks(chunks, [&](int k) {
      std::vector<int>& mark = marks[static_cast<size_t>(k)];
      if (mark.empty()) mark.assign(static_cast<size_t>(a.n), -1);
      for (int64_t s = Bound(a.num_sn, chunks, k); s < Bound(a.num_sn, chunks, k + 1); ++s) {
        const int first = a.sn_first[static_cast<size_t>(s)];
        const int last = a.sn_first[static_cast<size_t>(s) + 1] - 1;
        std::vector<int32_t>& up = q_upper[static_cast<size_t>(s)];
        auto add = [&](int i) {
          if (i <= last) return;
          const int t = a.snode_of[static_cast<size_t>(i)];
          if (mark[static_cast<size_t>(t)] != static_cast<int>(s)) {
            mark[static_cast<size_t>(t)] = static_cast<int>(s);
            up.push_back(t);
          }
        };
        for (int j = first; j <= last; ++j) {
          const int col = a.iperm[static_cast<size_t>(j)];
          if (col >= a.num_active && (col - a.num_active) % 3 != 0) continue;
          neighbors(col, add);
        }
      }
    });
    for (std::vector<int>& mark : marks) std::fill(mark.begin(), mark.end(), -1);
  }
  std::vector<int> q_parent(static_cast<size_t>(a.num_sn), -1);
  std::vector<int> sn_height(static_cast<size_t>(a.num_sn), 0);
  {
    std::vector<int64_t> lo_ptr(static_cast<size_t>(a.num_sn) + 1, 0);
    for (int s = 0; s < a.num_sn; ++s) {
      for (const int t : q_upper[static_cast<size_t>(s)]) ++lo_ptr[static_cast<size_t>(t) + 1];
    }
    for (int s = 0; s < a.num_sn; ++s) lo_ptr[static_cast<size_t>(s) + 1] += lo_ptr[static_cast<size_t>(s)];
    std::vector<int32_t> lo(static_cast<size_t>(lo_ptr[static_cast<size_t>(a.num_sn)]));
    std::vector<int64_t> fill(lo_ptr.begin(), lo_ptr.end() - 1);
    for (int s = 0; s < a.num_sn; ++s) {
      for (const int t : q_upper[static_cast<size_t>(s)]) lo[static_cast<size_t>(fill[static_cast<size_t>(t)]++)] = s;
    }
    std::vector<int> ancestor(static_cast<size_t>(a.num_sn), -1);
    for (int s = 0; s < a.num_sn; ++s) {
      for (int64_t e = lo_ptr[static_cast<size_t>(s)]; e < lo_ptr[static_cast<size_t>(s) + 1]; ++e) {
        int r = lo[static_cast<size_t>(e)];
        while (ancestor[static_cast<size_t>(r)] != -1 && ancestor[static_cast<size_t>(r)] != s) {
          const int up = ancestor[static_cast<size_t>(r)];
          ancestor[static_cast<size_t>(r)] = s;
          r = up;
        }
        if (ancestor[static_cast<size_t>(r)] == -1) {
          ancestor[static_cast<size_t>(r)] = s;
          q_parent[static_cast<size_t>(r)] = s;
        }
      }
    }
    for (int s = 0; s < a.num_sn; ++s) {
      const int p = q_parent[static_cast<size_t>(s)];
      if (p >= 0) {
        sn_height[static_cast<size_t>(p)] =
            std::max(sn_height[static_cast<size_t>(p)], sn_height[static_cast<size_t>(s)] + 1);
      }
    }
  }
  int max_height = 0;
  for (const int h : sn_height) max_height = std::max(max_height, h);
  std::vector<std::vector<int>> by_height(static_cast<size_t>(max_height) + 1);
  for (int s = 0; s < a.num_sn; ++s) by_height[static_cast<size_t>(sn_height[static_cast<size_t>(s)])].push_back(s);
  for (const std::vector<int>& list : by_height) {
    const int count = static_cast<int>(list.size());
    const int chunks = std::max(1, std::min(threads, count / 4));
    RunChunks(chunks, [&](int k) {
      std::vector<int>& mark = marks[static_cast<size_t>(k)];
      if (mark.empty()) mark.assign(static_cast<size_t>(a.n), -1);
      for (int64_t idx = Bound(count, chunks, k); idx < Bound(count, chunks, k + 1); ++idx) {
        const int s = list[static_cast<size_t>(idx)];
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
          const int col = a.iperm[static_cast<size_t>(j)];
          // The three columns of a pose share one supernode and one
          // neighbour set beyond it: visit the first only.
          if (col >= a.num_active && (col - a.num_active) % 3 != 0) continue;
          neighbors(col, add);
        }
        int lvl = 0;
        for (const int c : a.children[static_cast<size_t>(s)]) {
          for (const int i : R[static_cast<size_t>(c)]) add(i);
          lvl = std::max(lvl, a.sn_level[static_cast<size_t>(c)] + 1);
        }
        std::sort(rs.begin(), rs.end());
        if (!rs.empty()) a.sn_parent[static_cast<size_t>(s)] = a.snode_of[static_cast<size_t>(rs.front())];
        a.sn_level[static_cast<size_t>(s)] = lvl;
      }
    });
    for (const int s : list) {
      const int p = a.sn_parent[static_cast<size_t>(s)];
      if (p != q_parent[static_cast<size_t>(s)]) {
        throw std::logic_error("supernode parent differs from the quotient-graph elimination tree");
      }
      if (p >= 0) a.children[static_cast<size_t>(p)].push_back(s);
      a.num_levels = std::max(a.num_levels, a.sn_level[static_cast<size_t>(s)] + 1);
    }
  }
  for (std::vector<int32_t>& ch : a.children) std::sort(ch.begin(), ch.end());
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
  ParallelFor(a.num_sn, threads, 32, [&](int64_t ci) {
    const int c = static_cast<int>(ci);
    const int p = a.sn_parent[static_cast<size_t>(c)];
    if (p < 0) return;
    const int32_t* pb = a.sn_idx.data() + a.sn_idx_ptr[static_cast<size_t>(p)];
    const int32_t* pe = a.sn_idx.data() + a.sn_idx_ptr[static_cast<size_t>(p) + 1];
    const int pn = a.sn_first[static_cast<size_t>(p) + 1] - a.sn_first[static_cast<size_t>(p)];
    int split = 0;
    int64_t q = a.rel_off[static_cast<size_t>(c)];
    // Both lists ascend: one merge walk.
    const int32_t* it = pb;
    for (const int i : R[static_cast<size_t>(c)]) {
      while (it != pe && *it < i) ++it;
      if (it == pe || *it != i) {
        throw std::logic_error("child front row missing from parent front");
      }
      const auto pos = static_cast<int32_t>(it - pb);
      a.rel[static_cast<size_t>(q++)] = pos;
      if (pos < pn) ++split;
    }
    a.rel_split[static_cast<size_t>(c)] = split;
  });
  // Node-node slots.
  a.nn_off.assign(5 * static_cast<size_t>(num_nodes), -1);
  const int dws[5] = {0, 1, -1, 0, 1};
  const int dhs[5] = {0, 0, 1, 1, 1};
  // By supernode of the slot's column (the smaller position): a dense map
  // of the supernode's front positions, filled once per supernode, replaces
  // a binary search per slot.
  {
    std::vector<std::vector<int32_t>> cols_of_sn(static_cast<size_t>(a.num_sn));
    for (int g = 0; g < num_nodes; ++g) {
      const int cu = v.node_to_active[g];
      if (cu < 0) continue;
      const int pu = a.perm[static_cast<size_t>(cu)];
      const int w = g % a.nx;
      const int h = g / a.nx;
      for (int k = 0; k < 5; ++k) {
        const int ww = w + dws[k];
        const int hh = h + dhs[k];
        if (ww < 0 || hh < 0 || ww >= a.nx || hh >= a.ny) continue;
        const int cv = v.node_to_active[hh * a.nx + ww];
        if (cv < 0) continue;
        const int pv = a.perm[static_cast<size_t>(cv)];
        cols_of_sn[static_cast<size_t>(a.snode_of[static_cast<size_t>(std::min(pu, pv))])].push_back(5 * g + k);
      }
    }
    const int chunks = std::max(1, std::min(threads, a.num_sn / 64));
    RunChunks(chunks, [&](int t) {
      std::vector<int32_t> pos_in(static_cast<size_t>(a.n), -1);
      for (int64_t sn = Bound(a.num_sn, chunks, t); sn < Bound(a.num_sn, chunks, t + 1); ++sn) {
        const std::vector<int32_t>& slots = cols_of_sn[static_cast<size_t>(sn)];
        if (slots.empty()) continue;
        const int32_t* idx = a.sn_idx.data() + a.sn_idx_ptr[static_cast<size_t>(sn)];
        const int f = a.sn_f[static_cast<size_t>(sn)];
        for (int r = 0; r < f; ++r) pos_in[static_cast<size_t>(idx[r])] = r;
        for (const int32_t q : slots) {
          const int g = q / 5;
          const int k = q % 5;
          const int cu = v.node_to_active[g];
          const int cv = v.node_to_active[(g / a.nx + dhs[k]) * a.nx + g % a.nx + dws[k]];
          const int pu = a.perm[static_cast<size_t>(cu)];
          const int pv = a.perm[static_cast<size_t>(cv)];
          const int j = std::min(pu, pv);
          const int32_t row = pos_in[static_cast<size_t>(std::max(pu, pv))];
          if (row < 0) throw std::logic_error("node-node slot missing from the factor");
          a.nn_off[static_cast<size_t>(q)] = a.panel_off[static_cast<size_t>(sn)] +
                                             static_cast<int64_t>(j - a.sn_first[static_cast<size_t>(sn)]) * f + row;
        }
        for (int r = 0; r < f; ++r) pos_in[static_cast<size_t>(idx[r])] = -1;
      }
    });
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
  {
    // Offsets without a search per entry: the position of the pose's rows
    // inside a node supernode's front is looked up once per (pose,
    // supernode), and the pose supernode's front positions sit in a dense
    // per-thread map, rebuilt when the supernode changes.
    const int chunks = std::max(1, std::min(threads, (N - 1) / 8));
    RunChunks(chunks, [&](int k) {
      std::vector<int32_t> pos_in(static_cast<size_t>(a.n), -1);
      int cur_sp = -1;
      constexpr int kMemo = 256;
      int memo_s[kMemo];
      int32_t memo_pos[kMemo];
      for (int P = 1 + static_cast<int>(Bound(N - 1, chunks, k)); P < 1 + Bound(N - 1, chunks, k + 1); ++P) {
        const int p0 = a.perm[static_cast<size_t>(a.num_active + 3 * (P - 1))];
        const int sp = a.snode_of[static_cast<size_t>(p0)];
        const int32_t* sp_idx = a.sn_idx.data() + a.sn_idx_ptr[static_cast<size_t>(sp)];
        const int sp_f = a.sn_f[static_cast<size_t>(sp)];
        if (sp != cur_sp) {
          if (cur_sp >= 0) {
            const int32_t* old = a.sn_idx.data() + a.sn_idx_ptr[static_cast<size_t>(cur_sp)];
            for (int r = 0; r < a.sn_f[static_cast<size_t>(cur_sp)]; ++r) pos_in[static_cast<size_t>(old[r])] = -1;
          }
          for (int r = 0; r < sp_f; ++r) pos_in[static_cast<size_t>(sp_idx[r])] = r;
          cur_sp = sp;
        }
        std::fill(memo_s, memo_s + kMemo, -1);
        const int64_t p_col = a.panel_off[static_cast<size_t>(sp)] +
                              static_cast<int64_t>(p0 - a.sn_first[static_cast<size_t>(sp)]) * sp_f;
        int64_t q = a.fp_ptr[static_cast<size_t>(P)];
        for (const int g : footprint[static_cast<size_t>(P)]) {
          const int pu = a.perm[static_cast<size_t>(v.node_to_active[g])];
          int64_t off;
          int stride;
          if (pu < p0) {
            // Rows p0..p0+2 of column pu, in the node's supernode.
            const int su = a.snode_of[static_cast<size_t>(pu)];
            const int slot = su & (kMemo - 1);
            const int32_t* su_idx = a.sn_idx.data() + a.sn_idx_ptr[static_cast<size_t>(su)];
            const int su_f = a.sn_f[static_cast<size_t>(su)];
            if (memo_s[slot] != su) {
              const int32_t* it = std::lower_bound(su_idx, su_idx + su_f, p0);
              memo_s[slot] = su;
              memo_pos[slot] = (it == su_idx + su_f || *it != p0) ? -1 : static_cast<int32_t>(it - su_idx);
            }
            const int32_t r = memo_pos[slot];
            if (r < 0) throw std::logic_error("footprint slot missing from the factor");
            if (r + 2 >= su_f || su_idx[r + 2] != p0 + 2) throw std::logic_error("pose rows not adjacent");
            off = a.panel_off[static_cast<size_t>(su)] +
                  static_cast<int64_t>(pu - a.sn_first[static_cast<size_t>(su)]) * su_f + r;
            stride = 1;
          } else {
            // Row pu of columns p0..p0+2, in the pose's supernode.
            const int32_t r = pos_in[static_cast<size_t>(pu)];
            if (r < 0) throw std::logic_error("footprint slot missing from the factor");
            off = p_col + r;
            stride = sp_f;
          }
          a.fp_node[static_cast<size_t>(q)] = g;
          a.fp_off[static_cast<size_t>(q)] = off;
          a.fp_stride[static_cast<size_t>(q)] = stride;
          ++q;
        }
      }
    });
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
  for (int L = 0; L < a.num_levels; ++L) {
    if (!a.level_sn[static_cast<size_t>(L)].empty()) {
      a.panel_step_launches += 1;
      // The last step's launch only moves the factored diagonal blocks
      // into the panels; the update-matrix launch does that instead where
      // every front of the level has an update matrix.
      bool all_update = true;
      for (const int s : a.level_sn[static_cast<size_t>(L)]) all_update = all_update && a.sn_m[static_cast<size_t>(s)] > 0;
      a.gemm_launches += (a.level_max_nb[static_cast<size_t>(L)] + 63) / 64 - (all_update ? 1 : 0) +
                         (a.level_max_m[static_cast<size_t>(L)] > 0 ? 1 : 0);
    }
  }
  for (int s = 0; s < a.num_sn; ++s) {
    const int ns = a.sn_first[static_cast<size_t>(s) + 1] - a.sn_first[static_cast<size_t>(s)];
    const int f = a.sn_f[static_cast<size_t>(s)];
    for (int c0 = 0; c0 < ns; c0 += 64) {
      const double nb = std::min(64, ns - c0);
      const double rows = f - c0 - nb;
      // Diagonal block LDL^T and the TRSM of the rows below it: the first
      // step's in the panel-step launch, every later step's in the previous
      // step's trailing update (counted once; its copies are not work).
      const double step = 2.0 / 3.0 * nb * nb * nb + rows * nb * nb;
      if (c0 == 0) {
        a.panel_step_flops += step;
      } else {
        a.gemm_flops += step;
      }
      // Trailing update: the lower part of the rows and columns beyond the block.
      const double big_n = std::max(0.0, ns - c0 - nb);
      const double big_m = f - c0 - nb;
      a.gemm_flops += 2.0 * nb * (big_n * (big_n + 1.0) / 2.0 + (big_m - big_n) * big_n);
    }
    const double m = a.sn_m[static_cast<size_t>(s)];
    a.gemm_flops += 2.0 * ns * m * (m + 1.0) / 2.0;
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

void ReferenceFactorCounts(const Analysis& a, int64_t rows, const int32_t* jcols, int64_t* nnz, double* flops) {
  const int n = a.n;
  // Lower neighbours (in elimination positions) of every column.
  std::vector<std::vector<int32_t>> below(static_cast<size_t>(n));
  for (int64_t r = 0; r < rows; ++r) {
    int32_t q[7];
    int m = 0;
    for (int t = 0; t < 7; ++t) {
      const int32_t c = jcols[7 * r + t];
      if (c >= 0) q[m++] = a.perm[static_cast<size_t>(c)];
    }
    for (int x = 0; x < m; ++x) {
      for (int y = 0; y < m; ++y) {
        if (q[y] < q[x]) below[static_cast<size_t>(q[x])].push_back(q[y]);
      }
    }
  }
  // Elimination tree and column counts (row k's pattern is the union of the
  // tree paths from its entries up to k).
  std::vector<int> parent(static_cast<size_t>(n), -1);
  std::vector<int> flag(static_cast<size_t>(n), -1);
  std::vector<int64_t> count(static_cast<size_t>(n), 0);
  for (int k = 0; k < n; ++k) {
    flag[static_cast<size_t>(k)] = k;
    for (int i : below[static_cast<size_t>(k)]) {
      for (; flag[static_cast<size_t>(i)] != k; i = parent[static_cast<size_t>(i)]) {
        if (parent[static_cast<size_t>(i)] == -1) parent[static_cast<size_t>(i)] = k;
        ++count[static_cast<size_t>(i)];
        flag[static_cast<size_t>(i)] = k;
      }
    }
  }
  int64_t total = 0;
  double f = 0.0;
  for (int k = 0; k < n; ++k) {
    const double c = static_cast<double>(count[static_cast<size_t>(k)]);
    total += count[static_cast<size_t>(k)];
    f += c * c + 3.0 * c;
  }
  *nnz = total;
  *flops = f;
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
