// Host-only check of the symbolic analysis (no GPU): for each named workload
// the supernodal fronts, parents and levels that Analyze computes (threaded,
// one elimination-tree height at a time) are compared with a serial block
// symbolic factorization written here from the lower pattern alone:
// R_s = (rows of A beyond supernode s in its columns) united with the R_c of
// its children, parent = supernode of min(R_s), level = 1 + max child level.
//
// Usage: analysis_check <seed> <workload>...   prints one line per workload
// and exits non-zero on the first mismatch.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <string>
#include <vector>

#include "analysis.hpp"
#include "problem.hpp"
#include "scene.hpp"

namespace {

bool Check(const std::string& name, uint64_t seed) {
  const sdf_slam::Workload wl = sdf_slam::MakeWorkload(name, seed);
  const sdf_slam::Problem problem(wl.map, wl.initial, wl.scans, wl.odometry, wl.problem, wl.relations);
  const sdfslam_problem_view v = problem.View();
  std::vector<double> map_values;
  std::vector<double> poses;
  problem.GetState(map_values, poses);
  const sdfslam::Analysis a = sdfslam::Analyze(v, map_values.data(), poses.data(), sdfslam::AnalysisOptions{});
  const std::vector<sdfslam::PatternEntry> pat = sdfslam::LowerPattern(a, v);
  // Lower pattern in elimination order, by column.
  std::vector<std::vector<int>> below(static_cast<size_t>(a.n));
  for (const sdfslam::PatternEntry& e : pat) {
    const int i = a.perm[static_cast<size_t>(e.row)];
    const int j = a.perm[static_cast<size_t>(e.col)];
    below[static_cast<size_t>(std::min(i, j))].push_back(std::max(i, j));
  }
  std::vector<std::vector<int>> R(static_cast<size_t>(a.num_sn));
  std::vector<std::vector<int>> kids(static_cast<size_t>(a.num_sn));
  std::vector<int> level(static_cast<size_t>(a.num_sn), 0);
  for (int s = 0; s < a.num_sn; ++s) {
    const int first = a.sn_first[static_cast<size_t>(s)];
    const int last = a.sn_first[static_cast<size_t>(s) + 1] - 1;
    std::vector<int>& rs = R[static_cast<size_t>(s)];
    for (int j = first; j <= last; ++j) {
      for (const int i : below[static_cast<size_t>(j)]) {
        if (i > last) rs.push_back(i);
      }
    }
    for (const int c : kids[static_cast<size_t>(s)]) {
      for (const int i : R[static_cast<size_t>(c)]) {
        if (i > last) rs.push_back(i);
      }
      level[static_cast<size_t>(s)] = std::max(level[static_cast<size_t>(s)], level[static_cast<size_t>(c)] + 1);
    }
    std::sort(rs.begin(), rs.end());
    rs.erase(std::unique(rs.begin(), rs.end()), rs.end());
    const int parent = rs.empty() ? -1 : a.snode_of[static_cast<size_t>(rs.front())];
    if (parent >= 0) kids[static_cast<size_t>(parent)].push_back(s);
    if (parent != a.sn_parent[static_cast<size_t>(s)]) {
      std::printf("%s: supernode %d parent %d, serial %d\n", name.c_str(), s, a.sn_parent[static_cast<size_t>(s)],
                  parent);
      return false;
    }
    if (level[static_cast<size_t>(s)] != a.sn_level[static_cast<size_t>(s)]) {
      std::printf("%s: supernode %d level %d, serial %d\n", name.c_str(), s, a.sn_level[static_cast<size_t>(s)],
                  level[static_cast<size_t>(s)]);
      return false;
    }
    const int32_t* front = a.sn_idx.data() + a.sn_idx_ptr[static_cast<size_t>(s)] + (last - first + 1);
    const std::vector<int> got(front, front + a.sn_m[static_cast<size_t>(s)]);
    if (got != rs) {
      std::printf("%s: supernode %d front rows differ (%zu vs serial %zu)\n", name.c_str(), s, got.size(), rs.size());
      return false;
    }
  }
  int cross = 0;  // supernodes whose parent is not the next dissection ancestor's
  for (int s = 0; s < a.num_sn; ++s) {
    if (!kids[static_cast<size_t>(s)].empty() && kids[static_cast<size_t>(s)].size() != 2) ++cross;
  }
  std::printf("%s: n=%d supernodes=%d levels=%d factor_nnz=%lld non-binary=%d ok\n", name.c_str(), a.n, a.num_sn,
              a.num_levels, static_cast<long long>(a.factor_nnz), cross);
  return true;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s <seed> <workload>...\n", argv[0]);
    return 2;
  }
  const uint64_t seed = std::strtoull(argv[1], nullptr, 10);
  for (int k = 2; k < argc; ++k) {
    try {
      if (!Check(argv[k], seed)) return 1;
    } catch (const std::exception& e) {
      std::printf("%s: %s\n", argv[k], e.what());
      return 1;
    }
  }
  return 0;
}
