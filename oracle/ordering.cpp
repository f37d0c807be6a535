// Approximate minimum degree ordering for the oracle's LDL^T - TEST
// INFRASTRUCTURE ONLY (oracle.h).
//
// The reference factors with Eigen's SimplicialLDLT, whose default ordering
// is AMDOrdering (Eigen/src/OrderingMethods/Amd.h, Eigen3 version unpinned;
// parallel_ldlt.hpp:32-51 calls Base::analyzePattern, which applies it).
// Eigen is not in this image, so the ordering is restated here from the
// published algorithm (Amestoy, Davis and Duff, "An approximate minimum
// degree ordering algorithm", SIAM J. Matrix Anal. Appl. 17(4), 1996):
// a quotient graph of variables and elements; the pivot is a variable of
// least approximate external degree; its element absorbs the elements
// adjacent to it; variables of the new element get the approximate degree
//   min(n_left - |i|, d_i + |L_p \ i|, |A_i \ i| + |L_p \ i| + sum_e |L_e \ L_p|)
// (|L_e \ L_p| from one scan of the new element's variables); elements with
// L_e inside L_p are absorbed; indistinguishable variables (equal variable and
// element adjacency after the update) merge into one supervariable and are
// ordered together; rows with more than 10 sqrt(n) entries are ordered last
// (CSparse's cs_amd rule, which Eigen's port keeps). The exact permutation is
// not Eigen's (ties and bookkeeping differ), so it changes the rounding of
// the factor, not the solution beyond the parity tolerances.
#include "ordering.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <unordered_map>
#include <vector>

namespace oracle {

namespace {

enum Kind : uint8_t { kVariable = 0, kElement = 1, kAbsorbed = 2, kMerged = 3, kDense = 4 };

class Amd {
 public:
  Amd(int n, const std::vector<int64_t>& ptr, const std::vector<int32_t>& idx)
      : n_(n), adj_(static_cast<size_t>(n)), elems_(static_cast<size_t>(n)), vars_(static_cast<size_t>(n)),
        members_(static_cast<size_t>(n)), weight_(static_cast<size_t>(n), 1), kind_(static_cast<size_t>(n), kVariable),
        degree_(static_cast<size_t>(n), 0), elen_(static_cast<size_t>(n), 0), mark_(static_cast<size_t>(n), 0),
        wstamp_(static_cast<size_t>(n), 0), w_(static_cast<size_t>(n), 0), head_(static_cast<size_t>(n) + 1, -1),
        next_(static_cast<size_t>(n), -1), prev_(static_cast<size_t>(n), -1), bucket_(static_cast<size_t>(n), -1) {
    for (int i = 0; i < n; ++i) {
      auto& a = adj_[static_cast<size_t>(i)];
      for (int64_t q = ptr[static_cast<size_t>(i)]; q < ptr[static_cast<size_t>(i) + 1]; ++q) {
        if (idx[static_cast<size_t>(q)] != i) a.push_back(idx[static_cast<size_t>(q)]);
      }
      std::sort(a.begin(), a.end());
      a.erase(std::unique(a.begin(), a.end()), a.end());
    }
    // Dense rows: ordered last, their edges ignored.
    const int dense = std::min(n - 2, std::max(16, static_cast<int>(10.0 * std::sqrt(static_cast<double>(n)))));
    for (int i = 0; i < n; ++i) {
      if (static_cast<int>(adj_[static_cast<size_t>(i)].size()) > dense) kind_[static_cast<size_t>(i)] = kDense;
    }
    for (int i = 0; i < n; ++i) {
      if (kind_[static_cast<size_t>(i)] == kDense) {
        dense_.push_back(i);
        adj_[static_cast<size_t>(i)].clear();
        continue;
      }
      auto& a = adj_[static_cast<size_t>(i)];
      a.erase(std::remove_if(a.begin(), a.end(), [&](int j) { return kind_[static_cast<size_t>(j)] == kDense; }),
              a.end());
      degree_[static_cast<size_t>(i)] = static_cast<int>(a.size());
      live_ += 1;
      Insert(i);
    }
  }

  std::vector<int32_t> Run() {
    std::vector<int32_t> order;
    order.reserve(static_cast<size_t>(n_));
    int left = live_;  // weight not yet eliminated
    while (left > 0) {
      const int p = PopMin();
      left -= weight_[static_cast<size_t>(p)];
      order.push_back(p);
      for (const int m : members_[static_cast<size_t>(p)]) order.push_back(m);
      Eliminate(p, left);
    }
    for (const int d : dense_) order.push_back(d);
    return order;
  }

 private:
  void Insert(int i) {
    const int d = std::clamp(degree_[static_cast<size_t>(i)], 0, n_);
    bucket_[static_cast<size_t>(i)] = d;
    prev_[static_cast<size_t>(i)] = -1;
    next_[static_cast<size_t>(i)] = head_[static_cast<size_t>(d)];
    if (head_[static_cast<size_t>(d)] >= 0) prev_[static_cast<size_t>(head_[static_cast<size_t>(d)])] = i;
    head_[static_cast<size_t>(d)] = i;
    min_degree_ = std::min(min_degree_, d);
  }
  void Remove(int i) {
    const int d = bucket_[static_cast<size_t>(i)];
    if (d < 0) return;
    const int pv = prev_[static_cast<size_t>(i)];
    const int nx = next_[static_cast<size_t>(i)];
    if (pv >= 0) {
      next_[static_cast<size_t>(pv)] = nx;
    } else {
      head_[static_cast<size_t>(d)] = nx;
    }
    if (nx >= 0) prev_[static_cast<size_t>(nx)] = pv;
    bucket_[static_cast<size_t>(i)] = -1;
  }
  int PopMin() {
    while (head_[static_cast<size_t>(min_degree_)] < 0) ++min_degree_;
    const int p = head_[static_cast<size_t>(min_degree_)];
    Remove(p);
    return p;
  }
  bool Principal(int j) const { return kind_[static_cast<size_t>(j)] == kVariable; }

  void Eliminate(int p, int left) {
    const int tag = ++stamp_;
    mark_[static_cast<size_t>(p)] = tag;
    // The new element: every principal variable adjacent to p directly or
    // through one of p's elements (which p absorbs).
    std::vector<int32_t> lp;
    int lp_weight = 0;
    auto take = [&](int j) {
      if (Principal(j) && mark_[static_cast<size_t>(j)] != tag) {
        mark_[static_cast<size_t>(j)] = tag;
        lp.push_back(j);
        lp_weight += weight_[static_cast<size_t>(j)];
      }
    };
    for (const int j : adj_[static_cast<size_t>(p)]) take(j);
    for (const int e : elems_[static_cast<size_t>(p)]) {
      if (kind_[static_cast<size_t>(e)] != kElement) continue;
      for (const int j : vars_[static_cast<size_t>(e)]) take(j);
      kind_[static_cast<size_t>(e)] = kAbsorbed;
      std::vector<int32_t>().swap(vars_[static_cast<size_t>(e)]);
    }
    kind_[static_cast<size_t>(p)] = kElement;
    std::vector<int32_t>().swap(adj_[static_cast<size_t>(p)]);
    std::vector<int32_t>().swap(elems_[static_cast<size_t>(p)]);
    elen_[static_cast<size_t>(p)] = lp_weight;
    // |L_e \ L_p| for the elements next to the new element's variables.
    for (const int i : lp) {
      for (const int e : elems_[static_cast<size_t>(i)]) {
        if (kind_[static_cast<size_t>(e)] != kElement) continue;
        if (wstamp_[static_cast<size_t>(e)] != tag) {
          wstamp_[static_cast<size_t>(e)] = tag;
          w_[static_cast<size_t>(e)] = elen_[static_cast<size_t>(e)];
        }
        w_[static_cast<size_t>(e)] -= weight_[static_cast<size_t>(i)];
      }
    }
    // Update the variables of the new element.
    for (const int i : lp) {
      auto& ei = elems_[static_cast<size_t>(i)];
      int external = 0;
      size_t out = 0;
      for (const int e : ei) {
        if (kind_[static_cast<size_t>(e)] != kElement) continue;
        if (w_[static_cast<size_t>(e)] <= 0) {
          kind_[static_cast<size_t>(e)] = kAbsorbed;  // L_e inside L_p: absorbed
          std::vector<int32_t>().swap(vars_[static_cast<size_t>(e)]);
          continue;
        }
        external += w_[static_cast<size_t>(e)];
        ei[out++] = e;
      }
      ei.resize(out);
      ei.push_back(p);
      auto& ai = adj_[static_cast<size_t>(i)];
      int a_weight = 0;
      out = 0;
      for (const int j : ai) {
        if (!Principal(j) || mark_[static_cast<size_t>(j)] == tag) continue;  // covered by element p
        a_weight += weight_[static_cast<size_t>(j)];
        ai[out++] = j;
      }
      ai.resize(out);
      const int wi = weight_[static_cast<size_t>(i)];
      const int lp_ext = lp_weight - wi;
      int d = std::min(degree_[static_cast<size_t>(i)] + lp_ext, a_weight + lp_ext + external);
      d = std::min(d, left - wi);
      degree_[static_cast<size_t>(i)] = std::max(d, 0);
    }
    vars_[static_cast<size_t>(p)] = lp;
    // Mass elimination: variables of L_p with the same variable and element
    // adjacency become one supervariable.
    std::unordered_map<uint64_t, std::vector<int32_t>> by_hash;
    for (const int i : lp) {
      uint64_t h = 0;
      for (const int j : adj_[static_cast<size_t>(i)]) h += static_cast<uint64_t>(j) * 0x9E3779B97F4A7C15ULL + 1;
      for (const int e : elems_[static_cast<size_t>(i)]) h += static_cast<uint64_t>(e) * 0xC2B2AE3D27D4EB4FULL + 7;
      h ^= (adj_[static_cast<size_t>(i)].size() << 20) ^ elems_[static_cast<size_t>(i)].size();
      by_hash[h].push_back(i);
    }
    for (auto& kv : by_hash) {
      std::vector<int32_t>& group = kv.second;
      if (group.size() < 2) continue;
      for (size_t x = 0; x < group.size(); ++x) {
        const int i = group[x];
        if (!Principal(i)) continue;
        std::vector<int32_t> ai = adj_[static_cast<size_t>(i)];
        std::vector<int32_t> ei = elems_[static_cast<size_t>(i)];
        std::sort(ai.begin(), ai.end());
        std::sort(ei.begin(), ei.end());
        for (size_t y = x + 1; y < group.size(); ++y) {
          const int j = group[y];
          if (!Principal(j)) continue;
          if (adj_[static_cast<size_t>(j)].size() != ai.size() || elems_[static_cast<size_t>(j)].size() != ei.size()) {
            continue;
          }
          std::vector<int32_t> aj = adj_[static_cast<size_t>(j)];
          std::vector<int32_t> ej = elems_[static_cast<size_t>(j)];
          std::sort(aj.begin(), aj.end());
          std::sort(ej.begin(), ej.end());
          if (aj != ai || ej != ei) continue;
          // j joins i: ordered right after it, its weight added to i's and
          // no longer external to it.
          weight_[static_cast<size_t>(i)] += weight_[static_cast<size_t>(j)];
          degree_[static_cast<size_t>(i)] = std::max(0, degree_[static_cast<size_t>(i)] - weight_[static_cast<size_t>(j)]);
          weight_[static_cast<size_t>(j)] = 0;
          kind_[static_cast<size_t>(j)] = kMerged;
          auto& mi = members_[static_cast<size_t>(i)];
          mi.push_back(j);
          for (const int m : members_[static_cast<size_t>(j)]) mi.push_back(m);
          std::vector<int32_t>().swap(members_[static_cast<size_t>(j)]);
          std::vector<int32_t>().swap(adj_[static_cast<size_t>(j)]);
          std::vector<int32_t>().swap(elems_[static_cast<size_t>(j)]);
          Remove(j);
        }
      }
    }
    // New degrees into the buckets.
    for (const int i : lp) {
      if (!Principal(i)) continue;
      Remove(i);
      Insert(i);
    }
  }

  int n_;
  std::vector<std::vector<int32_t>> adj_, elems_, vars_, members_;
  std::vector<int> weight_;
  std::vector<uint8_t> kind_;
  std::vector<int> degree_, elen_, mark_, wstamp_, w_;
  std::vector<int> head_, next_, prev_, bucket_;
  std::vector<int32_t> dense_;
  int live_{0};
  int stamp_{0};
  int min_degree_{0};
};

}  // namespace

std::vector<int32_t> AmdOrder(int n, const std::vector<int64_t>& ptr, const std::vector<int32_t>& idx) {
  if (n <= 0) return {};
  if (n <= 2) {
    std::vector<int32_t> order(static_cast<size_t>(n));
    std::iota(order.begin(), order.end(), 0);
    return order;
  }
  return Amd(n, ptr, idx).Run();
}

}  // namespace oracle
