// Fill-reducing ordering of the oracle's LDL^T (test infrastructure only).
#pragma once

#include <cstdint>
#include <vector>

namespace oracle {

/// Approximate minimum degree order of the symmetric pattern given as
/// adjacency lists (ptr, idx: both triangles; diagonal entries ignored).
/// order[k] = the vertex eliminated k-th.
std::vector<int32_t> AmdOrder(int n, const std::vector<int64_t>& ptr, const std::vector<int32_t>& idx);

}  // namespace oracle
