// Scan normals and node normals on the device: the per-increment normal
// estimation of the problem build, restated from normals.cpp:70-103 (kNN
// PCA normal of a scan point, oriented toward its sensor origin),
// normals.cpp:128-160 (one neighbourhood structure over all scans) and
// normals.cpp:162-220 (each grid node takes the normal of its nearest scan
// point; pca / hybrid / weighted).
//
// Neighbour search: a uniform grid over the points (counting sort by cell
// with integer atomics; the order inside a cell does not matter) and, per
// query, rings of cells around the query's cell until no unvisited point can
// enter the k best. The k best are kept sorted by (squared distance, point
// index), the order an exact search with ties broken by index returns, and
// the PCA sums run in that order; with contraction off (-ffp-contract=off)
// every value equals the host builder's bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "normals.hpp"

namespace sdfslam {

namespace {

constexpr int kThreads = 256;
constexpr int kMaxK = 32;

struct PointGridDev {
  double x0, y0, cell;
  int gx, gy;
};

// Where a kernel finds the point grid: by value (computed on the host from
// host arrays) or in device memory (computed on the device by KPointGrid).
struct GridByValue {
  PointGridDev g;
  __device__ PointGridDev Load() const { return g; }
};
struct GridInMemory {
  const PointGridDev* p;
  __device__ PointGridDev Load() const { return *p; }
};

__device__ __forceinline__ int CellCoord(double v, double v0, double cell, int n) {
  const double f = floor((v - v0) / cell);
  return f < 0.0 ? 0 : (f >= n ? n - 1 : static_cast<int>(f));
}

template <class GridSrc>
__global__ void KCellCount(const double* __restrict__ px, const double* __restrict__ py, int64_t n,
                           GridSrc src, int* __restrict__ cell_of, int* __restrict__ count) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const PointGridDev g = src.Load();
  const int c = CellCoord(py[i], g.y0, g.cell, g.gy) * g.gx + CellCoord(px[i], g.x0, g.cell, g.gx);
  cell_of[i] = c;
  atomicAdd(&count[c], 1);
}

// Exclusive scan of count[0 .. cells) into start[0 .. cells], cells = gx * gy
// (one workgroup of kScanThreads; serial over chunks).
constexpr int kScanThreads = 1024;
template <class GridSrc>
__global__ void __launch_bounds__(kScanThreads) KCellScan(const int* __restrict__ count, GridSrc src,
                                                          int* __restrict__ start) {
  const PointGridDev g = src.Load();
  const int cells = g.gx * g.gy;
  __shared__ int part[kScanThreads];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < cells; base += kScanThreads) {
    const int i = base + static_cast<int>(threadIdx.x);
    const int v = i < cells ? count[i] : 0;
    part[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
      const int t = threadIdx.x >= static_cast<unsigned>(off) ? part[threadIdx.x - off] : 0;
      __syncthreads();
      part[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < cells) start[i] = carry + part[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == kScanThreads - 1) carry += part[kScanThreads - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) start[cells] = carry;
}

__global__ void KCellScatter(const int* __restrict__ cell_of, int64_t n, const int* __restrict__ start,
                             int* __restrict__ fill, int* __restrict__ sorted) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const int c = cell_of[i];
  sorted[start[c] + atomicAdd(&fill[c], 1)] = static_cast<int>(i);
}

// KCellScatter that also stores the points' coordinates in cell order, so a
// range of cells is one contiguous range of coordinates.
__global__ void KCellScatterXY(const int* __restrict__ cell_of, int64_t n, const int* __restrict__ start,
                               int* __restrict__ fill, const double* __restrict__ px,
                               const double* __restrict__ py, int* __restrict__ sorted, double* __restrict__ sx,
                               double* __restrict__ sy) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const int c = cell_of[i];
  const int t = start[c] + atomicAdd(&fill[c], 1);
  sorted[t] = static_cast<int>(i);
  sx[t] = px[i];
  sy[t] = py[i];
}

// Inserts point i at squared distance dd into the k best, kept sorted by
// (distance, index) (the host builder's rule).
__device__ __forceinline__ void Insert(int i, double dd, int k, int& found, int* idx, double* d2) {
  if (found == k) {
    if (dd > d2[k - 1] || (dd == d2[k - 1] && i > idx[k - 1])) return;
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

// The k nearest points to (qx, qy), sorted by (distance, index); returns
// min(k, n). Ring r holds the cells at Chebyshev distance r from the
// query's cell (clamped into the grid); a point in ring r + 1 or beyond is
// farther than r cells from the query, and farther than r - 1 cells even if
// rounding put it in a neighbouring cell, so after ring r the search stops
// once the k-th distance is below (r - 1) cells.
__device__ int Knn(double qx, double qy, int k, const double* __restrict__ px, const double* __restrict__ py,
                   const PointGridDev& g, const int* __restrict__ start, const int* __restrict__ sorted,
                   int* idx, double* d2) {
  const int cx = CellCoord(qx, g.x0, g.cell, g.gx);
  const int cy = CellCoord(qy, g.y0, g.cell, g.gy);
  const int rmax = max(max(cx, g.gx - 1 - cx), max(cy, g.gy - 1 - cy));
  int found = 0;
  for (int r = 0; r <= rmax; ++r) {
    for (int yy = max(cy - r, 0); yy <= min(cy + r, g.gy - 1); ++yy) {
      const bool edge = yy == cy - r || yy == cy + r;
      const int step = edge || r == 0 ? 1 : 2 * r;
      for (int xx = cx - r; xx <= cx + r; xx += step) {
        if (xx < 0 || xx >= g.gx) continue;
        const int c = yy * g.gx + xx;
        for (int t = start[c]; t < start[c + 1]; ++t) {
          const int i = sorted[t];
          const double dx = px[i] - qx;
          const double dy = py[i] - qy;
          Insert(i, dx * dx + dy * dy, k, found, idx, d2);
        }
      }
    }
    if (found == k && r >= 2) {
      const double lim = static_cast<double>(r - 1) * g.cell;
      if (d2[k - 1] < lim * lim) break;
    }
  }
  return found;
}

// The k best of a search held in registers: K is the compile-time k, so
// every index into the arrays is a constant and nothing goes to scratch
// memory. Entries are kept sorted by (squared distance, point index), the
// order an exact search with ties broken by index returns; unused entries
// hold +inf.
template <int K>
struct TopK {
  double d[K];
  int id[K];
  int found;
  __device__ void Init() {
#pragma unroll
    for (int j = 0; j < K; ++j) {
      d[j] = __builtin_huge_val();
      id[j] = 0x7fffffff;
    }
    found = 0;
  }
  static __device__ bool Less(double da, int ia, double db, int ib) { return da < db || (da == db && ia < ib); }
  __device__ void Insert(int i, double dd) {
    if (!Less(dd, i, d[K - 1], id[K - 1])) return;
    // Entry j takes the old entry j - 1 if the new point sorts before it,
    // else the new point if it sorts before the old entry j (descending j
    // reads only old entries).
#pragma unroll
    for (int j = K - 1; j > 0; --j) {
      const bool before_prev = Less(dd, i, d[j - 1], id[j - 1]);
      const bool before_here = Less(dd, i, d[j], id[j]);
      const double nd = before_prev ? d[j - 1] : (before_here ? dd : d[j]);
      const int ni = before_prev ? id[j - 1] : (before_here ? i : id[j]);
      d[j] = nd;
      id[j] = ni;
    }
    if (Less(dd, i, d[0], id[0])) {
      d[0] = dd;
      id[0] = i;
    }
    found = min(found + 1, K);
  }
};

// The k nearest points of one query searched by a whole wavefront: the
// lanes take the points of a run of cells 64 at a time (cells sorted row
// major, so a row of a ring is one contiguous run of sorted points), and the
// points closer than the current k-th join the k best, which every lane
// holds alike (candidates broadcast one by one from the lanes that found
// them). The same rings and stopping rule as Knn, so the same k best.
__device__ __forceinline__ int Lane() { return static_cast<int>(threadIdx.x & 63); }

template <int K>
__device__ void WaveScanRange(int a, int b, double qx, double qy, const double* __restrict__ sx,
                              const double* __restrict__ sy, const int* __restrict__ sid, TopK<K>& best) {
  for (int base = a; base < b; base += 64) {
    const int t = base + Lane();
    double dd = __builtin_huge_val();
    int id = 0x7fffffff;
    if (t < b) {
      const double dx = sx[t] - qx;
      const double dy = sy[t] - qy;
      dd = dx * dx + dy * dy;
      id = sid[t];
    }
    uint64_t m = __ballot(t < b && TopK<K>::Less(dd, id, best.d[K - 1], best.id[K - 1]));
    while (m != 0) {
      const int src = __builtin_ctzll(m);
      m &= m - 1;
      const int lo = __builtin_amdgcn_readlane(static_cast<int>(__double_as_longlong(dd)), src);
      const int hi = __builtin_amdgcn_readlane(static_cast<int>(__double_as_longlong(dd) >> 32), src);
      const double cd =
          __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
      const int ci = __builtin_amdgcn_readlane(id, src);
      best.Insert(ci, cd);
    }
  }
}

template <int K>
__device__ void WaveKnn(double qx, double qy, const PointGridDev& g, const int* __restrict__ start,
                        const double* __restrict__ sx, const double* __restrict__ sy, const int* __restrict__ sid,
                        TopK<K>& best) {
  best.Init();
  const int cx = CellCoord(qx, g.x0, g.cell, g.gx);
  const int cy = CellCoord(qy, g.y0, g.cell, g.gy);
  const int rmax = max(max(cx, g.gx - 1 - cx), max(cy, g.gy - 1 - cy));
  for (int r = 0; r <= rmax; ++r) {
    const int x0 = max(cx - r, 0);
    const int x1 = min(cx + r, g.gx - 1);
    // Rows cy - r and cy + r: whole runs of cells.
    if (cy - r >= 0) {
      WaveScanRange<K>(start[(cy - r) * g.gx + x0], start[(cy - r) * g.gx + x1 + 1], qx, qy, sx, sy, sid, best);
    }
    if (r > 0 && cy + r < g.gy) {
      WaveScanRange<K>(start[(cy + r) * g.gx + x0], start[(cy + r) * g.gx + x1 + 1], qx, qy, sx, sy, sid, best);
    }
    // Columns cx - r and cx + r between them: the lanes read the cells'
    // ranges side by side (lane j: row y0 + j / 2, left or right by j % 2)
    // and the non-empty ones are scanned one after another.
    if (r > 0) {
      const int y0 = max(cy - r + 1, 0);
      const int cells = 2 * (min(cy + r - 1, g.gy - 1) - y0 + 1);
      for (int base = 0; base < cells; base += 64) {
        const int j = base + Lane();
        const int xx = (j & 1) != 0 ? cx + r : cx - r;
        int a = 0;
        int b = 0;
        if (j < cells && xx >= 0 && xx < g.gx) {
          const int c = (y0 + (j >> 1)) * g.gx + xx;
          a = start[c];
          b = start[c + 1];
        }
        uint64_t m = __ballot(b > a);
        while (m != 0) {
          const int src = __builtin_ctzll(m);
          m &= m - 1;
          WaveScanRange<K>(__builtin_amdgcn_readlane(a, src), __builtin_amdgcn_readlane(b, src), qx, qy, sx, sy,
                           sid, best);
        }
      }
    }
    if (best.found == K && r >= 2) {
      const double lim = static_cast<double>(r - 1) * g.cell;
      if (best.d[K - 1] < lim * lim) break;
    }
  }
}

// PCA normal of point i from its k nearest (PcaNormal, normals.cpp:70-103):
// the eigenvector of the smaller covariance eigenvalue (closed form),
// oriented toward the point's origin, and the eigenvalue ratio.
__device__ __forceinline__ void PcaFromSums(double a, double b, double c, double qx, double qy, double ox,
                                            double oy, double& nx, double& ny, double& corner) {
  const double half_tr = 0.5 * (a + c);
  const double half_diff = 0.5 * (a - c);
  const double rad = sqrt(half_diff * half_diff + b * b);
  const double l_min = half_tr - rad;
  const double l_max = half_tr + rad;
  const double v1x = b;
  const double v1y = l_min - a;
  const double v2x = l_min - c;
  const double v2y = b;
  const double n1 = sqrt(v1x * v1x + v1y * v1y);
  const double n2 = sqrt(v2x * v2x + v2y * v2y);
  if (n1 < 1e-300 && n2 < 1e-300) {
    nx = a <= c ? 1.0 : 0.0;
    ny = a <= c ? 0.0 : 1.0;
  } else if (n1 >= n2) {
    nx = v1x / n1;
    ny = v1y / n1;
  } else {
    nx = v2x / n2;
    ny = v2y / n2;
  }
  if (nx * (ox - qx) + ny * (oy - qy) < 0.0) {
    nx = -nx;
    ny = -ny;
  }
  corner = fmax(l_min, 0.0) / fmax(l_max, 1e-12);
}

// kNN PCA normals with one workgroup per cell of the point grid: the cell's
// points are the queries (one per thread), and their candidates, the points
// of the ring cells around the cell, are staged in LDS by the whole
// workgroup and read by every thread alike: first rings 0 to 2 (the 5 x 5
// cells every search visits), then, while some query is not settled by
// Knn's stopping rule, one more ring at a time. Every query gets the k best
// Knn gives. Outputs at each point's own index.
constexpr int kCellThreads = 256;
constexpr int kCellCandidates = 2048;  // candidates staged in LDS per pass
constexpr int kCellRanges = 2 * kCellThreads + 8;  // ranges of one staged ring

// The k best over the staged candidates first, first + L, ... below count
// (a query's L threads share its candidates).
template <int K, int L>
__device__ __forceinline__ void TakeStaged(const double* cand_x, const double* cand_y, const int* cand_id, int first,
                                           int count, double qx, double qy, TopK<K>& best) {
  for (int k = first; k < count; k += L) {
    const double dx = cand_x[k] - qx;
    const double dy = cand_y[k] - qy;
    best.Insert(cand_id[k], dx * dx + dy * dy);
  }
}

// Merges the k best of a query's L threads (aligned groups of L lanes; each
// point in at most one thread's list) into every thread of the group: k
// rounds of a group minimum over the threads' heads, the winner's thread
// dropping its head.
template <int K, int L>
__device__ __forceinline__ void MergeGroup(TopK<K>& best) {
  double md[K];
  int mid[K];
#pragma unroll
  for (int r = 0; r < K; ++r) {
    double wd = best.d[0];
    int wi = best.id[0];
#pragma unroll
    for (int m = 1; m < L; m <<= 1) {
      const double od = __shfl_xor(wd, m, L);
      const int oi = __shfl_xor(wi, m, L);
      if (TopK<K>::Less(od, oi, wd, wi)) {
        wd = od;
        wi = oi;
      }
    }
    md[r] = wd;
    mid[r] = wi;
    if (best.d[0] == wd && best.id[0] == wi) {
#pragma unroll
      for (int j = 0; j + 1 < K; ++j) {
        best.d[j] = best.d[j + 1];
        best.id[j] = best.id[j + 1];
      }
      best.d[K - 1] = __builtin_huge_val();
      best.id[K - 1] = 0x7fffffff;
    }
  }
  best.found = 0;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    best.d[j] = md[j];
    best.id[j] = mid[j];
    best.found += md[j] < __builtin_huge_val() ? 1 : 0;
  }
}

template <int K, int L, class GridSrc>
__global__ void __launch_bounds__(kCellThreads)
    KPointNormalsCell(const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ ox,
                      const double* __restrict__ oy, GridSrc src, const int* __restrict__ start,
                      const double* __restrict__ sx, const double* __restrict__ sy, const int* __restrict__ sid,
                      double* __restrict__ nrm_x, double* __restrict__ nrm_y, double* __restrict__ corner) {
  __shared__ double cand_x[kCellCandidates];
  __shared__ double cand_y[kCellCandidates];
  __shared__ int cand_id[kCellCandidates];
  __shared__ int range_a[kCellRanges];
  __shared__ int range_off[kCellRanges + 1];
  __shared__ int num_ranges;
  const PointGridDev g = src.Load();
  const int c = static_cast<int>(blockIdx.x);
  if (c >= g.gx * g.gy) return;
  const int q0 = start[c];
  const int q1 = start[c + 1];
  if (q1 == q0) return;
  const int cx = c % g.gx;
  const int cy = c / g.gx;
  const int rmax = max(max(cx, g.gx - 1 - cx), max(cy, g.gy - 1 - cy));
  const int tid = static_cast<int>(threadIdx.x);
  // Lists the point ranges of rings r0 .. r1 into range_a / range_off, ring
  // by ring (ring 0: the cell; ring r: its two rows as runs, then its
  // 2 (2 r - 1) column cells), so that the nearest candidates come first
  // and later ones rarely enter the k best; returns the candidate count.
  auto ranges_before = [](int r) { return r == 0 ? 0 : 1 + 2 * r * (r - 1); };
  auto list_rings = [&](int r0, int r1) -> int {
    const int base0 = ranges_before(r0);
    const int nr = ranges_before(r1 + 1) - base0;
    __syncthreads();
    for (int j = tid; j < nr; j += kCellThreads) {
      int r = r0;
      while (ranges_before(r + 1) - base0 <= j) ++r;
      const int slot = j - (ranges_before(r) - base0);
      int a = 0;
      int b = 0;
      if (r == 0) {
        a = start[c];
        b = start[c + 1];
      } else if (slot < 2) {
        const int row = slot == 0 ? cy - r : cy + r;
        if (row >= 0 && row < g.gy) {
          a = start[row * g.gx + max(cx - r, 0)];
          b = start[row * g.gx + min(cx + r, g.gx - 1) + 1];
        }
      } else {
        const int k = slot - 2;
        const int yy = cy - r + 1 + (k >> 1);
        const int xx = (k & 1) != 0 ? cx + r : cx - r;
        if (yy >= 0 && yy < g.gy && xx >= 0 && xx < g.gx) {
          a = start[yy * g.gx + xx];
          b = start[yy * g.gx + xx + 1];
        }
      }
      range_a[j] = a;
      range_off[j + 1] = b - a;
    }
    if (tid == 0) num_ranges = nr;
    __syncthreads();
    if (tid == 0) {
      range_off[0] = 0;
      for (int k = 0; k < nr; ++k) range_off[k + 1] += range_off[k];
    }
    __syncthreads();
    return range_off[nr];
  };
  // Stages candidates [pass, pass + count) of the listed ranges.
  auto stage = [&](int pass, int count) {
    __syncthreads();
    const int nr = num_ranges;
    for (int k = tid; k < count; k += kCellThreads) {
      const int gi = pass + k;
      int lo = 0;
      int hi = nr - 1;
      while (lo < hi) {  // the last range whose offset is <= gi
        const int mid = (lo + hi + 1) >> 1;
        if (range_off[mid] <= gi) {
          lo = mid;
        } else {
          hi = mid - 1;
        }
      }
      const int t = range_a[lo] + (gi - range_off[lo]);
      cand_x[k] = sx[t];
      cand_y[k] = sy[t];
      cand_id[k] = sid[t];
    }
    __syncthreads();
  };
  constexpr int kQueries = kCellThreads / L;  // queries per pass, L threads each
  const int sub = tid % L;
  for (int qbase = q0; qbase < q1; qbase += kQueries) {
    const int q = qbase + tid / L;
    const bool active = q < q1;
    const double qx = active ? sx[q] : 0.0;
    const double qy = active ? sy[q] : 0.0;
    TopK<K> best;
    best.Init();
    const int r_first = min(2, rmax);
    int total = list_rings(0, r_first);
    for (int pass = 0; pass < total; pass += kCellCandidates) {
      const int count = min(kCellCandidates, total - pass);
      stage(pass, count);
      if (active) TakeStaged<K, L>(cand_x, cand_y, cand_id, sub, count, qx, qy, best);
    }
    MergeGroup<K, L>(best);
    auto settled = [&](int r) {
      if (r >= rmax) return true;  // every cell visited
      if (best.found < K || r < 2) return false;
      const double lim = static_cast<double>(r - 1) * g.cell;
      return best.d[K - 1] < lim * lim;
    };
    bool done = !active || settled(r_first);
    for (int r = r_first + 1; r <= rmax; ++r) {
      if (__syncthreads_or(done ? 0 : 1) == 0) break;
      if (4 * r > kCellRanges) {
        // A ring wider than the range table: the unsettled queries finish
        // on their own (first thread of each), ring by ring.
        for (int rr = r; !done && sub == 0 && rr <= rmax; ++rr) {
          for (int yy = max(cy - rr, 0); yy <= min(cy + rr, g.gy - 1); ++yy) {
            const bool edge = yy == cy - rr || yy == cy + rr;
            const int step = edge ? 1 : 2 * rr;
            for (int xx = cx - rr; xx <= cx + rr; xx += step) {
              if (xx < 0 || xx >= g.gx) continue;
              const int cc = yy * g.gx + xx;
              for (int t = start[cc]; t < start[cc + 1]; ++t) {
                const double dx = sx[t] - qx;
                const double dy = sy[t] - qy;
                best.Insert(sid[t], dx * dx + dy * dy);
              }
            }
          }
          done = settled(rr);
        }
        break;
      }
      // The group's merged list stays with its first thread; the others
      // collect the ring's candidates from empty lists.
      if (sub != 0) best.Init();
      total = list_rings(r, r);
      for (int pass = 0; pass < total; pass += kCellCandidates) {
        const int count = min(kCellCandidates, total - pass);
        stage(pass, count);
        if (!done) TakeStaged<K, L>(cand_x, cand_y, cand_id, sub, count, qx, qy, best);
      }
      MergeGroup<K, L>(best);
      if (!done) done = settled(r);
    }
    if (!active || sub != 0) continue;
    double nbx[K];
    double nby[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
      nbx[j] = px[best.id[j]];
      nby[j] = py[best.id[j]];
    }
    double mx = 0.0;
    double my = 0.0;
#pragma unroll
    for (int j = 0; j < K; ++j) {
      mx = mx + nbx[j];
      my = my + nby[j];
    }
    mx = mx / static_cast<double>(K);
    my = my / static_cast<double>(K);
    double a = 0.0;
    double b = 0.0;
    double cc = 0.0;
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const double dx = nbx[j] - mx;
      const double dy = nby[j] - my;
      a += dx * dx;
      b += dx * dy;
      cc += dy * dy;
    }
    const int i = sid[q];
    double nx, ny, cr;
    PcaFromSums(a, b, cc, qx, qy, ox[i], oy[i], nx, ny, cr);
    nrm_x[i] = nx;
    nrm_y[i] = ny;
    corner[i] = cr;
  }
}

// The same for any k up to kMaxK, with the k best in (scratch) arrays.
template <class GridSrc>
__global__ void KPointNormals(const double* __restrict__ px, const double* __restrict__ py,
                              const double* __restrict__ ox, const double* __restrict__ oy, int64_t n, int k,
                              GridSrc src, const int* __restrict__ start, const int* __restrict__ sorted,
                              double* __restrict__ nrm_x, double* __restrict__ nrm_y, double* __restrict__ corner) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const PointGridDev g = src.Load();
  int idx[kMaxK];
  double d2[kMaxK];
  const double qx = px[i];
  const double qy = py[i];
  const int found = Knn(qx, qy, k, px, py, g, start, sorted, idx, d2);
  double mx = 0.0;
  double my = 0.0;
  for (int j = 0; j < found; ++j) {
    mx = mx + px[idx[j]];
    my = my + py[idx[j]];
  }
  mx = mx / static_cast<double>(found);
  my = my / static_cast<double>(found);
  double a = 0.0;
  double b = 0.0;
  double c = 0.0;
  for (int j = 0; j < found; ++j) {
    const double dx = px[idx[j]] - mx;
    const double dy = py[idx[j]] - my;
    a += dx * dx;
    b += dx * dy;
    c += dy * dy;
  }
  double nx, ny, cr;
  PcaFromSums(a, b, c, qx, qy, ox[i], oy[i], nx, ny, cr);
  nrm_x[i] = nx;
  nrm_y[i] = ny;
  corner[i] = cr;
}

// Point normals with one workgroup per cell for k up to 16 (sx, sy, sid:
// the points in cell order; `cells` workgroups, at least gx * gy), the
// general kernel above that.
template <class GridSrc>
void LaunchPointNormalsCells(hipStream_t s, const double* px, const double* py, const double* ox, const double* oy,
                             int64_t n, int kk, GridSrc src, unsigned cells, const int* start, const int* sorted,
                             const double* sx, const double* sy, double* nrm_x, double* nrm_y, double* corner) {
#define SDFSLAM_POINT_NORMALS_CELL(K)                                                                       \
  case K:                                                                                                \
    KPointNormalsCell<K, 4, GridSrc><<<cells, kCellThreads, 0, s>>>(px, py, ox, oy, src, start, sx, sy, sorted, \
                                                                 nrm_x, nrm_y, corner);                    \
    return;
  switch (kk) {
    SDFSLAM_POINT_NORMALS_CELL(2)
    SDFSLAM_POINT_NORMALS_CELL(3)
    SDFSLAM_POINT_NORMALS_CELL(4)
    SDFSLAM_POINT_NORMALS_CELL(5)
    SDFSLAM_POINT_NORMALS_CELL(6)
    SDFSLAM_POINT_NORMALS_CELL(7)
    SDFSLAM_POINT_NORMALS_CELL(8)
    SDFSLAM_POINT_NORMALS_CELL(9)
    SDFSLAM_POINT_NORMALS_CELL(10)
    SDFSLAM_POINT_NORMALS_CELL(11)
    SDFSLAM_POINT_NORMALS_CELL(12)
    SDFSLAM_POINT_NORMALS_CELL(13)
    SDFSLAM_POINT_NORMALS_CELL(14)
    SDFSLAM_POINT_NORMALS_CELL(15)
    SDFSLAM_POINT_NORMALS_CELL(16)
    default: {
      const auto pb = static_cast<unsigned>((n + kThreads - 1) / kThreads);
      KPointNormals<GridSrc><<<pb, kThreads, 0, s>>>(px, py, ox, oy, n, kk, src, start, sorted, nrm_x, nrm_y,
                                                      corner);
    }
  }
#undef SDFSLAM_POINT_NORMALS_CELL
}


// Normal and eikonal scale of every grid node from its nearest scan point
// (AssignNodeNormals, normals.cpp:162-220).
// Normal and eikonal scale of a node at (qx, qy) from its nearest scan
// point (AssignNodeNormals, normals.cpp:162-220).
__device__ __forceinline__ void NodeNormalFrom(int nearest, double qx, double qy, const double* __restrict__ px,
                                               const double* __restrict__ py, const double* __restrict__ pnx,
                                               const double* __restrict__ pny, const double* __restrict__ pcorner,
                                               int method, double corner_threshold, double& rx, double& ry,
                                               double& scale) {
  const double n_x = pnx[nearest];
  const double n_y = pny[nearest];
  const double corner = fmin(fmax(pcorner[nearest] / corner_threshold, 0.0), 1.0);
  rx = n_x;
  ry = n_y;
  scale = 1.0;
  if (method == SDFSLAM_NORMALS_HYBRID) {
    const double tx = px[nearest] - qx;
    const double ty = py[nearest] - qy;
    const double tn = sqrt(tx * tx + ty * ty);
    if (!(tn < 1e-9)) {
      double cx = tx / tn;
      double cy = ty / tn;
      if (cx * n_x + cy * n_y < 0.0) {
        cx = -cx;
        cy = -cy;
      }
      const double bx = n_x * (1.0 - corner) + cx * corner;
      const double by = n_y * (1.0 - corner) + cy * corner;
      const double bn = sqrt(bx * bx + by * by);
      if (bn > 1e-9) {
        rx = bx / bn;
        ry = by / bn;
      }
    }
  } else if (method == SDFSLAM_NORMALS_WEIGHTED) {
    scale = 1.0 - corner;
  }
}

// One thread per node.
template <class GridSrc>
__global__ void KNodeNormals(int nx, int ny, double min_x, double min_y, double dx, double dy,
                             const double* __restrict__ px, const double* __restrict__ py,
                             const double* __restrict__ pnx, const double* __restrict__ pny,
                             const double* __restrict__ pcorner, GridSrc src, const int* __restrict__ start,
                             const int* __restrict__ sorted, int method, double corner_threshold,
                             double* __restrict__ out_nx, double* __restrict__ out_ny,
                             double* __restrict__ out_scale) {
  const int64_t node = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (node >= static_cast<int64_t>(nx) * ny) return;
  const PointGridDev g = src.Load();
  const int w = static_cast<int>(node % nx);
  const int h = static_cast<int>(node / nx);
  const double qx = min_x + w * dx;
  const double qy = min_y + h * dy;
  int nearest = 0;
  double dist = 0.0;
  Knn(qx, qy, 1, px, py, g, start, sorted, &nearest, &dist);
  double rx, ry, scale;
  NodeNormalFrom(nearest, qx, qy, px, py, pnx, pny, pcorner, method, corner_threshold, rx, ry, scale);
  out_nx[node] = rx;
  out_ny[node] = ry;
  out_scale[node] = scale;
}

// One wavefront per node (WaveKnn with k = 1; sx, sy, sid: the points in
// cell order); lane 0 stores.
template <class GridSrc>
__global__ void __launch_bounds__(kThreads)
    KNodeNormalsWave(int nx, int ny, double min_x, double min_y, double dx, double dy,
                     const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ pnx,
                     const double* __restrict__ pny, const double* __restrict__ pcorner, GridSrc src,
                     const int* __restrict__ start, const double* __restrict__ sx, const double* __restrict__ sy,
                     const int* __restrict__ sid, int method, double corner_threshold, double* __restrict__ out_nx,
                     double* __restrict__ out_ny, double* __restrict__ out_scale) {
  const int64_t node = blockIdx.x * static_cast<int64_t>(kThreads / 64) + (threadIdx.x >> 6);
  if (node >= static_cast<int64_t>(nx) * ny) return;
  const PointGridDev g = src.Load();
  const int w = static_cast<int>(node % nx);
  const int h = static_cast<int>(node / nx);
  const double qx = min_x + w * dx;
  const double qy = min_y + h * dy;
  TopK<1> best;
  WaveKnn<1>(qx, qy, g, start, sx, sy, sid, best);
  double rx, ry, scale;
  NodeNormalFrom(best.id[0], qx, qy, px, py, pnx, pny, pcorner, method, corner_threshold, rx, ry, scale);
  if (Lane() == 0) {
    out_nx[node] = rx;
    out_ny[node] = ry;
    out_scale[node] = scale;
  }
}

// Map seeding from scans (InitializeFromScans, map_init.cpp:31-69): every
// node takes the distance to its nearest scan point, negated behind the
// surface when signed (the nearest point's sensor origin decides: the node
// is in front when it is no farther from that origin than the point).
template <class GridSrc>
__global__ void KInitFromScans(int nx, int ny, double min_x, double min_y, double dx, double dy,
                               const double* __restrict__ px, const double* __restrict__ py,
                               const double* __restrict__ ox, const double* __restrict__ oy, GridSrc src,
                               const int* __restrict__ start, const int* __restrict__ sorted, int signed_distance,
                               double* __restrict__ out) {
  const int64_t node = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (node >= static_cast<int64_t>(nx) * ny) return;
  const PointGridDev g = src.Load();
  const int w = static_cast<int>(node % nx);
  const int h = static_cast<int>(node / nx);
  const double qx = min_x + w * dx;
  const double qy = min_y + h * dy;
  int nearest = 0;
  double squared = 0.0;
  Knn(qx, qy, 1, px, py, g, start, sorted, &nearest, &squared);
  double sign = 1.0;
  if (signed_distance != 0) {
    const double ax = qx - ox[nearest];
    const double ay = qy - oy[nearest];
    const double bx = px[nearest] - ox[nearest];
    const double by = py[nearest] - oy[nearest];
    const double to_node = sqrt(ax * ax + ay * ay);
    const double to_surface = sqrt(bx * bx + by * by);
    sign = to_node <= to_surface ? 1.0 : -1.0;
  }
  out[node] = sign * sqrt(squared);
}

/// Uniform grid over the points: about two points per cell over their
/// bounding box (at most 2048 cells a side).
PointGridDev MakePointGrid(const double* points_xy, int64_t n) {
  double x0 = points_xy[0];
  double y0 = points_xy[1];
  double x1 = x0;
  double y1 = y0;
  for (int64_t i = 1; i < n; ++i) {
    x0 = std::min(x0, points_xy[2 * i]);
    x1 = std::max(x1, points_xy[2 * i]);
    y0 = std::min(y0, points_xy[2 * i + 1]);
    y1 = std::max(y1, points_xy[2 * i + 1]);
  }
  const double extent = std::max({x1 - x0, y1 - y0, 1e-6});
  double cell = std::sqrt((x1 - x0 + 1e-6) * (y1 - y0 + 1e-6) * 2.0 / static_cast<double>(n));
  cell = std::max(cell, extent / 2048.0);
  return PointGridDev{x0, y0, cell, static_cast<int>((x1 - x0) / cell) + 1, static_cast<int>((y1 - y0) / cell) + 1};
}

#define HIPN(x)                                        \
  do {                                                 \
    const hipError_t e_ = (x);                         \
    if (e_ != hipSuccess) return Fail(e_, #x);         \
  } while (0)

thread_local std::string t_error;

int Fail(hipError_t e, const char* what) {
  t_error = std::string(what) + ": " + hipGetErrorString(e);
  return e == hipErrorOutOfMemory ? SDFSLAM_ERR_OOM : SDFSLAM_ERR_HIP;
}

/// Device buffers of one call, grown as needed and never freed while the
/// process runs (freeing device memory waits for the whole device);
/// borrowed from a process-wide pool so that concurrent calls do not share.
struct Workspace {
  size_t bytes{0};
  char* dev{nullptr};
  hipStream_t stream{nullptr};
};

class WorkspacePool {
 public:
  static WorkspacePool& Get() {
    static WorkspacePool* pool = new WorkspacePool();  // lives until the process ends
    return *pool;
  }
  Workspace* Borrow() {
    std::lock_guard<std::mutex> lock(mu_);
    if (!free_.empty()) {
      Workspace* w = free_.back();
      free_.pop_back();
      return w;
    }
    return new Workspace();
  }
  void Return(Workspace* w) {
    std::lock_guard<std::mutex> lock(mu_);
    free_.push_back(w);
  }

 private:
  std::mutex mu_;
  std::vector<Workspace*> free_;
};

struct Lease {
  Workspace* w;
  Lease() : w(WorkspacePool::Get().Borrow()) {}
  ~Lease() { WorkspacePool::Get().Return(w); }
};

}  // namespace

const char* NormalsLastError() { return t_error.c_str(); }

int NodeNormalsOnDevice(const double* points_xy, const double* origins_xy, int64_t num_points, int32_t k,
                        const sdfslam_grid& grid, int32_t method, double corner_threshold,
                        double* node_normal_xy, double* node_scale) {
  const int64_t nodes = static_cast<int64_t>(grid.nx) * grid.ny;
  if (num_points <= 0) {
    // No scan points: every node keeps the default (zero normal, scale 1).
    for (int64_t v = 0; v < nodes; ++v) {
      node_normal_xy[2 * v] = 0.0;
      node_normal_xy[2 * v + 1] = 0.0;
      node_scale[v] = 1.0;
    }
    return SDFSLAM_OK;
  }
  const int64_t n = num_points;
  if (k > kMaxK) {
    t_error = "k_neighbors above " + std::to_string(kMaxK);
    return SDFSLAM_ERR_INVALID_ARGUMENT;
  }
  if (n > (int64_t{1} << 30) || static_cast<int64_t>(grid.nx) * grid.ny > (int64_t{1} << 30)) {
    t_error = "too many points or nodes";
    return SDFSLAM_ERR_INVALID_ARGUMENT;
  }
  const int kk = static_cast<int>(std::min<int64_t>(std::max(2, k), n));
  const PointGridDev g = MakePointGrid(points_xy, n);
  const int64_t cells = static_cast<int64_t>(g.gx) * g.gy;
  // Layout of the workspace (256-byte aligned parts).
  auto align = [](size_t b) { return (b + 255) & ~static_cast<size_t>(255); };
  const size_t nd = align(static_cast<size_t>(n) * sizeof(double));
  const size_t ni = align(static_cast<size_t>(n) * sizeof(int));
  const size_t ci = align(static_cast<size_t>(cells + 1) * sizeof(int));
  const size_t vd = align(static_cast<size_t>(nodes) * sizeof(double));
  const size_t need = 9 * nd + 2 * ni + 3 * ci + 3 * vd;
  Lease lease;
  Workspace& ws = *lease.w;
  if (ws.stream == nullptr) HIPN(hipStreamCreateWithFlags(&ws.stream, hipStreamNonBlocking));
  if (need > ws.bytes) {
    // The outgrown buffer is kept (never freed while the process runs).
    ws.dev = nullptr;
    ws.bytes = 0;
    const size_t want = need + need / 2;
    HIPN(hipMalloc(&ws.dev, want));
    ws.bytes = want;
  }
  char* p = ws.dev;
  auto take = [&p](size_t b) {
    char* q = p;
    p += b;
    return q;
  };
  double* px = reinterpret_cast<double*>(take(nd));
  double* py = reinterpret_cast<double*>(take(nd));
  double* ox = reinterpret_cast<double*>(take(nd));
  double* oy = reinterpret_cast<double*>(take(nd));
  double* pnx = reinterpret_cast<double*>(take(nd));
  double* pny = reinterpret_cast<double*>(take(nd));
  double* pcorner = reinterpret_cast<double*>(take(nd));
  double* sx = reinterpret_cast<double*>(take(nd));
  double* sy = reinterpret_cast<double*>(take(nd));
  int* cell_of = reinterpret_cast<int*>(take(ni));
  int* sorted = reinterpret_cast<int*>(take(ni));
  int* count = reinterpret_cast<int*>(take(ci));
  int* start = reinterpret_cast<int*>(take(ci));
  int* fill = reinterpret_cast<int*>(take(ci));
  double* vnx = reinterpret_cast<double*>(take(vd));
  double* vny = reinterpret_cast<double*>(take(vd));
  double* vscale = reinterpret_cast<double*>(take(vd));
  // Split the interleaved host arrays.
  std::vector<double> hx(static_cast<size_t>(n)), hy(static_cast<size_t>(n)), hox(static_cast<size_t>(n)),
      hoy(static_cast<size_t>(n));
  for (int64_t i = 0; i < n; ++i) {
    hx[static_cast<size_t>(i)] = points_xy[2 * i];
    hy[static_cast<size_t>(i)] = points_xy[2 * i + 1];
    hox[static_cast<size_t>(i)] = origins_xy[2 * i];
    hoy[static_cast<size_t>(i)] = origins_xy[2 * i + 1];
  }
  hipStream_t s = ws.stream;
  const size_t nb = static_cast<size_t>(n) * sizeof(double);
  HIPN(hipMemcpyAsync(px, hx.data(), nb, hipMemcpyHostToDevice, s));
  HIPN(hipMemcpyAsync(py, hy.data(), nb, hipMemcpyHostToDevice, s));
  HIPN(hipMemcpyAsync(ox, hox.data(), nb, hipMemcpyHostToDevice, s));
  HIPN(hipMemcpyAsync(oy, hoy.data(), nb, hipMemcpyHostToDevice, s));
  HIPN(hipMemsetAsync(count, 0, static_cast<size_t>(cells + 1) * sizeof(int), s));
  HIPN(hipMemsetAsync(fill, 0, static_cast<size_t>(cells + 1) * sizeof(int), s));
  const unsigned pb = static_cast<unsigned>((n + kThreads - 1) / kThreads);
  KCellCount<<<pb, kThreads, 0, s>>>(px, py, n, GridByValue{g}, cell_of, count);
  KCellScan<<<1, kScanThreads, 0, s>>>(count, GridByValue{g}, start);
  KCellScatterXY<<<pb, kThreads, 0, s>>>(cell_of, n, start, fill, px, py, sorted, sx, sy);
  LaunchPointNormalsCells(s, px, py, ox, oy, n, kk, GridByValue{g}, static_cast<unsigned>(cells), start, sorted, sx,
                         sy, pnx, pny, pcorner);
  const auto wb = static_cast<unsigned>((nodes + kThreads / 64 - 1) / (kThreads / 64));
  KNodeNormalsWave<<<wb, kThreads, 0, s>>>(grid.nx, grid.ny, grid.min_x, grid.min_y, grid.dx, grid.dy, px, py, pnx,
                                           pny, pcorner, GridByValue{g}, start, sx, sy, sorted, method,
                                           corner_threshold, vnx, vny, vscale);
  HIPN(hipGetLastError());
  std::vector<double> rx(static_cast<size_t>(nodes)), ry(static_cast<size_t>(nodes));
  const size_t vb_bytes = static_cast<size_t>(nodes) * sizeof(double);
  HIPN(hipMemcpyAsync(rx.data(), vnx, vb_bytes, hipMemcpyDeviceToHost, s));
  HIPN(hipMemcpyAsync(ry.data(), vny, vb_bytes, hipMemcpyDeviceToHost, s));
  HIPN(hipMemcpyAsync(node_scale, vscale, vb_bytes, hipMemcpyDeviceToHost, s));
  HIPN(hipStreamSynchronize(s));
  for (int64_t v = 0; v < nodes; ++v) {
    node_normal_xy[2 * v] = rx[static_cast<size_t>(v)];
    node_normal_xy[2 * v + 1] = ry[static_cast<size_t>(v)];
  }
  return SDFSLAM_OK;
}

int InitFromScansOnDevice(const double* points_xy, const double* origins_xy, int64_t num_points,
                          const sdfslam_grid& grid, int32_t signed_distance, double* map_values) {
  if (num_points <= 0) return SDFSLAM_OK;  // an empty scan list leaves the map unchanged
  const int64_t n = num_points;
  const int64_t nodes = static_cast<int64_t>(grid.nx) * grid.ny;
  if (n > (int64_t{1} << 30) || nodes > (int64_t{1} << 30)) {
    t_error = "too many points or nodes";
    return SDFSLAM_ERR_INVALID_ARGUMENT;
  }
  const PointGridDev g = MakePointGrid(points_xy, n);
  const int64_t cells = static_cast<int64_t>(g.gx) * g.gy;
  auto align = [](size_t b) { return (b + 255) & ~static_cast<size_t>(255); };
  const size_t nd = align(static_cast<size_t>(n) * sizeof(double));
  const size_t ni = align(static_cast<size_t>(n) * sizeof(int));
  const size_t ci = align(static_cast<size_t>(cells + 1) * sizeof(int));
  const size_t vd = align(static_cast<size_t>(nodes) * sizeof(double));
  const size_t need = 4 * nd + 2 * ni + 3 * ci + vd;
  Lease lease;
  Workspace& ws = *lease.w;
  if (ws.stream == nullptr) HIPN(hipStreamCreateWithFlags(&ws.stream, hipStreamNonBlocking));
  if (need > ws.bytes) {
    ws.dev = nullptr;
    ws.bytes = 0;
    const size_t want = need + need / 2;
    HIPN(hipMalloc(&ws.dev, want));
    ws.bytes = want;
  }
  char* p = ws.dev;
  auto take = [&p](size_t b) {
    char* q = p;
    p += b;
    return q;
  };
  double* px = reinterpret_cast<double*>(take(nd));
  double* py = reinterpret_cast<double*>(take(nd));
  double* ox = reinterpret_cast<double*>(take(nd));
  double* oy = reinterpret_cast<double*>(take(nd));
  int* cell_of = reinterpret_cast<int*>(take(ni));
  int* sorted = reinterpret_cast<int*>(take(ni));
  int* count = reinterpret_cast<int*>(take(ci));
  int* start = reinterpret_cast<int*>(take(ci));
  int* fill = reinterpret_cast<int*>(take(ci));
  double* out = reinterpret_cast<double*>(take(vd));
  std::vector<double> hx(static_cast<size_t>(n)), hy(static_cast<size_t>(n)), hox(static_cast<size_t>(n)),
      hoy(static_cast<size_t>(n));
  for (int64_t i = 0; i < n; ++i) {
    hx[static_cast<size_t>(i)] = points_xy[2 * i];
    hy[static_cast<size_t>(i)] = points_xy[2 * i + 1];
    hox[static_cast<size_t>(i)] = origins_xy[2 * i];
    hoy[static_cast<size_t>(i)] = origins_xy[2 * i + 1];
  }
  hipStream_t s = ws.stream;
  const size_t nb = static_cast<size_t>(n) * sizeof(double);
  HIPN(hipMemcpyAsync(px, hx.data(), nb, hipMemcpyHostToDevice, s));
  HIPN(hipMemcpyAsync(py, hy.data(), nb, hipMemcpyHostToDevice, s));
  HIPN(hipMemcpyAsync(ox, hox.data(), nb, hipMemcpyHostToDevice, s));
  HIPN(hipMemcpyAsync(oy, hoy.data(), nb, hipMemcpyHostToDevice, s));
  HIPN(hipMemsetAsync(count, 0, static_cast<size_t>(cells + 1) * sizeof(int), s));
  HIPN(hipMemsetAsync(fill, 0, static_cast<size_t>(cells + 1) * sizeof(int), s));
  const unsigned pb = static_cast<unsigned>((n + kThreads - 1) / kThreads);
  KCellCount<<<pb, kThreads, 0, s>>>(px, py, n, GridByValue{g}, cell_of, count);
  KCellScan<<<1, kScanThreads, 0, s>>>(count, GridByValue{g}, start);
  KCellScatter<<<pb, kThreads, 0, s>>>(cell_of, n, start, fill, sorted);
  const unsigned vb = static_cast<unsigned>((nodes + kThreads - 1) / kThreads);
  KInitFromScans<<<vb, kThreads, 0, s>>>(grid.nx, grid.ny, grid.min_x, grid.min_y, grid.dx, grid.dy, px, py, ox,
                                         oy, GridByValue{g}, start, sorted, signed_distance, out);
  HIPN(hipGetLastError());
  HIPN(hipMemcpyAsync(map_values, out, static_cast<size_t>(nodes) * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPN(hipStreamSynchronize(s));
  return SDFSLAM_OK;
}

// ---- Map seeding without allocation ---------------------------------------
// InitFromScansNoAlloc keeps every intermediate buffer in `scratch`, one
// file-scope device array per GPU (kNoAllocScratchBytes, held until the
// process ends), so it calls no hipMalloc: for callers under graph capture
// or in serving loops. Inputs and output are device memory of the caller;
// the point grid is computed on the device. Every call on a GPU uses that
// GPU's one array: calls that overlap on one GPU race.

constexpr int64_t kNoAllocMaxCells = kNoAllocMaxPoints / 2 + 4097;
constexpr size_t Align256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }
constexpr size_t kNoAllocPointBytes = Align256(sizeof(double) * static_cast<size_t>(kNoAllocMaxPoints));
constexpr size_t kNoAllocIndexBytes = Align256(sizeof(int) * static_cast<size_t>(kNoAllocMaxPoints));
constexpr size_t kNoAllocCellBytes = Align256(sizeof(int) * static_cast<size_t>(kNoAllocMaxCells + 1));
constexpr size_t kNoAllocScratchBytes =
    256 + 4 * kNoAllocPointBytes + 2 * kNoAllocIndexBytes + 3 * kNoAllocCellBytes;

static_assert(kNoAllocScratchBytes <= (size_t{SDFSLAM_NO_ALLOC_SCRATCH_MIB} << 20),
              "scratch beyond the size the header states");

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

// points_xy / origins_xy (x, y pairs) into separate coordinate arrays.
__global__ void KSplitPoints(const double* __restrict__ points_xy, const double* __restrict__ origins_xy,
                             int64_t n, double* __restrict__ px, double* __restrict__ py,
                             double* __restrict__ ox, double* __restrict__ oy) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  px[i] = points_xy[2 * i];
  py[i] = points_xy[2 * i + 1];
  ox[i] = origins_xy[2 * i];
  oy[i] = origins_xy[2 * i + 1];
}

// The point grid of MakePointGrid, on the device (one workgroup): the
// bounding box by a min/max reduction, then the same cell size; the cell
// grows while gx * gy exceeds max_cells (the bound the size allows).
constexpr int kGridThreads = 1024;
__global__ void __launch_bounds__(kGridThreads) KPointGrid(const double* __restrict__ px,
                                                           const double* __restrict__ py, int64_t n,
                                                           int64_t max_cells, PointGridDev* __restrict__ out) {
  __shared__ double red[4][kGridThreads];
  double x0 = px[0], x1 = px[0], y0 = py[0], y1 = py[0];
  for (int64_t i = threadIdx.x; i < n; i += kGridThreads) {
    x0 = fmin(x0, px[i]);
    x1 = fmax(x1, px[i]);
    y0 = fmin(y0, py[i]);
    y1 = fmax(y1, py[i]);
  }
  red[0][threadIdx.x] = x0;
  red[1][threadIdx.x] = x1;
  red[2][threadIdx.x] = y0;
  red[3][threadIdx.x] = y1;
  __syncthreads();
  for (int off = kGridThreads / 2; off > 0; off >>= 1) {
    if (threadIdx.x < static_cast<unsigned>(off)) {
      red[0][threadIdx.x] = fmin(red[0][threadIdx.x], red[0][threadIdx.x + off]);
      red[1][threadIdx.x] = fmax(red[1][threadIdx.x], red[1][threadIdx.x + off]);
      red[2][threadIdx.x] = fmin(red[2][threadIdx.x], red[2][threadIdx.x + off]);
      red[3][threadIdx.x] = fmax(red[3][threadIdx.x], red[3][threadIdx.x + off]);
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  x0 = red[0][0];
  x1 = red[1][0];
  y0 = red[2][0];
  y1 = red[3][0];
  const double extent = fmax(fmax(x1 - x0, y1 - y0), 1e-6);
  double cell = sqrt((x1 - x0 + 1e-6) * (y1 - y0 + 1e-6) * 2.0 / static_cast<double>(n));
  cell = fmax(cell, extent / 2048.0);
  int gx = static_cast<int>((x1 - x0) / cell) + 1;
  int gy = static_cast<int>((y1 - y0) / cell) + 1;
  for (int t = 0; t < 64 && static_cast<int64_t>(gx) * gy > max_cells; ++t) {
    cell *= 2.0;
    gx = static_cast<int>((x1 - x0) / cell) + 1;
    gy = static_cast<int>((y1 - y0) / cell) + 1;
  }
  *out = PointGridDev{x0, y0, cell, gx, gy};
}

__global__ void KClearInts(int* __restrict__ a, int* __restrict__ b, int64_t n) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  a[i] = 0;
  b[i] = 0;
}

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_dev[kMaxScratchDevices] = {};

/// The device address of `scratch` on the current device, looked up once
/// per device (with that device current).
int ScratchOfCurrentDevice(unsigned char** out) {
  int device = 0;
  HIPN(hipGetDevice(&device));
  if (device < 0 || device >= kMaxScratchDevices) {
    t_error = "device index beyond the scratch table";
    return SDFSLAM_ERR_INVALID_ARGUMENT;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[device] == nullptr) {
    void* ptr = nullptr;
    HIPN(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_dev[device] = static_cast<unsigned char*>(ptr);
  }
  *out = g_scratch_dev[device];
  return SDFSLAM_OK;
}

}  // namespace

int InitFromScansNoAlloc(const double* points_xy, const double* origins_xy, int64_t num_points,
                         const sdfslam_grid& grid, int32_t signed_distance, double* map_values, hipStream_t s) {
  if (num_points <= 0) return SDFSLAM_OK;  // an empty scan list leaves the map unchanged
  const int64_t n = num_points;
  const int64_t nodes = static_cast<int64_t>(grid.nx) * grid.ny;
  if (n > kNoAllocMaxPoints) {
    t_error = "num_points " + std::to_string(n) + " above the allocation-free maximum of " +
              std::to_string(kNoAllocMaxPoints);
    return SDFSLAM_ERR_INVALID_ARGUMENT;
  }
  if (grid.nx < 2 || grid.ny < 2 || nodes > (int64_t{1} << 30)) {
    t_error = "grid needs 2..2^30 nodes with nx, ny >= 2";
    return SDFSLAM_ERR_INVALID_ARGUMENT;
  }
  if (points_xy == nullptr || origins_xy == nullptr || map_values == nullptr) {
    t_error = "null device pointer";
    return SDFSLAM_ERR_INVALID_ARGUMENT;
  }
  unsigned char* base = nullptr;
  const int st = ScratchOfCurrentDevice(&base);
  if (st != SDFSLAM_OK) return st;
  unsigned char* p = base;
  auto take = [&p](size_t b) {
    unsigned char* q = p;
    p += b;
    return q;
  };
  auto* gp = reinterpret_cast<PointGridDev*>(take(256));
  auto* px = reinterpret_cast<double*>(take(kNoAllocPointBytes));
  auto* py = reinterpret_cast<double*>(take(kNoAllocPointBytes));
  auto* ox = reinterpret_cast<double*>(take(kNoAllocPointBytes));
  auto* oy = reinterpret_cast<double*>(take(kNoAllocPointBytes));
  auto* cell_of = reinterpret_cast<int*>(take(kNoAllocIndexBytes));
  auto* sorted = reinterpret_cast<int*>(take(kNoAllocIndexBytes));
  auto* count = reinterpret_cast<int*>(take(kNoAllocCellBytes));
  auto* start = reinterpret_cast<int*>(take(kNoAllocCellBytes));
  auto* fill = reinterpret_cast<int*>(take(kNoAllocCellBytes));
  // At most n/2 + 4097 cells for n points (MakePointGrid's cell size), so
  // count and fill need n/2 + 4098 entries cleared.
  const int64_t max_cells = n / 2 + 4097;
  const unsigned pb = static_cast<unsigned>((n + kThreads - 1) / kThreads);
  const unsigned cb = static_cast<unsigned>((max_cells + 1 + kThreads - 1) / kThreads);
  const unsigned vb = static_cast<unsigned>((nodes + kThreads - 1) / kThreads);
  KSplitPoints<<<pb, kThreads, 0, s>>>(points_xy, origins_xy, n, px, py, ox, oy);
  KPointGrid<<<1, kGridThreads, 0, s>>>(px, py, n, max_cells, gp);
  KClearInts<<<cb, kThreads, 0, s>>>(count, fill, max_cells + 1);
  KCellCount<<<pb, kThreads, 0, s>>>(px, py, n, GridInMemory{gp}, cell_of, count);
  KCellScan<<<1, kScanThreads, 0, s>>>(count, GridInMemory{gp}, start);
  KCellScatter<<<pb, kThreads, 0, s>>>(cell_of, n, start, fill, sorted);
  KInitFromScans<<<vb, kThreads, 0, s>>>(grid.nx, grid.ny, grid.min_x, grid.min_y, grid.dx, grid.dy, px, py, ox,
                                         oy, GridInMemory{gp}, start, sorted, signed_distance, map_values);
  HIPN(hipGetLastError());
  return SDFSLAM_OK;
}


// ---- Eikonal normals of a device-resident problem ---------------------------
// The eikonal specs' normals of a problem whose specs already live on the
// device, computed there from its scan points (the point rows with expected
// SDF 0, in row order: ComputeScanNormals' order) at its poses: the
// per-increment normal estimation of the problem build, without a host
// round trip. World points use the host's per-frame cos and sin and its
// Pose2::Apply arithmetic, so every value equals the host builder's.

namespace {

// Scan point q (row rows[q] of frame f) in world coordinates,
// (c p.x - s p.y) + t.x and (s p.x + c p.y) + t.y as Pose2::Apply, and its
// sensor origin t.
__global__ void KScanWorld(const int32_t* __restrict__ rows, int64_t n, const int32_t* __restrict__ frame,
                           const double* __restrict__ sx, const double* __restrict__ sy,
                           const double* __restrict__ cos_sin, const double* __restrict__ poses,
                           double* __restrict__ px, double* __restrict__ py, double* __restrict__ ox,
                           double* __restrict__ oy) {
  const int64_t q = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (q >= n) return;
  const int r = rows[q];
  const int f = frame[r];
  const double c = cos_sin[2 * f];
  const double sn = cos_sin[2 * f + 1];
  const double tx = poses[3 * f];
  const double ty = poses[3 * f + 1];
  px[q] = (c * sx[r] - sn * sy[r]) + tx;
  py[q] = (sn * sx[r] + c * sy[r]) + ty;
  ox[q] = tx;
  oy[q] = ty;
}

// Eikonal spec e at node (w, h) takes that node's normal.
__global__ void KEikonalNormals(const int32_t* __restrict__ eik_w, const int32_t* __restrict__ eik_h, int64_t m,
                                int nx, const double* __restrict__ vnx, const double* __restrict__ vny,
                                double* __restrict__ eik_nx, double* __restrict__ eik_ny) {
  const int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (e >= m) return;
  const int64_t node = static_cast<int64_t>(eik_h[e]) * nx + eik_w[e];
  eik_nx[e] = vnx[node];
  eik_ny[e] = vny[node];
}

}  // namespace

size_t EikonalNormalsWorkBytes(int64_t scan_points, int64_t nodes) {
  const int64_t max_cells = scan_points / 2 + 4097;
  const size_t nd = Align256(sizeof(double) * static_cast<size_t>(std::max<int64_t>(scan_points, 1)));
  const size_t ni = Align256(sizeof(int) * static_cast<size_t>(std::max<int64_t>(scan_points, 1)));
  const size_t ci = Align256(sizeof(int) * static_cast<size_t>(max_cells + 1));
  const size_t vd = Align256(sizeof(double) * static_cast<size_t>(nodes));
  return 256 + 9 * nd + 2 * ni + 3 * ci + 3 * vd;
}

namespace {

// Rows of one appended frame per workgroup (FrameRowsJob). The scan rows are
// written directly; a hallucinated sample's row follows from the count of
// points of length >= 1e-9 before its point (a workgroup scan per chunk).
constexpr int kFrameThreads = 1024;
__global__ void __launch_bounds__(kFrameThreads) KFrameRows(FrameRowsJob j) {
  const int f = static_cast<int>(blockIdx.x);
  const int64_t p0 = j.scan_off[f];
  const int64_t n = j.scan_off[f + 1] - p0;
  const int64_t r0 = j.row_off[f];
  const int64_t l0 = j.list_off[f];
  const int frame = j.first_frame + f;
  const int per_point = j.per_side * (j.both_directions != 0 ? 2 : 1);
  __shared__ int part[kFrameThreads];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t base = 0; base < n; base += kFrameThreads) {
    const int64_t i = base + threadIdx.x;
    double px = 0.0;
    double py = 0.0;
    double norm = 0.0;
    int valid = 0;
    if (i < n) {
      px = j.scan_xy[2 * (p0 + i)];
      py = j.scan_xy[2 * (p0 + i) + 1];
      const int64_t r = r0 + i;
      j.frame[r] = frame;
      j.x[r] = px;
      j.y[r] = py;
      j.expected_sdf[r] = 0.0;
      j.sqrt_weight[r] = j.sqrt_w_scan;
      j.scan_list[l0 + i] = static_cast<int32_t>(r);
      norm = sqrt(px * px + py * py);
      valid = (per_point > 0 && !(norm < 1e-9)) ? 1 : 0;
    }
    // Inclusive scan of the valid flags over the chunk.
    part[threadIdx.x] = valid;
    __syncthreads();
    for (int off = 1; off < kFrameThreads; off <<= 1) {
      const int t = threadIdx.x >= static_cast<unsigned>(off) ? part[threadIdx.x - off] : 0;
      __syncthreads();
      part[threadIdx.x] += t;
      __syncthreads();
    }
    if (valid != 0) {
      const int64_t before = carry + part[threadIdx.x] - 1;
      int64_t r = r0 + n + before * per_point;
      // toward_origin = -point / norm * step (component by component).
      const double tx = -px / norm * j.step;
      const double ty = -py / norm * j.step;
      for (int s = 1; s <= j.per_side; ++s) {
        const double ds = static_cast<double>(s);
        j.frame[r] = frame;
        j.x[r] = px + tx * ds;
        j.y[r] = py + ty * ds;
        j.expected_sdf[r] = j.step * ds;
        j.sqrt_weight[r] = j.sqrt_w_hallucination;
        ++r;
        if (j.both_directions != 0) {
          j.frame[r] = frame;
          j.x[r] = px - tx * ds;
          j.y[r] = py - ty * ds;
          j.expected_sdf[r] = -j.step * ds;
          j.sqrt_weight[r] = j.sqrt_w_hallucination;
          ++r;
        }
      }
    }
    __syncthreads();
    if (threadIdx.x == kFrameThreads - 1) carry += part[kFrameThreads - 1];
    __syncthreads();
  }
}

__global__ void KClearBytes(uint8_t* __restrict__ p, int64_t n) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i < n) p[i] = 0;
}

// One thread per point row: the row placed at its frame's pose
// (Pose2::Apply), located in the grid (GridMap::Locate, grid_map.hpp:58-72:
// outside the grid marks nothing), marks the nodes of the dilated cell.
// Every write stores 1, so concurrent writes to a node agree.
__global__ void KActiveMarks(const int32_t* __restrict__ frame, const double* __restrict__ sx,
                             const double* __restrict__ sy, int64_t n, const double* __restrict__ cos_sin,
                             const double* __restrict__ poses, sdfslam_grid g, int margin,
                             uint8_t* __restrict__ marks) {
  const int64_t r = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (r >= n) return;
  const int f = frame[r];
  const double c = cos_sin[2 * f];
  const double s = cos_sin[2 * f + 1];
  const double gx = (c * sx[r] - s * sy[r]) + poses[3 * f];
  const double gy = (s * sx[r] + c * sy[r]) + poses[3 * f + 1];
  const double fx = (gx - g.min_x) / g.dx;
  const double fy = (gy - g.min_y) / g.dy;
  if (fx < 0.0 || fy < 0.0 || fx > g.nx - 1 || fy > g.ny - 1) return;
  const int cw = min(static_cast<int>(fx), g.nx - 2);
  const int ch = min(static_cast<int>(fy), g.ny - 2);
  for (int h = max(0, ch - margin); h <= min(g.ny - 1, ch + 1 + margin); ++h) {
    for (int w = max(0, cw - margin); w <= min(g.nx - 1, cw + 1 + margin); ++w) {
      marks[static_cast<int64_t>(h) * g.nx + w] = 1;
    }
  }
}

}  // namespace

int EnqueueFrameRows(const FrameRowsJob& j, hipStream_t s) {
  if (j.frames <= 0) return SDFSLAM_OK;
  if (j.frames > 65535 || j.per_side < 0 || (j.per_side > 0 && !(j.step > 0.0))) {
    t_error = "appended frames need at most 65535 frames and a positive hallucination step";
    return SDFSLAM_ERR_INVALID_ARGUMENT;
  }
  KFrameRows<<<static_cast<unsigned>(j.frames), kFrameThreads, 0, s>>>(j);
  HIPN(hipGetLastError());
  return SDFSLAM_OK;
}

int EnqueueActiveMarks(const int32_t* frame, const double* x, const double* y, int64_t num_rows,
                       const double* cos_sin, const double* poses, const sdfslam_grid& grid, int32_t margin,
                       uint8_t* marks, hipStream_t s) {
  const int64_t nodes = static_cast<int64_t>(grid.nx) * grid.ny;
  KClearBytes<<<static_cast<unsigned>((nodes + kThreads - 1) / kThreads), kThreads, 0, s>>>(marks, nodes);
  if (num_rows > 0) {
    KActiveMarks<<<static_cast<unsigned>((num_rows + kThreads - 1) / kThreads), kThreads, 0, s>>>(
        frame, x, y, num_rows, cos_sin, poses, grid, margin, marks);
  }
  HIPN(hipGetLastError());
  return SDFSLAM_OK;
}

int EnqueueEikonalNormals(const EikonalNormalsJob& j, void* work, hipStream_t s) {
  const int64_t n = j.scan_points;
  const int64_t nodes = static_cast<int64_t>(j.nx) * j.ny;
  if (n <= 0) {
    t_error = "no scan points for the eikonal normals";
    return SDFSLAM_ERR_INVALID_ARGUMENT;
  }
  if (j.k > kMaxK || (j.method != SDFSLAM_NORMALS_PCA && j.method != SDFSLAM_NORMALS_HYBRID)) {
    t_error = "device eikonal normals need the pca or hybrid method and k_neighbors <= 32";
    return SDFSLAM_ERR_INVALID_ARGUMENT;
  }
  const int64_t max_cells = n / 2 + 4097;
  const size_t nd = Align256(sizeof(double) * static_cast<size_t>(n));
  const size_t ni = Align256(sizeof(int) * static_cast<size_t>(n));
  const size_t ci = Align256(sizeof(int) * static_cast<size_t>(max_cells + 1));
  const size_t vd = Align256(sizeof(double) * static_cast<size_t>(nodes));
  unsigned char* p = static_cast<unsigned char*>(work);
  auto take = [&p](size_t b) {
    unsigned char* q = p;
    p += b;
    return q;
  };
  auto* gp = reinterpret_cast<PointGridDev*>(take(256));
  auto* px = reinterpret_cast<double*>(take(nd));
  auto* py = reinterpret_cast<double*>(take(nd));
  auto* ox = reinterpret_cast<double*>(take(nd));
  auto* oy = reinterpret_cast<double*>(take(nd));
  auto* pnx = reinterpret_cast<double*>(take(nd));
  auto* pny = reinterpret_cast<double*>(take(nd));
  auto* pcorner = reinterpret_cast<double*>(take(nd));
  auto* sx = reinterpret_cast<double*>(take(nd));
  auto* sy = reinterpret_cast<double*>(take(nd));
  auto* cell_of = reinterpret_cast<int*>(take(ni));
  auto* sorted = reinterpret_cast<int*>(take(ni));
  auto* count = reinterpret_cast<int*>(take(ci));
  auto* start = reinterpret_cast<int*>(take(ci));
  auto* fill = reinterpret_cast<int*>(take(ci));
  auto* vnx = reinterpret_cast<double*>(take(vd));
  auto* vny = reinterpret_cast<double*>(take(vd));
  auto* vscale = reinterpret_cast<double*>(take(vd));
  const int kk = static_cast<int>(std::min<int64_t>(std::max(2, j.k), n));
  const unsigned pb = static_cast<unsigned>((n + kThreads - 1) / kThreads);
  const unsigned cb = static_cast<unsigned>((max_cells + 1 + kThreads - 1) / kThreads);
  KScanWorld<<<pb, kThreads, 0, s>>>(j.scan_rows, n, j.frame, j.sensor_x, j.sensor_y, j.cos_sin, j.poses, px, py, ox,
                                      oy);
  KPointGrid<<<1, kGridThreads, 0, s>>>(px, py, n, max_cells, gp);
  KClearInts<<<cb, kThreads, 0, s>>>(count, fill, max_cells + 1);
  KCellCount<<<pb, kThreads, 0, s>>>(px, py, n, GridInMemory{gp}, cell_of, count);
  KCellScan<<<1, kScanThreads, 0, s>>>(count, GridInMemory{gp}, start);
  KCellScatterXY<<<pb, kThreads, 0, s>>>(cell_of, n, start, fill, px, py, sorted, sx, sy);
  // At most max_cells cells (KPointGrid): the workgroups past gx * gy exit.
  LaunchPointNormalsCells(s, px, py, ox, oy, n, kk, GridInMemory{gp}, static_cast<unsigned>(max_cells), start, sorted,
                         sx, sy, pnx, pny, pcorner);
  const auto wb = static_cast<unsigned>((nodes + kThreads / 64 - 1) / (kThreads / 64));
  KNodeNormalsWave<<<wb, kThreads, 0, s>>>(j.nx, j.ny, j.min_x, j.min_y, j.dx, j.dy, px, py, pnx, pny, pcorner,
                                           GridInMemory{gp}, start, sx, sy, sorted, j.method, j.corner_threshold,
                                           vnx, vny, vscale);
  if (j.num_eikonal > 0) {
    const unsigned eb = static_cast<unsigned>((j.num_eikonal + kThreads - 1) / kThreads);
    KEikonalNormals<<<eb, kThreads, 0, s>>>(j.eik_w, j.eik_h, j.num_eikonal, j.nx, vnx, vny, j.eik_nx, j.eik_ny);
  }
  HIPN(hipGetLastError());
  return SDFSLAM_OK;
}

}  // namespace sdfslam
