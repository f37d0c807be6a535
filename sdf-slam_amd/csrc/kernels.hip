// Device kernels of the joint-LM backend (gfx950, FP64).
//
// Assembly follows the reference's residual model term by term
// (src/problem/residual_math.hpp:27-205, src/core/grid_map.hpp:58-115,
// src/solvers/lm_solver.cpp:30-82, 275-296). Normal equations are formed
// into fixed factor slots with fixed-order reductions (no float atomics),
// the numeric factorization is a level-scheduled multifrontal LDL^T, and the
// LM control of lm_solver.cpp:366-425 runs in a one-block kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <type_traits>

#include "kernels.hpp"

namespace sdfslam {

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;  // wavefronts per workgroup
constexpr int kNB = 64;     // block column width of the panel factorization
constexpr int kTile = 64;   // GEMM tile edge
constexpr int kTk = 16;     // GEMM k chunk

__device__ __forceinline__ bool Halted(const LmState* st) { return st->halt != 0; }

__device__ __forceinline__ double WrapAngle(double a) { return atan2(sin(a), cos(a)); }

typedef double v4d __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double WaveSum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    v += __shfl_xor(v, off, 64);
  }
  return v;
}

// The pose-block and panel extend-add kernels' LDS forms (SetLdsForms).
std::atomic<bool> g_lds_forms{true};

// Where a kernel finds its slot: one slot passed by value as a kernel
// argument (a single problem: no memory indirection before the kernel's
// first loads), or a table of slots in device memory (a batch).
struct OneSlot {
  Slot s;
  __device__ __forceinline__ const Slot& get(int) const { return s; }
};
struct SlotTable {
  const Slot* __restrict__ t;
  __device__ __forceinline__ const Slot& get(int i) const { return t[i]; }
};
// Likewise for the list a level kernel works on (see ListRef).
struct OneList {
  ListRef r;
  __device__ __forceinline__ ListRef get(int) const { return r; }
};
struct ListTable {
  const ListRef* __restrict__ t;
  __device__ __forceinline__ ListRef get(int i) const { return t[i]; }
};

// The slot of this workgroup: its problem, analysis and work buffers.
#define BIND_SLOT(index)                             \
  const Slot& S = slots.get(index);                  \
  [[maybe_unused]] const DeviceProblem& P = S.P;     \
  [[maybe_unused]] const DeviceAnalysis& A = S.A;    \
  [[maybe_unused]] const DeviceWork& W = S.W

// Item `i` of a launch plan's list for `slot`, or -1 beyond its length.
template <class Lists>
__device__ __forceinline__ int PlanItem(const Lists& plan, int slot, int i) {
  const ListRef lr = plan.get(slot);
  return i < lr.count ? lr.items[i] : -1;
}

unsigned Blocks(int64_t n, int threads = kThreads, int64_t cap = 1 << 20) {
  int64_t b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return static_cast<unsigned>(b);
}

// ---------------------------------------------------------------- utilities

// Clears the slot's factor panels (the normal-equation kernels then write
// only the pattern's slots).
template <class Slots>
__global__ void KZeroPanels(Slots slots) {
  BIND_SLOT(blockIdx.y);
  if (Halted(W.state)) return;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < S.panel_total;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    W.panel[i] = 0.0;
  }
}

// ---------------------------------------------------------------- assembly

struct CellW {
  int w, h;
  double a, b;
  double wt[4];
};

// GridMap::FillCell (grid_map.hpp:109-115)
__device__ __forceinline__ void FillWeights(CellW& c) {
  c.wt[0] = (1.0 - c.a) * (1.0 - c.b);
  c.wt[1] = c.a * (1.0 - c.b);
  c.wt[2] = (1.0 - c.a) * c.b;
  c.wt[3] = c.a * c.b;
}

// GridMap::Interpolate (grid_map.hpp:86-93)
__device__ __forceinline__ double InterpolateCell(const GridParams& g, const double* map,
                                                  const CellW& c) {
  const int n00 = c.h * g.nx + c.w;
  double value = 0.0;
  value += c.wt[0] * map[n00];
  value += c.wt[1] * map[n00 + 1];
  value += c.wt[2] * map[n00 + g.nx];
  value += c.wt[3] * map[n00 + g.nx + 1];
  return value;
}

// GridMap::LocateClamped (grid_map.hpp:75-83) + Interpolate
__device__ double InterpolateClamped(const GridParams& g, const double* map, double x, double y) {
  const double fx = (x - g.min_x) / g.dx;
  const double fy = (y - g.min_y) / g.dy;
  CellW c;
  c.w = min(max(static_cast<int>(fx), 0), g.nx - 2);
  c.h = min(max(static_cast<int>(fy), 0), g.ny - 2);
  c.a = fx - c.w;
  c.b = fy - c.h;
  FillWeights(c);
  return InterpolateCell(g, map, c);
}

// Point residual and Jacobian (residual_math.hpp:75-106) with the IRLS
// scaling of AssemblePointRows (lm_solver.cpp:30-61). One thread per row.
template <class Slots>
__global__ void KAssemblePoints(Slots slots, int check) {
  BIND_SLOT(blockIdx.y);
  if (Halted(W.state)) return;
  const int64_t s = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (s >= P.num_points) return;
  const GridParams g = P.grid;
  const int f = P.pt_frame[s];
  const double px = P.pt_x[s];
  const double py = P.pt_y[s];
  const double delta = P.pt_delta[s];
  const double sws = P.pt_sw[s];
  const double tx = W.poses[3 * f];
  const double ty = W.poses[3 * f + 1];
  const double th = W.poses[3 * f + 2];
  const double c = cos(th);
  const double sn = sin(th);
  const double gx = c * px + (-sn) * py + tx;
  const double gy = sn * px + c * py + ty;
  const double fx = (gx - g.min_x) / g.dx;
  const double fy = (gy - g.min_y) / g.dy;
  const bool valid = !(fx < 0.0 || fy < 0.0 || fx > g.nx - 1 || fy > g.ny - 1);
  double residual;
  double wt[4] = {0.0, 0.0, 0.0, 0.0};
  double dp[3] = {0.0, 0.0, 0.0};
  uint32_t key = static_cast<uint32_t>(P.num_cells);
  if (!valid) {
    residual = -delta;
  } else {
    CellW cell;
    cell.w = min(static_cast<int>(fx), g.nx - 2);
    cell.h = min(static_cast<int>(fy), g.ny - 2);
    cell.a = fx - cell.w;
    cell.b = fy - cell.h;
    FillWeights(cell);
    residual = InterpolateCell(g, W.map, cell) - delta;
    double grx, gry;
    if (P.smooth_gradient != 0 && P.smooth_gradient_step > 0.0) {
      // SmoothGradientMeters (residual_math.hpp:63-73)
      const double st = P.smooth_gradient_step;
      grx = (InterpolateClamped(g, W.map, gx + st, gy) - InterpolateClamped(g, W.map, gx - st, gy)) /
            (2.0 * st);
      gry = (InterpolateClamped(g, W.map, gx, gy + st) - InterpolateClamped(g, W.map, gx, gy - st)) /
            (2.0 * st);
    } else if (P.smooth_gradient != 0) {
      // SmoothGradient (residual_math.hpp:40-57)
      double ngx[4], ngy[4];
      const int ws[4] = {cell.w, cell.w + 1, cell.w, cell.w + 1};
      const int hs[4] = {cell.h, cell.h, cell.h + 1, cell.h + 1};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int w = ws[k];
        const int h = hs[k];
        const int w_lo = max(w - 1, 0);
        const int w_hi = min(w + 1, g.nx - 1);
        const int h_lo = max(h - 1, 0);
        const int h_hi = min(h + 1, g.ny - 1);
        ngx[k] = (W.map[h * g.nx + w_hi] - W.map[h * g.nx + w_lo]) / ((w_hi - w_lo) * g.dx);
        ngy[k] = (W.map[h_hi * g.nx + w] - W.map[h_lo * g.nx + w]) / ((h_hi - h_lo) * g.dy);
      }
      const double a = cell.a;
      const double b = cell.b;
      const double s00 = (1.0 - a) * (1.0 - b);
      const double s10 = a * (1.0 - b);
      const double s01 = (1.0 - a) * b;
      const double s11 = a * b;
      grx = s00 * ngx[0] + s10 * ngx[1] + s01 * ngx[2] + s11 * ngx[3];
      gry = s00 * ngy[0] + s10 * ngy[1] + s01 * ngy[2] + s11 * ngy[3];
    } else {
      // GridMap::Gradient (grid_map.hpp:98-106)
      const int n00 = cell.h * g.nx + cell.w;
      const double d00 = W.map[n00];
      const double d10 = W.map[n00 + 1];
      const double d01 = W.map[n00 + g.nx];
      const double d11 = W.map[n00 + g.nx + 1];
      grx = ((1.0 - cell.b) * (d10 - d00) + cell.b * (d11 - d01)) / g.dx;
      gry = ((1.0 - cell.a) * (d01 - d00) + cell.a * (d11 - d10)) / g.dy;
    }
    const double dxdt = -sn * px - c * py;
    const double dydt = c * px - sn * py;
    dp[0] = grx;
    dp[1] = gry;
    dp[2] = grx * dxdt + gry * dydt;
#pragma unroll
    for (int k = 0; k < 4; ++k) wt[k] = cell.wt[k];
    key = static_cast<uint32_t>(cell.h * (g.nx - 1) + cell.w);
  }
  // Huber IRLS weight (problem.hpp:115-121, lm_solver.cpp:39-47).
  const double weighted = sws * residual;
  double hw = 1.0;
  if (P.huber_delta > 0.0) {
    const double ar = fabs(weighted);
    hw = ar <= P.huber_delta ? 1.0 : P.huber_delta / ar;
  }
  const double swh = sqrt(hw);
  W.r_pt[s] = swh * weighted;
  const double scale = swh * sws;
  const int64_t np = P.num_points;
#pragma unroll
  for (int k = 0; k < 4; ++k) W.jw[k * np + s] = valid ? scale * wt[k] : 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) W.jp[k * np + s] = valid ? scale * dp[k] : 0.0;
  W.cellkey[s] = key;
  if (check != 0 && valid && f >= 1) {
    const uint32_t word = A.fp_cellbits[static_cast<int64_t>(f) * A.words_per_pose + (key >> 5)];
    if (((word >> (key & 31U)) & 1U) == 0U) {
      W.pose_overflow[f] = 1;
      W.state->overflow = 1;
      W.state->halt = 1;
    }
  }
}

// EvalEikonalResidual (residual_math.hpp:158-171), AssembleEikonalRows
// (lm_solver.cpp:67-82): residual and the three scaled Jacobian entries.
template <class Slots>
__global__ void KAssembleEikonal(Slots slots, double* j_eik) {
  BIND_SLOT(blockIdx.y);
  if (Halted(W.state)) return;
  const int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (e >= P.num_eikonal) return;
  const GridParams g = P.grid;
  const int w = P.eik_w[e];
  const int h = P.eik_h[e];
  const double nx = P.eik_nx[e];
  const double ny = P.eik_ny[e];
  const double sw = P.eik_sw[e];
  const double d00 = W.map[h * g.nx + w];
  const double d10 = W.map[h * g.nx + w + 1];
  const double d01 = W.map[(h + 1) * g.nx + w];
  const double grx = (d10 - d00) / g.dx;
  const double gry = (d01 - d00) / g.dy;
  W.r_eik[e] = sw * (1.0 - (nx * grx + ny * gry));
  if (j_eik != nullptr) {
    j_eik[3 * e] = sw * (nx / g.dx + ny / g.dy);
    j_eik[3 * e + 1] = sw * (-nx / g.dx);
    j_eik[3 * e + 2] = sw * (-ny / g.dy);
  }
}

// EvalOdomResidual (residual_math.hpp:182-205), scaled as in the odometry
// tail of SolveLm (lm_solver.cpp:275-296).
template <class Slots>
__global__ void KAssembleOdom(Slots slots) {
  BIND_SLOT(blockIdx.y);
  if (Halted(W.state)) return;
  const int64_t o = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (o >= P.num_odom) return;
  const int i = P.od_i[o];
  const int j = P.od_j[o];
  const double* pi = W.poses + 3 * i;
  const double* pj = W.poses + 3 * j;
  const double c = cos(pi[2]);
  const double s = sin(pi[2]);
  const double dtx = pj[0] - pi[0];
  const double dty = pj[1] - pi[1];
  const double relx = c * dtx + s * dty;
  const double rely = (-s) * dtx + c * dty;
  const double sw = P.od_sw[o];
  W.od_r[3 * o] = sw * (relx - P.od_dx[o]);
  W.od_r[3 * o + 1] = sw * (rely - P.od_dy[o]);
  W.od_r[3 * o + 2] = sw * WrapAngle(pj[2] - pi[2] - P.od_dth[o]);
  double* ji = W.od_ji + 9 * o;
  double* jj = W.od_jj + 9 * o;
  ji[0] = sw * (-c);
  ji[1] = sw * (-s);
  ji[2] = sw * ((-s) * dtx + c * dty);
  ji[3] = sw * s;
  ji[4] = sw * (-c);
  ji[5] = sw * ((-c) * dtx + (-s) * dty);
  ji[6] = sw * 0.0;
  ji[7] = sw * 0.0;
  ji[8] = sw * -1.0;
  jj[0] = sw * c;
  jj[1] = sw * s;
  jj[2] = sw * 0.0;
  jj[3] = sw * (-s);
  jj[4] = sw * c;
  jj[5] = sw * 0.0;
  jj[6] = sw * 0.0;
  jj[7] = sw * 0.0;
  jj[8] = sw * 1.0;
}

// ---------------------------------------------------------------- J^T J

// Per cell: the 10 products of its node weights and the 4 weight x residual
// sums over the cell's points of every frame. Each frame's pose kernel has
// summed its own points per occupied cell; here the cell's candidate frames
// are added in a fixed order: slot s (0..63) sums the candidates s, s + 64,
// ... in order, and the 64 slot sums meet in an xor butterfly (pairs 32
// apart, then 16, 8, 4, 2, 1). Sixteen lanes take a cell (four cells per
// wavefront): lane j holds slots j, j + 16, j + 32 and j + 48, forms the
// butterfly's first two levels itself ((s_j + s_j+32) + (s_j+16 + s_j+48))
// and the last four with its group; the result is the same sum as with a
// wavefront per cell. The four candidates a lane takes per pass have their
// loads issued together.
template <class Slots>
__global__ void __launch_bounds__(kThreads) KCellGather(Slots slots) {
  BIND_SLOT(blockIdx.y);
  if (Halted(W.state)) return;
  const int lane = threadIdx.x & 63;
  const int j = lane & 15;
  const int cell = (blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6)) * 4 + (lane >> 4);
  const bool valid = cell < P.num_cells;
  const int64_t np = P.num_points;
  double acc[4][14];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int k = 0; k < 14; ++k) acc[t][k] = 0.0;
  // Candidate frame q holds the cell at rank cp_rank[q] of its dilated
  // cells; KPoseBlock mapped the rank to the frame's segment of the cell.
  const int64_t q_lo = valid ? A.cp_ptr[cell] : 0;
  const int64_t q_hi = valid ? A.cp_ptr[cell + 1] : 0;
  for (int64_t q0 = q_lo; q0 < q_hi; q0 += 64) {
    int64_t base[4];
    int u[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int64_t q = min(q0 + j + 16 * t, q_hi - 1);
      const int pose = A.cp_pose[q];
      u[t] = W.seg_map[A.dl_ptr[pose] + A.cp_rank[q]];
      base[t] = A.pr_ptr[pose];
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      if (q0 + j + 16 * t >= q_hi || u[t] < 0) continue;
#pragma unroll
      for (int k = 0; k < 14; ++k) acc[t][k] += W.seg_part[k * np + base[t] + u[t]];
    }
  }
  double sum[14];
#pragma unroll
  for (int k = 0; k < 14; ++k) {
    double v = (acc[0][k] + acc[2][k]) + (acc[1][k] + acc[3][k]);
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    sum[k] = v;
  }
  if (valid && j == 0) {
#pragma unroll
    for (int k = 0; k < 10; ++k) W.nn[static_cast<int64_t>(k) * P.num_cells + cell] = sum[k];
#pragma unroll
    for (int k = 0; k < 4; ++k) W.rn[static_cast<int64_t>(k) * P.num_cells + cell] = sum[10 + k];
  }
}

__device__ __forceinline__ double Damp(double diag, double lambda, int marquardt) {
  // lm_solver.cpp:321-330
  return marquardt != 0 ? diag + lambda * fmax(diag, 1e-12) : diag + lambda;
}

// Node-node slots (9-point stencil, lower triangle owned by the node with the
// smaller grid id) and the node entries of -J^T r. One thread per grid node.
// The eikonal rows' J^T J terms at node g's five node-node slots (self,
// right, up-left, up, up-right). The row of spec e at node (w, h) is
// sw * (nx/dx + ny/dy, -nx/dx, -ny/dy) on nodes g, g + 1, g + nx (each where
// active); a slot's terms come from the specs at g, g - 1 and g - nx and are
// summed in spec order, as the reference's product accumulates them
// (lm_solver.cpp:113-172).
__device__ __forceinline__ void EikonalSlots(const DeviceProblem& P, const DeviceAnalysis& A, int g, int w, int h,
                                             double (&out)[5]) {
  const int nx = P.grid.nx;
  const double dx = P.grid.dx;
  const double dy = P.grid.dy;
  auto active = [&](int node) { return P.node_to_active[node] >= 0; };
  // Terms of slot 0 with their spec indices, sorted by spec before summing.
  int te[3];
  double tv[3];
  int nt = 0;
  for (int k = 0; k < 5; ++k) out[k] = 0.0;
  if (h >= 1) {  // spec at g - nx: g is its third node
    const int e = A.eik_at[g - nx];
    if (e >= 0 && active(g)) {
      const double j2 = P.eik_sw[e] * (-P.eik_ny[e] / dy);
      te[nt] = e;
      tv[nt++] = j2 * j2;
    }
  }
  if (w >= 1) {  // spec at g - 1: g is its second node
    const int e = A.eik_at[g - 1];
    if (e >= 0 && active(g)) {
      const double j1 = P.eik_sw[e] * (-P.eik_nx[e] / dx);
      te[nt] = e;
      tv[nt++] = j1 * j1;
      if (active(g - 1 + nx)) {
        const double j2 = P.eik_sw[e] * (-P.eik_ny[e] / dy);
        out[2] = 0.0 + j1 * j2;
      }
    }
  }
  const int e = A.eik_at[g];
  if (e >= 0) {  // spec at g: g is its first node
    const double sw = P.eik_sw[e];
    const double j0 = sw * (P.eik_nx[e] / dx + P.eik_ny[e] / dy);
    const double j1 = sw * (-P.eik_nx[e] / dx);
    const double j2 = sw * (-P.eik_ny[e] / dy);
    if (active(g)) {
      te[nt] = e;
      tv[nt++] = j0 * j0;
      if (active(g + 1)) out[1] = 0.0 + j0 * j1;
      if (active(g + nx)) out[3] = 0.0 + j0 * j2;
    }
  }
  for (int a = 1; a < nt; ++a) {
    for (int b = a; b > 0 && te[b - 1] > te[b]; --b) {
      const int ti = te[b];
      te[b] = te[b - 1];
      te[b - 1] = ti;
      const double tvv = tv[b];
      tv[b] = tv[b - 1];
      tv[b - 1] = tvv;
    }
  }
  for (int a = 0; a < nt; ++a) out[0] += tv[a];
}

template <class Slots>
__global__ void KNodeSlots(Slots slots, int marquardt) {
  BIND_SLOT(blockIdx.y);
  if (Halted(W.state)) return;
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = P.grid.nx;
  const int ny = P.grid.ny;
  if (g >= nx * ny) return;
  const int cu = P.node_to_active[g];
  if (cu < 0) return;
  const int w = g % nx;
  const int h = g / nx;
  const int ncx = nx - 1;
  const int ncy = ny - 1;
  const int nc = P.num_cells;
  const double lambda = W.state->lambda;
  auto ok = [&](int cw, int ch) { return cw >= 0 && ch >= 0 && cw < ncx && ch < ncy; };
  auto nn = [&](int cw, int ch, int t) { return W.nn[static_cast<int64_t>(t) * nc + ch * ncx + cw]; };
  auto rn = [&](int cw, int ch, int t) { return W.rn[static_cast<int64_t>(t) * nc + ch * ncx + cw]; };
  const int64_t* off = A.nn_off + 5 * static_cast<int64_t>(g);
  double eik[5];
  EikonalSlots(P, A, g, w, h, eik);
  // Upper-triangle index of corner pair (a <= b) in the 10 cell products.
  // (0,0)0 (0,1)1 (0,2)2 (0,3)3 (1,1)4 (1,2)5 (1,3)6 (2,2)7 (2,3)8 (3,3)9
  {
    double v = 0.0;
    if (ok(w - 1, h - 1)) v += nn(w - 1, h - 1, 9);
    if (ok(w, h - 1)) v += nn(w, h - 1, 7);
    if (ok(w - 1, h)) v += nn(w - 1, h, 4);
    if (ok(w, h)) v += nn(w, h, 0);
    v += eik[0];
    W.panel[off[0]] = Damp(v, lambda, marquardt);
  }
  if (off[1] >= 0) {  // right neighbour (w+1, h)
    double v = 0.0;
    if (ok(w, h - 1)) v += nn(w, h - 1, 8);
    if (ok(w, h)) v += nn(w, h, 1);
    W.panel[off[1]] = v + eik[1];
  }
  if (off[2] >= 0) {  // up-left neighbour (w-1, h+1)
    double v = 0.0;
    if (ok(w - 1, h)) v += nn(w - 1, h, 5);
    W.panel[off[2]] = v + eik[2];
  }
  if (off[3] >= 0) {  // up neighbour (w, h+1)
    double v = 0.0;
    if (ok(w - 1, h)) v += nn(w - 1, h, 6);
    if (ok(w, h)) v += nn(w, h, 2);
    W.panel[off[3]] = v + eik[3];
  }
  if (off[4] >= 0) {  // up-right neighbour (w+1, h+1)
    double v = 0.0;
    if (ok(w, h)) v += nn(w, h, 3);
    W.panel[off[4]] = v + eik[4];
  }
  double gsum = 0.0;
  if (ok(w - 1, h - 1)) gsum += rn(w - 1, h - 1, 3);
  if (ok(w, h - 1)) gsum += rn(w, h - 1, 2);
  if (ok(w - 1, h)) gsum += rn(w - 1, h, 1);
  if (ok(w, h)) gsum += rn(w, h, 0);
  const double dx = P.grid.dx;
  const double dy = P.grid.dy;
  int e = A.eik_at[g];
  if (e >= 0) gsum += (P.eik_sw[e] * (P.eik_nx[e] / dx + P.eik_ny[e] / dy)) * W.r_eik[e];
  if (w >= 1) {
    e = A.eik_at[g - 1];
    if (e >= 0) gsum += (P.eik_sw[e] * (-P.eik_nx[e] / dx)) * W.r_eik[e];
  }
  if (h >= 1) {
    e = A.eik_at[g - nx];
    if (e >= 0) gsum += (P.eik_sw[e] * (-P.eik_ny[e] / dy)) * W.r_eik[e];
  }
  W.rhs[cu] = -gsum;
}

// The value itself, opaque to the optimizer. A load whose result feeds a
// select on a wave-uniform condition is otherwise moved under a branch on
// that condition, where its data is awaited before the next load issues;
// loads feeding Opaque are issued where they stand.
__device__ __forceinline__ double Opaque(double v) {
  __asm__ volatile("" : "+v"(v));
  return v;
}

// base[byte_off / 8] for a byte offset below 4 GiB: a wave-uniform base and
// a 32-bit per-lane offset, which the load takes as its scalar base and
// vector offset (no 64-bit address per lane).
__device__ __forceinline__ double LoadAt(const double* base, uint32_t byte_off) {
  return *reinterpret_cast<const double*>(reinterpret_cast<const char*>(base) + byte_off);
}

// The wavefront's index in the workgroup, as a wave-uniform value.
__device__ __forceinline__ int WaveIndex() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

// Opaque for a 32-bit integer.
__device__ __forceinline__ uint32_t OpaqueU(uint32_t v) {
  __asm__ volatile("" : "+v"(v));
  return v;
}

// Occupied cells of a frame whose pose sums KPoseBlock keeps in the LDS (12
// per cell, in `red`).
constexpr int kLdsCells = 192;
static_assert(12 * kLdsCells <= 9 * kThreads, "pose sums fit the reduction buffer");

// Pose rows: footprint (pose x node) entries, the pose's 3x3 diagonal block
// and its -J^T r entries. One workgroup per pose; the pose's point rows are
// sorted by cell in LDS (bitonic) so every entry sums its points in a fixed
// order.
// Exclusive scan of one value per thread over a kThreads-thread workgroup
// (Hillis-Steele in `buf`, kThreads + 1 entries): buf[tid] receives the sum
// over threads before tid; returns the total. Contains barriers.
__device__ __forceinline__ int BlockExclusiveScan(int32_t* buf, int v) {
  const int tid = static_cast<int>(threadIdx.x);
  __syncthreads();
  buf[tid] = v;
  __syncthreads();
  for (int off = 1; off < kThreads; off <<= 1) {
    const int t = tid >= off ? buf[tid - off] : 0;
    __syncthreads();
    buf[tid] += t;
    __syncthreads();
  }
  const int total = buf[kThreads - 1];
  const int excl = buf[tid] - v;
  __syncthreads();
  buf[tid] = excl;
  __syncthreads();
  return total;
}

template <class Slots>
__global__ void KPoseBlock(Slots slots, int marquardt, int sort_size, int lds_ok) {
  BIND_SLOT(blockIdx.y);
  if (Halted(W.state)) return;
  const int pose = blockIdx.x;
  if (pose >= P.num_poses) return;
  const int tid = threadIdx.x;
  extern __shared__ unsigned char smem[];
  uint64_t* keys = reinterpret_cast<uint64_t*>(smem);
  int32_t* rows = reinterpret_cast<int32_t*>(keys + sort_size);
  int32_t* ucell = rows + sort_size;
  int32_t* ustart = ucell + sort_size;
  __shared__ int32_t counts[kThreads + 1];
  __shared__ double red[9][kThreads];
  // Until the diagonal block's reduction `red` is scratch: the counting
  // sort's rank bitmap and fill counts, then a bitmap of the frame's
  // occupied cells (grids of up to 9 * kThreads * 64 cells), so footprint
  // nodes with no occupied cell around them skip the searches.
  uint32_t* occupied = reinterpret_cast<uint32_t*>(&red[0][0]);
  const int occ_words = (P.num_cells + 31) / 32;
  const bool use_occ = occ_words <= 9 * kThreads * 2;
  const int64_t rb = A.pr_ptr[pose];
  const int m = static_cast<int>(A.pr_ptr[pose + 1] - rb);
  // This frame's sort size: its row count rounded up to a power of two
  // (the LDS is sized for the largest frame of the launch).
  int ss = 1;
  while (ss < m) ss <<= 1;
  sort_size = min(sort_size, ss);
  const int64_t np = P.num_points;
  const uint32_t invalid = static_cast<uint32_t>(P.num_cells);
  // The frame's rows grouped by cell, cells ascending and rows in their
  // frame order within a cell: order[t] is the local index (into rows) of
  // the t-th grouped row, ucell[u] / ustart[u] the u-th occupied cell and
  // its first position (ustart[num_u] = the valid rows). Rows without a
  // Jacobian row (cell key num_cells) are left out.
  int32_t* order = reinterpret_cast<int32_t*>(keys);
  const int64_t nrank = A.dl_ptr[pose + 1] - A.dl_ptr[pose];
  const int rwords = static_cast<int>((nrank + 31) / 32);
  const int rgroups = (rwords + 7) / 8;  // groups of 8 bitmap words (256 ranks)
  constexpr int kRedWords = 9 * kThreads * 2;
  int num_u = 0;
  if (rwords <= kRedWords && rgroups <= sort_size && sort_size <= kRedWords) {
    // Counting sort over the occupied cells: every valid row's cell has a
    // rank among the frame's dilated cells (a point outside them ends the
    // pattern epoch, and the assembly then halts the iteration before this
    // kernel); the occupied ranks, in a bitmap (in `red`), give each row a
    // dense cell index in ascending cell order (a prefix count per group of
    // 8 words, in order[sort_size ..]); rows go to their cell's range in
    // frame order. ucell holds each row's cell index until the end.
    uint32_t* rbits = reinterpret_cast<uint32_t*>(&red[0][0]);
    int32_t* gpre = order + sort_size;
    int32_t* rowu = ucell;
    const uint32_t* fbits = A.fp_cellbits + static_cast<int64_t>(pose) * A.words_per_pose;
    const int32_t* frank = A.fp_cellrank + static_cast<int64_t>(pose) * A.words_per_pose;
    for (int w = tid; w < rwords; w += kThreads) rbits[w] = 0u;
    __syncthreads();
    // Four rows per thread at a time: their row ids, then their cells, then
    // their bitmap words are loaded together (clamped; rows past m unused).
    for (int base = 0; base < m; base += 4 * kThreads) {
      int32_t rw[4];
      uint32_t ck[4], wd[4];
      int32_t fr[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) rw[e] = A.pr_row[rb + min(base + tid + e * kThreads, m - 1)];
#pragma unroll
      for (int e = 0; e < 4; ++e) ck[e] = W.cellkey[rw[e]];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const uint32_t cc = ck[e] < invalid ? ck[e] : 0u;
        wd[e] = OpaqueU(fbits[cc >> 5]);
        fr[e] = static_cast<int32_t>(OpaqueU(static_cast<uint32_t>(frank[cc >> 5])));
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int i = base + tid + e * kThreads;
        if (i >= m) continue;
        rows[i] = rw[e];
        const uint32_t c = ck[e];
        int r = -1;
        if (c < invalid) {
          const uint32_t bit = 1u << (c & 31);
          if ((wd[e] & bit) != 0u) {
            r = fr[e] + __popc(wd[e] & (bit - 1u));
            atomicOr(&rbits[r >> 5], 1u << (r & 31));
          }
        }
        rowu[i] = r;
      }
    }
    __syncthreads();
    // Exclusive prefix count per group of 8 words (a chunk of groups per
    // thread, then the chunk totals in order).
    const int per = (rgroups + kThreads - 1) / kThreads;
    const int g0 = min(tid * per, rgroups);
    const int g1 = min(g0 + per, rgroups);
    auto group_count = [&](int g) {
      int n = 0;
      for (int w = 8 * g; w < min(8 * g + 8, rwords); ++w) n += __popc(rbits[w]);
      return n;
    };
    int chunk = 0;
    for (int g = g0; g < g1; ++g) chunk += group_count(g);
    num_u = BlockExclusiveScan(counts, chunk);
    int run = counts[tid];
    for (int g = g0; g < g1; ++g) {
      gpre[g] = run;
      run += group_count(g);
    }
    for (int u = tid; u <= num_u; u += kThreads) ustart[u] = 0;
    __syncthreads();
    for (int i = tid; i < m; i += kThreads) {
      const int r = rowu[i];
      if (r < 0) continue;
      int u = gpre[r >> 8];
      for (int w = (r >> 8) * 8; w < (r >> 5); ++w) u += __popc(rbits[w]);
      u += __popc(rbits[r >> 5] & ((1u << (r & 31)) - 1u));
      rowu[i] = u;
      atomicAdd(&ustart[u + 1], 1);
    }
    __syncthreads();
    // Cell starts: the per-cell counts (moved to fill, which takes the
    // bitmap's place) scanned, a chunk of cells per thread; fill is then
    // cleared for the placement.
    int32_t* fill = reinterpret_cast<int32_t*>(&red[0][0]);
    for (int u = tid; u < num_u; u += kThreads) fill[u] = ustart[u + 1];
    __syncthreads();
    {
      const int cper = (num_u + kThreads - 1) / kThreads;
      const int u0 = min(tid * cper, num_u);
      const int u1 = min(u0 + cper, num_u);
      int csum = 0;
      for (int u = u0; u < u1; ++u) csum += fill[u];
      const int rows_valid = BlockExclusiveScan(counts, csum);
      int at = counts[tid];
      for (int u = u0; u < u1; ++u) {
        ustart[u] = at;
        at += fill[u];
      }
      if (tid == 0) ustart[num_u] = rows_valid;
    }
    __syncthreads();
    for (int u = tid; u < num_u; u += kThreads) fill[u] = 0;
    __syncthreads();
    // Rows to their cells in frame order (a row goes after the rows of its
    // cell with smaller frame index). The rows are cut into four runs of
    // 64-row chunks, one per wavefront; a wavefront walks its chunks in
    // order and ranks each row among its cell's rows in the run (one ballot
    // per distinct cell of a chunk, the run's count per cell in wcnt), then
    // every row adds the counts of the earlier runs. The LDS operations of
    // one wavefront complete in order, so a count is read before its
    // increment lands. wcnt (4 num_u counts) follows fill in `red`; the
    // ranks go where the group prefix counts were (order + sort_size).
    int32_t* wcnt = fill + num_u;
    int32_t* rrank = gpre;
    const int nchunk = (m + 63) >> 6;
    const int cper = (nchunk + kWaves - 1) / kWaves;
    const bool waves4 = 5 * num_u <= kRedWords;
    if (waves4) {
      for (int x = tid; x < kWaves * num_u; x += kThreads) wcnt[x] = 0;
      __syncthreads();
      const int wv = tid >> 6;
      const int ln = tid & 63;
      int32_t* cnt = wcnt + wv * num_u;
      for (int c = wv * cper; c < min(nchunk, (wv + 1) * cper); ++c) {
        const int i = 64 * c + ln;
        const int u = i < m ? rowu[i] : -1;
        uint64_t pending = __ballot(u >= 0);
        while (pending != 0ull) {
          const int leader = __ffsll(static_cast<unsigned long long>(pending)) - 1;
          const int v = __builtin_amdgcn_readlane(u, leader);
          const uint64_t same = __ballot(u == v);
          if (u == v) rrank[i] = cnt[v] + __popcll(same & ((1ull << ln) - 1ull));
          if (ln == leader) cnt[v] += __popcll(same);
          pending &= ~same;
        }
      }
      __syncthreads();
      for (int i = tid; i < m; i += kThreads) {
        const int u = rowu[i];
        if (u < 0) continue;
        const int run = (i >> 6) / cper;
        int at = ustart[u] + rrank[i];
        for (int w2 = 0; w2 < run; ++w2) at += wcnt[w2 * num_u + u];
        order[at] = i;
      }
    } else if (tid < 64) {
      // More cells than the counts fit: wavefront 0 alone, fill as the
      // running count.
      for (int base = 0; base < m; base += 64) {
        const int i = base + tid;
        const int u = i < m ? rowu[i] : -1;
        uint64_t pending = __ballot(u >= 0);
        while (pending != 0ull) {
          const int leader = __ffsll(static_cast<unsigned long long>(pending)) - 1;
          const int v = __builtin_amdgcn_readlane(u, leader);
          const uint64_t same = __ballot(u == v);
          if (u == v) order[ustart[v] + fill[v] + __popcll(same & ((1ull << tid) - 1ull))] = i;
          if (tid == leader) fill[v] += __popcll(same);
          pending &= ~same;
        }
      }
    }
    __syncthreads();
    for (int u = tid; u < num_u; u += kThreads) ucell[u] = static_cast<int32_t>(W.cellkey[rows[order[ustart[u]]]]);
    __syncthreads();
  } else {
    // A frame with more dilated cells than the bitmap holds: bitonic sort
    // of (cell, local row) keys.
    for (int i = tid; i < sort_size; i += kThreads) {
      if (i < m) {
        const int32_t row = A.pr_row[rb + i];
        rows[i] = row;
        keys[i] = (static_cast<uint64_t>(W.cellkey[row]) << 32) | static_cast<uint32_t>(i);
      } else {
        keys[i] = ~0ULL;
      }
    }
    __syncthreads();
    for (int k = 2; k <= sort_size; k <<= 1) {
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int i = tid; i < sort_size; i += kThreads) {
          const int l = i ^ j;
          if (l > i) {
            const uint64_t a = keys[i];
            const uint64_t b = keys[l];
            const bool up = (i & k) == 0;
            if ((a > b) == up) {
              keys[i] = b;
              keys[l] = a;
            }
          }
        }
        __syncthreads();
      }
    }
    // Unique valid cells and their start positions (fixed-order compaction).
    const int per = (sort_size + kThreads - 1) / kThreads;
    const int lo = tid * per;
    const int hi = min(lo + per, sort_size);
    int cnt = 0;
    for (int t = lo; t < hi; ++t) {
      const uint32_t c = static_cast<uint32_t>(keys[t] >> 32);
      if (c < invalid && (t == 0 || static_cast<uint32_t>(keys[t - 1] >> 32) != c)) ++cnt;
    }
    counts[tid] = cnt;
    __syncthreads();
    if (tid == 0) {
      int run = 0;
      for (int i = 0; i < kThreads; ++i) {
        const int c = counts[i];
        counts[i] = run;
        run += c;
      }
      counts[kThreads] = run;
    }
    __syncthreads();
    int pos = counts[tid];
    int valid_end_local = -1;
    for (int t = lo; t < hi; ++t) {
      const uint32_t c = static_cast<uint32_t>(keys[t] >> 32);
      if (c < invalid) {
        if (t == 0 || static_cast<uint32_t>(keys[t - 1] >> 32) != c) {
          ucell[pos] = static_cast<int32_t>(c);
          ustart[pos] = t;
          ++pos;
        }
        if (t + 1 == sort_size || static_cast<uint32_t>(keys[t + 1] >> 32) >= invalid) {
          valid_end_local = t + 1;
        }
      }
    }
    num_u = counts[kThreads];
    __syncthreads();
    if (valid_end_local >= 0) ustart[num_u] = valid_end_local;
    if (num_u == 0 && tid == 0) ustart[0] = 0;
    // The keys' local indices as order[], a chunk of kRedWords at a time
    // through `red` (order aliases keys).
    int32_t* tmp = reinterpret_cast<int32_t*>(&red[0][0]);
    for (int c0 = 0; c0 < sort_size; c0 += kRedWords) {
      const int c1 = min(c0 + kRedWords, sort_size);
      for (int t = c0 + tid; t < c1; t += kThreads) tmp[t - c0] = static_cast<int32_t>(keys[t] & 0xffffffffULL);
      __syncthreads();
      for (int t = c0 + tid; t < c1; t += kThreads) order[t] = tmp[t - c0];
      __syncthreads();
    }
  }
  // With at most kLdsCells occupied cells (and room for a half-empty table
  // in the sort keys' second half, free after the sort) the cells' pose
  // sums stay in the LDS (in `red`) and the footprint entries find a cell's
  // segment in an open-addressing table (cell -> segment); otherwise the
  // sums go through global memory and a node's cells are found by binary
  // search behind a bitmap of the occupied cells (in `red`).
  const int hslots = sort_size / 2;
  const bool lds_fp = lds_ok != 0 && num_u <= kLdsCells && hslots >= 2 && 2 * num_u <= hslots;
  int32_t* hkey = order + sort_size;
  int32_t* hval = hkey + hslots;
  const int hshift = 32 - __builtin_ctz(static_cast<unsigned>(max(hslots, 2)));
  auto hslot = [&](int c) { return static_cast<int>((static_cast<uint32_t>(c) * 0x9E3779B1u) >> hshift); };
  double* jp_lds = &red[0][0];
  if (lds_fp) {
    for (int h = tid; h < hslots; h += kThreads) hkey[h] = -1;
    __syncthreads();
    for (int u = tid; u < num_u; u += kThreads) {
      const int c = ucell[u];
      int h = hslot(c);
      while (atomicCAS(&hkey[h], -1, c) != -1) h = (h + 1) & (hslots - 1);
      hval[h] = u;
    }
  } else if (use_occ) {
    for (int i = tid; i < occ_words; i += kThreads) occupied[i] = 0u;
    __syncthreads();
    for (int u = tid; u < num_u; u += kThreads) {
      const int c = ucell[u];
      atomicOr(&occupied[c >> 5], 1u << (c & 31));
    }
  }
  // The frame's dilated cells, by rank: their segment, or -1 where the frame
  // has no row (KCellGather reads these). A cell outside the dilated cells
  // (a point that left the footprint, which ends the pattern epoch) has no
  // rank and is left out.
  int32_t* smap = W.seg_map + A.dl_ptr[pose];
  for (int64_t r = tid; r < nrank; r += kThreads) smap[r] = -1;
  __syncthreads();
  {
    const uint32_t* bits = A.fp_cellbits + static_cast<int64_t>(pose) * A.words_per_pose;
    const int32_t* wrank = A.fp_cellrank + static_cast<int64_t>(pose) * A.words_per_pose;
    for (int u = tid; u < num_u; u += kThreads) {
      const int c = ucell[u];
      const uint32_t word = bits[c >> 5];
      const uint32_t bit = 1u << (c & 31);
      if ((word & bit) == 0u) continue;
      smap[wrank[c >> 5] + __popc(word & (bit - 1u))] = u;
    }
  }
  __syncthreads();

  const double* jp0 = W.jp;
  const double* jp1 = W.jp + np;
  const double* jp2 = W.jp + 2 * np;
  // This frame's sums per occupied cell: the 14 node sums (10 weight
  // products, 4 weight x residual) and the 12 pose-partial x weight sums
  // the footprint entries below add up over the node's cells, all entries
  // of the Gram matrix of the cell's rows x = (w0..w3, r, p0..p2), formed on
  // FP64 MFMA: one 16x16x4 product takes four rows each of two cells (lane
  // l supplies feature l % 8 of row l / 16 of cell 2 p + (l % 16) / 8 as
  // both operands, so D(i, j) = sum_k x_i x_j within each cell's 8 x 8
  // diagonal block; rows past a cell's last are zero). A wavefront takes
  // cell pairs p = wave, wave + 4, ...; every sum is formed in the same
  // order whatever the launch.
  {
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int f16 = lane & 15;
    const int feat = f16 & 7;
    const int kr = lane >> 4;
    const double* fbase = feat < 4 ? W.jw + static_cast<int64_t>(feat) * np
                                   : (feat == 4 ? W.r_pt : W.jp + static_cast<int64_t>(feat - 5) * np);
    for (int pair = wave; 2 * pair < num_u; pair += kWaves) {
      const int ua = 2 * pair;
      const bool has_b = ua + 1 < num_u;
      const int my = f16 < 8 ? ua : (has_b ? ua + 1 : ua);
      const int t0 = ustart[my];
      const int t1 = (f16 < 8 || has_b) ? ustart[my + 1] : t0;
      const int na = ustart[ua + 1] - ustart[ua];
      const int nb = has_b ? ustart[ua + 2] - ustart[ua + 1] : 0;
      const int groups = (max(na, nb) + 3) >> 2;
      v4d g = v4d{0.0, 0.0, 0.0, 0.0};
      for (int q0 = 0; q0 < groups; q0 += 4) {
        double x[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int t = t0 + 4 * (q0 + q) + kr;
          const int tc = max(min(t, t1 - 1), 0);
          x[q] = fbase[rows[order[tc]]];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const double v = Opaque(x[q]);
          const double xv = t0 + 4 * (q0 + q) + kr < t1 ? v : 0.0;
          if (q0 + q < groups) g = __builtin_amdgcn_mfma_f64_16x16x4f64(xv, xv, g, 0, 0, 0);
        }
      }
      // Lane l holds D(4 e + l / 16, l % 16) in g[e]: cell a's block in
      // e = 0, 1 of lanes with l % 16 < 8, cell b's in e = 2, 3 of the rest.
      const bool lane_b = f16 >= 8;
      if (!lane_b || has_b) {
        const int u = lane_b ? ua + 1 : ua;
        const int j = feat;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int i = 4 * h + kr;
          const double v = lane_b ? g[2 + h] : g[h];
          if (i <= j && j < 4) {
            const int k = i * 4 - (i * (i - 1)) / 2 + (j - i);
            W.seg_part[k * np + rb + u] = v;
          } else if (j == 4 && i < 4) {
            W.seg_part[(10 + i) * np + rb + u] = v;
          } else if (i >= 5 && j < 4) {
            const int k = 3 * j + (i - 5);
            if (lds_fp) {
              jp_lds[k * kLdsCells + u] = v;
            } else {
              W.seg_jp[k * np + rb + u] = v;
            }
          }
        }
        if (f16 == 0 || f16 == 8) {
          if (kr == 0) W.seg_cell[rb + u] = ucell[u];
        }
      }
    }
  }
  if (tid == 0) W.seg_count[pose] = num_u;
  if (pose == 0) return;  // frame 0 is fixed: no pose rows

  // Footprint entries: sum over the (up to) four cells around the node of
  // the cell's pose-partial x corner-weight sums (this workgroup wrote them
  // above; the barrier makes them visible).
  __syncthreads();
  const int nx = P.grid.nx;
  const int ncx = nx - 1;
  const int ncy = P.grid.ny - 1;
  if (lds_fp) {
    // Driven by the occupied cells: the nodes with an occupied cell around
    // them are the corners of the occupied cells; item (u, cc) is corner cc
    // of cell u, and the node goes to the item of its first occupied cell
    // in the order below (cells (w-1,h-1), (w,h-1), (w-1,h), (w,h) of node
    // (w, h)). That item sums the node's occupied cells in that order and
    // finds the node's footprint entry by binary search over the pose's
    // footprint nodes (ascending); an inactive node has none. Every other
    // footprint entry keeps the zeros KZeroPanels stored.
    const int64_t q_lo = A.fp_ptr[pose];
    const int64_t q_hi = A.fp_ptr[pose + 1];
    auto find_u = [&](int cw, int ch) {
      if (cw < 0 || ch < 0 || cw >= ncx || ch >= ncy) return -1;
      const int cid = ch * ncx + cw;
      int hh = hslot(cid);
      int key = hkey[hh];
      while (key != cid && key != -1) {
        hh = (hh + 1) & (hslots - 1);
        key = hkey[hh];
      }
      return key == -1 ? -1 : hval[hh];
    };
    for (int x = tid; x < 4 * num_u; x += kThreads) {
      const int cc = x & 3;
      const int c = ucell[x >> 2];
      const int w = c % ncx + (cc & 1);
      const int h = c / ncx + (cc >> 1);
      bool owner = true;
      for (int k2 = 0; k2 < 3 - cc; ++k2) {
        if (find_u(w - 1 + (k2 & 1), h - 1 + (k2 >> 1)) >= 0) {
          owner = false;
          break;
        }
      }
      if (!owner) continue;
      double v0 = 0.0, v1 = 0.0, v2 = 0.0;
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) {
        const int uu = find_u(w - 1 + (k2 & 1), h - 1 + (k2 >> 1));
        if (uu < 0) continue;
        const double* sp = jp_lds + 3 * (3 - k2) * kLdsCells + uu;
        v0 += sp[0];
        v1 += sp[kLdsCells];
        v2 += sp[2 * kLdsCells];
      }
      const int g = h * nx + w;
      int64_t lo = q_lo, hi = q_hi;
      while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (A.fp_node[mid] < g) {
          lo = mid + 1;
        } else {
          hi = mid;
        }
      }
      if (lo >= q_hi || A.fp_node[lo] != g) continue;
      const int64_t o = A.fp_off[lo];
      const int64_t st = A.fp_stride[lo];
      W.panel[o] = v0;
      W.panel[o + st] = v1;
      W.panel[o + 2 * st] = v2;
    }
  }
  for (int64_t q = A.fp_ptr[pose] + tid; !lds_fp && q < A.fp_ptr[pose + 1]; q += kThreads) {
    const int g = A.fp_node[q];
    const int w = g % nx;
    const int h = g / nx;
    double v0 = 0.0, v1 = 0.0, v2 = 0.0;
    bool any = false;
    const int cws[4] = {w - 1, w, w - 1, w};
    const int chs[4] = {h - 1, h - 1, h, h};
    const int corner[4] = {3, 2, 1, 0};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int cw = cws[k];
      const int ch = chs[k];
      if (cw < 0 || ch < 0 || cw >= ncx || ch >= ncy) continue;
      const int cid = ch * ncx + cw;
      if (use_occ && ((occupied[cid >> 5] >> (cid & 31)) & 1u) == 0u) continue;
      int lo2 = 0, hi2 = num_u;
      while (lo2 < hi2) {
        const int mid = (lo2 + hi2) >> 1;
        if (ucell[mid] < cid) lo2 = mid + 1; else hi2 = mid;
      }
      if (lo2 >= num_u || ucell[lo2] != cid) continue;
      const double* sp = W.seg_jp + static_cast<int64_t>(3 * corner[k]) * np + rb + lo2;
      v0 += sp[0];
      v1 += sp[np];
      v2 += sp[2 * np];
      any = true;
    }
    if (!any) continue;  // the zeros KZeroPanels stored
    const int64_t o = A.fp_off[q];
    const int64_t st = A.fp_stride[q];
    W.panel[o] = v0;
    W.panel[o + st] = v1;
    W.panel[o + 2 * st] = v2;
  }
  // Diagonal block and -J^T r over the pose's rows (strided, then a fixed
  // tree); `red` is the occupied-cell bitmap until every thread is past the
  // footprint entries.
  __syncthreads();
  double acc[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) acc[k] = 0.0;
  for (int base = 0; base < m; base += 4 * kThreads) {
    // Four rows per thread: every load issued first (clamped), then the
    // sums in row order.
    int rw[4];
    uint32_t ck[4];
    double la[4][3], lr[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) rw[e] = rows[min(base + tid + e * kThreads, m - 1)];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      ck[e] = W.cellkey[rw[e]];
      la[e][0] = jp0[rw[e]];
      la[e][1] = jp1[rw[e]];
      la[e][2] = jp2[rw[e]];
      lr[e] = W.r_pt[rw[e]];
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
    const double a0 = Opaque(la[e][0]);
    const double a1 = Opaque(la[e][1]);
    const double a2 = Opaque(la[e][2]);
    const double r = Opaque(lr[e]);
    if (base + tid + e * kThreads >= m) break;
    if (ck[e] >= invalid) continue;
    acc[0] += a0 * a0;
    acc[1] += a1 * a0;
    acc[2] += a1 * a1;
    acc[3] += a2 * a0;
    acc[4] += a2 * a1;
    acc[5] += a2 * a2;
    acc[6] += a0 * r;
    acc[7] += a1 * r;
    acc[8] += a2 * r;
    }
  }
#pragma unroll
  for (int k = 0; k < 9; ++k) red[k][tid] = acc[k];
  __syncthreads();
  for (int stride = kThreads / 2; stride > 0; stride >>= 1) {
    if (tid < stride) {
#pragma unroll
      for (int k = 0; k < 9; ++k) red[k][tid] += red[k][tid + stride];
    }
    __syncthreads();
  }
  if (tid == 0) {
    double d[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) d[k] = red[k][0];
    // Odometry/relation rows touching this pose, in spec order.
    for (int64_t q = A.po_ptr[pose]; q < A.po_ptr[pose + 1]; ++q) {
      const int o = A.po_spec[q];
      const double* J = (P.od_i[o] == pose) ? W.od_ji + 9 * static_cast<int64_t>(o)
                                            : W.od_jj + 9 * static_cast<int64_t>(o);
      const double* r = W.od_r + 3 * static_cast<int64_t>(o);
      const int ij[6][2] = {{0, 0}, {1, 0}, {1, 1}, {2, 0}, {2, 1}, {2, 2}};
      for (int k = 0; k < 6; ++k) {
        const int a = ij[k][0];
        const int b = ij[k][1];
        d[k] += J[a] * J[b] + J[3 + a] * J[3 + b] + J[6 + a] * J[6 + b];
      }
      for (int a = 0; a < 3; ++a) {
        d[6 + a] += J[a] * r[0] + J[3 + a] * r[1] + J[6 + a] * r[2];
      }
    }
    const double lambda = W.state->lambda;
    d[0] = Damp(d[0], lambda, marquardt);
    d[2] = Damp(d[2], lambda, marquardt);
    d[5] = Damp(d[5], lambda, marquardt);
    const int64_t* po = A.pd_off + 6 * static_cast<int64_t>(pose);
#pragma unroll
    for (int k = 0; k < 6; ++k) W.panel[po[k]] = d[k];
    const int pc = P.num_active + 3 * (pose - 1);
    W.rhs[pc] = -d[6];
    W.rhs[pc + 1] = -d[7];
    W.rhs[pc + 2] = -d[8];
  }
}

// Off-diagonal pose-pose blocks from odometry/relation rows (spec order).
template <class Slots>
__global__ void KOdomPairs(Slots slots) {
  BIND_SLOT(blockIdx.y);
  if (Halted(W.state)) return;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= A.num_pairs) return;
  const int pp = A.op_pose[2 * k];
  double v[9];
#pragma unroll
  for (int t = 0; t < 9; ++t) v[t] = 0.0;
  for (int64_t q = A.op_ptr[k]; q < A.op_ptr[k + 1]; ++q) {
    const int o = A.op_spec[q];
    const bool p_is_i = P.od_i[o] == pp;
    const double* JP = p_is_i ? W.od_ji + 9 * static_cast<int64_t>(o) : W.od_jj + 9 * static_cast<int64_t>(o);
    const double* JQ = p_is_i ? W.od_jj + 9 * static_cast<int64_t>(o) : W.od_ji + 9 * static_cast<int64_t>(o);
    for (int a = 0; a < 3; ++a) {
      for (int b = 0; b < 3; ++b) {
        v[3 * a + b] += JP[a] * JQ[b] + JP[3 + a] * JQ[3 + b] + JP[6 + a] * JQ[6 + b];
      }
    }
  }
#pragma unroll
  for (int t = 0; t < 9; ++t) W.panel[A.op_off[9 * static_cast<int64_t>(k) + t]] = v[t];
}

// ---------------------------------------------------------------- factorization

// dst[col + rel[a] - off] += src[a] for a = b + lane, b + lane + 64, ... < mc
// (column b of a child's update matrix, one wavefront). In groups of four
// per lane: the group's list and source loads are issued together, then its
// destination loads (at indices clamped into the lists), then the sums are
// stored for a < mc; a per-element loop would wait on two loads per element.
__device__ __forceinline__ void ScatterAddColumn(double* dst, int64_t col, int off, const int32_t* rel,
                                                 const double* src, int b, int lane, int mc) {
  // Up to kPer entries per lane per round (1,024 rows of the column): the
  // map and source loads of a round are all issued (clamped) before the
  // destination loads, so a round costs two memory latencies.
  constexpr int kPer = 4;
  for (int g = b; g < mc; g += kPer * 64) {
    int r[kPer];
    double u[kPer], d[kPer];
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int a = min(g + lane + 64 * q, mc - 1);
      r[q] = rel[a];
      u[q] = src[a];
    }
#pragma unroll
    for (int q = 0; q < kPer; ++q) d[q] = dst[col + r[q] - off];
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      if (g + lane + 64 * q < mc) dst[col + r[q] - off] = d[q] + u[q];
    }
  }
}

// Extend-add into the panels of the level's fronts: the columns of their
// children's update matrices that land in a front's own columns (each
// child's columns [0, split); the rest are gathered by the front's
// update-matrix kernel). One wavefront per own column of a front: it adds
// the column of each child that lands there (at most one per child, found
// through the child's map from the parent's own columns to its own), child
// by child in ascending order, so no two wavefronts write the same entry.
template <class Slots, class Lists>
__global__ void KExtendAddPanel(Slots slots, Lists plan) {
  BIND_SLOT(blockIdx.z);
  if (Halted(W.state)) return;
  const int p = PlanItem(plan, blockIdx.z, blockIdx.y);
  if (p < 0) return;
  const int j = static_cast<int>(blockIdx.x) * kWaves + static_cast<int>(threadIdx.x >> 6);
  if (j >= A.sn_first[p + 1] - A.sn_first[p]) return;
  const int fp = A.sn_f[p];
  double* fpanel = W.panel + A.panel_off[p];
  const int64_t col = static_cast<int64_t>(j) * fp;
  for (int64_t ci = A.children_ptr[p]; ci < A.children_ptr[p + 1]; ++ci) {
    const int c = A.children[ci];
    const int64_t poff = A.pinv_off[c];
    if (poff < 0) continue;
    const int b = A.pinv[poff + j];
    if (b < 0) continue;
    const int mc = A.sn_m[c];
    ScatterAddColumn(fpanel, col, 0, A.rel + A.rel_off[c], W.upd + A.upd_off[c] + static_cast<int64_t>(b) * mc, b,
                     static_cast<int>(threadIdx.x & 63), mc);
    __threadfence_block();  // this child's stores before the next child's loads of the column
  }
}

// Fronts up to this many rows take the LDS form of the panel extend-add.
constexpr int kExtendLdsRows = 2048;
static_assert(kWaves * kExtendLdsRows * 8 <= 65536, "the LDS form fits the default LDS limit");

// KExtendAddPanel with the parent's column staged in the LDS (one column of
// at most ldf rows per wavefront): the column is read once (coalesced), the
// children's maps from the parent's own columns are read together, then
// each child's entries (rows b .. mc - 1 of its column b, in rounds of 1024:
// the round's map and value loads of two children issued together) are
// added in the LDS child by child in ascending order, and the column is
// written back once. The same additions in the same order as the global
// form; a column's wavefront has no other writer, so no barrier is needed
// (the LDS operations of one wavefront complete in order).
template <class Slots, class Lists>
__global__ void __launch_bounds__(kThreads) KExtendAddPanelLds(Slots slots, Lists plan, int ldf) {
  BIND_SLOT(blockIdx.z);
  if (Halted(W.state)) return;
  const int p = PlanItem(plan, blockIdx.z, blockIdx.y);
  if (p < 0) return;
  const int wave = static_cast<int>(threadIdx.x >> 6);
  const int lane = static_cast<int>(threadIdx.x & 63);
  const int j = static_cast<int>(blockIdx.x) * kWaves + wave;
  if (j >= A.sn_first[p + 1] - A.sn_first[p]) return;
  const int fp = A.sn_f[p];
  double* fcol = W.panel + A.panel_off[p] + static_cast<int64_t>(j) * fp;
  extern __shared__ double ext_lds[];
  double* colv = ext_lds + wave * ldf;
  constexpr int kPer = 16;  // entries per lane per round
  for (int i0 = 0; i0 < fp; i0 += kPer * 64) {
    double v[kPer];
#pragma unroll
    for (int q = 0; q < kPer; ++q) v[q] = fcol[min(i0 + lane + 64 * q, fp - 1)];
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      if (i0 + lane + 64 * q < fp) colv[i0 + lane + 64 * q] = v[q];
    }
  }
  const int64_t c_lo = A.children_ptr[p];
  const int64_t c_hi = A.children_ptr[p + 1];
  for (int64_t ci = c_lo; ci < c_hi; ci += 2) {
    // Two children at a time: their columns (or none) and their entries.
    int b[2], mc[2];
    const int32_t* rel[2];
    const double* src[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      b[h] = -1;
      mc[h] = 0;
      rel[h] = nullptr;
      src[h] = nullptr;
      if (ci + h < c_hi) {
        const int c = A.children[ci + h];
        const int64_t poff = A.pinv_off[c];
        if (poff >= 0) {
          b[h] = A.pinv[poff + j];
          mc[h] = A.sn_m[c];
          rel[h] = A.rel + A.rel_off[c];
          src[h] = W.upd + A.upd_off[c] + static_cast<int64_t>(max(b[h], 0)) * mc[h];
        }
      }
    }
    const int span = max(b[0] >= 0 ? mc[0] - b[0] : 0, b[1] >= 0 ? mc[1] - b[1] : 0);
    for (int g = 0; g < span; g += kPer * 64) {
      int r[2][kPer];
      double u[2][kPer];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (b[h] < 0) continue;
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
          const int a = min(b[h] + g + lane + 64 * q, mc[h] - 1);
          r[h][q] = rel[h][a];
          u[h][q] = src[h][a];
        }
      }
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (b[h] < 0) continue;
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
          if (b[h] + g + lane + 64 * q < mc[h]) colv[r[h][q]] = colv[r[h][q]] + u[h][q];
        }
      }
    }
  }
  for (int i0 = 0; i0 < fp; i0 += kPer * 64) {
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int i = i0 + lane + 64 * q;
      if (i < fp) fcol[i] = colv[i];
    }
  }
}

// Row stride (doubles) of the panel step's LDS tiles: two doubles of
// padding put the four k rows an MFMA operand read touches in different
// banks.
constexpr int kLdT = kNB + 2;

// Staging slot of supernode s's factored diagonal block of step kb: two
// per supernode, alternating with the step, so the block of step kb + 1
// can be written while step kb's is still being read.
__device__ __forceinline__ double* StageSlot(const DeviceWork& W, int s, int kb) {
  return W.diag_stage + (2 * static_cast<int64_t>(s) + (kb & 1)) * kNB * kNB;
}

// L_kk^-1 of supernode s's kb-th diagonal block (column-major 64 x 64,
// identity beyond the block's nb columns), kept for the solves.
__device__ __forceinline__ double* DiagInv(const DeviceAnalysis& A, const DeviceWork& W, int s, int kb) {
  return W.diag_inv + A.inv_off[s] + static_cast<int64_t>(kb) * kNB * kNB;
}

// D of supernode s's kb-th diagonal block (64 entries, padded as the
// inverses), kept for the update matrices.
__device__ __forceinline__ double* DiagD(const DeviceAnalysis& A, const DeviceWork& W, int s, int kb) {
  return W.diag_d + A.inv_off[s] / (kNB * kNB) * kNB + static_cast<int64_t>(kb) * kNB;
}

// Reads lane `lane`'s value of v (a wave-uniform lane index).
__device__ __forceinline__ double ReadLane(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), lane);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), lane);
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<long long>(static_cast<unsigned>(lo)));
}

// 1 / d: hardware reciprocal estimate refined by two Newton steps (each
// doubles the correct bits; the result is within an ulp of the quotient).
__device__ __forceinline__ double Recip(double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = fma(fma(-d, r, 1.0), r, r);
  r = fma(fma(-d, r, 1.0), r, r);
  return r;
}

// Pivots kC .. kC + 15 of the blocked LDL^T of a 64x64 block, one
// wavefront, row `lane` of the panel in r[j] = T(lane, kC + j): right-looking
// over the panel's 16 columns only. Step k: every lane broadcasts its
// unscaled c = r[k] through the LDS vector `col` while the pivot d = c of
// lane kC + k comes with v_readlane, so the broadcast overlaps the
// reciprocal (v_rcp_f64 and two Newton steps); then r[j] -= (c / d) T(kC + j,
// kC + k) for j > k. Step k + 1's column is updated first and its broadcast,
// pivot and reciprocal issued before the rest of step k's update. LDS
// operations of one wavefront complete in order, so no barrier is needed.
// On return r[j] holds the multiplier L(lane, kC + j) (scratch where
// lane <= kC + j) and dvec[kC + k] = D(kC + k).
template <int kC>
__device__ __forceinline__ void PanelPivots(double (&r)[16], int lane, double* col, double* dvec, bool& zero) {
  double c_next = r[0];
  col[lane] = c_next;
  double d_next = ReadLane(c_next, kC);
  double r_next = Recip(d_next);
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const double c = c_next;
    const double d = d_next;
    zero = zero || d == 0.0;
    dvec[kC + k] = d;
    const double l = c * r_next;
    r[k] = l;
    double cj[16];
#pragma unroll
    for (int j = k + 1; j < 16; ++j) cj[j] = col[kC + j];
    if (k + 1 < 16) {
      r[k + 1] = fma(-l, cj[k + 1], r[k + 1]);
      c_next = r[k + 1];
      col[lane] = c_next;
      d_next = ReadLane(c_next, kC + k + 1);
      r_next = Recip(d_next);
    }
#pragma unroll
    for (int j = k + 2; j < 16; ++j) r[j] = fma(-l, cj[j], r[j]);
  }
}

// Panel kP (columns 16 kP .. + 15) of the blocked LDL^T, wave 0: loads the
// panel's rows (identity pad beyond nb), runs its pivots and writes the
// multipliers back over the panel (every row: below the diagonal they are
// L, on and above it scratch) and 1 / D on the diagonal.
template <int kLd, int kP>
__device__ __forceinline__ void FactorPanel(double* lt, double* dvec, double* col, int nb, int lane, bool& zero) {
  constexpr int c = 16 * kP;
  double r[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const double v = lt[(c + j) * kLd + lane];
    r[j] = (lane < nb && c + j < nb && c + j <= lane) ? v : (c + j == lane ? 1.0 : 0.0);
  }
  PanelPivots<c>(r, lane, col, dvec, zero);
#pragma unroll
  for (int j = 0; j < 16; ++j) lt[(c + j) * kLd + lane] = r[j];
  if (lane >= c && lane < c + 16) lt[lane * kLd + lane] = 1.0 / dvec[lane];
}

// Trailing update after panel kP: every lower 16x16 block (I, J),
// kP < J <= I < np, loses L_I,kP D_kP L_J,kP^T, one FP64 MFMA chain of
// four (16x16x4: lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16] and
// holds D[4 e + l / 16][l % 16] in element e); the blocks go round-robin to
// the wavefronts.
template <int kLd, int kP>
__device__ __forceinline__ void UpdatePanel(double* lt, const double* dvec, int np, int wave, int lane) {
  constexpr int c = 16 * kP;
  int b = 0;
#pragma unroll
  for (int I = kP + 1; I < 4; ++I) {
#pragma unroll
    for (int J = kP + 1; J <= I; ++J, ++b) {
      if (I >= np || (b & (kWaves - 1)) != wave) continue;
      v4d acc = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int kc = 0; kc < 16; kc += 4) {
        const int t = c + kc + (lane >> 4);
        const double a = lt[t * kLd + 16 * I + (lane & 15)];
        const double bl = lt[t * kLd + 16 * J + (lane & 15)] * dvec[t];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bl, acc, 0, 0, 0);
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        double* x = lt + (16 * J + (lane & 15)) * kLd + 16 * I + 4 * e + (lane >> 4);
        *x = *x - acc[e];
      }
    }
  }
}

// M_pp = L_pp^-1 of the 16x16 unit lower diagonal block at `base`, lanes
// 0 .. 15 of a wavefront (lane c forms column c by forward substitution);
// M(r, c) goes to lt[r * kLd + c] (r > c), the other triangle of the tile.
// Row j of the column is y_j = -sum_t L(j, t) y_t over t = 0 .. j - 1 in
// ascending order (one FMA chain per row, rows in order), so only one row
// of L is read at a time.
template <int kLd>
__device__ __forceinline__ void InvertDiag16(double* lt, int base, int c) {
  double y[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    double acc = j == c ? 1.0 : 0.0;
#pragma unroll
    for (int t = 0; t < j; ++t) acc = fma(-lt[(base + t) * kLd + base + j], y[t], acc);
    y[j] = acc;
    __asm__ volatile("" ::: "memory");
  }
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    if (j > c) lt[(base + j) * kLd + base + c] = y[j];
  }
}

// M(r, c) of a block on or below the diagonal of the inverse (unit
// diagonal, zero above).
template <int kLd>
__device__ __forceinline__ double InvAt(const double* lt, int r, int c) {
  const double v = lt[r * kLd + c];
  return r > c ? v : (r == c ? 1.0 : 0.0);
}

// Off-diagonal block M_pq = -M_pp (sum_{k = q}^{p - 1} L_pk M_kq) of the
// inverse, one wavefront, two FP64 MFMA chains: the accumulator of the sum
// (lane l holds S[4 e + l / 16][l % 16] in element e) is already the B
// operand of the product with M_pp. Needs M_pp and every M_kq, k < p.
template <int kLd>
__device__ __forceinline__ void InvertOffDiag(double* lt, int p, int q, int lane) {
  const int lr = lane & 15;  // A row / B column within a block
  const int lk = lane >> 4;  // k within a 4-wide chunk
  v4d sum = v4d{0.0, 0.0, 0.0, 0.0};
  for (int k = q; k < p; ++k) {
    double a[4], b[4];
#pragma unroll
    for (int x = 0; x < 4; ++x) {
      a[x] = lt[(16 * k + 4 * x + lk) * kLd + 16 * p + lr];  // L_pk
      b[x] = InvAt<kLd>(lt, 16 * k + 4 * x + lk, 16 * q + lr);  // M_kq
    }
#pragma unroll
    for (int x = 0; x < 4; ++x) sum = __builtin_amdgcn_mfma_f64_16x16x4f64(a[x], b[x], sum, 0, 0, 0);
  }
  double mp[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) mp[e] = InvAt<kLd>(lt, 16 * p + lr, 16 * p + 4 * e + lk);  // M_pp
  v4d prod = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int e = 0; e < 4; ++e) prod = __builtin_amdgcn_mfma_f64_16x16x4f64(mp[e], sum[e], prod, 0, 0, 0);
#pragma unroll
  for (int e = 0; e < 4; ++e) lt[(16 * p + 4 * e + lk) * kLd + 16 * q + lr] = -prod[e];
}

// LDL^T of a 64x64 block and M = L^-1 of its unit lower triangle by a
// whole workgroup (kThreads threads; contains barriers) in the LDS tile lt,
// column-major kLd apart: on entry T(i, j) at lt[j * kLd + i] for
// j <= i < nb (everything else is ignored; rows and columns from nb on are
// an identity pad). Blocked by 16 columns: wave 0 runs a panel's pivots
// (PanelPivots), then every wavefront takes lower 16x16 blocks of the
// trailing update on FP64 MFMA. The inverse's diagonal block of a panel is
// formed by wave 1 while wave 0 factors the next panel, and a block row's
// off-diagonal blocks by wave 2 one panel later; after the last panel its
// diagonal block (wave 0) and the row before (wave 1), then its own row
// (one block per wavefront). Panels past nb are written as identity.
// On return lt[j * kLd + j] = 1 / D(j), lt[j * kLd + i] = L(i, j) and
// lt[i * kLd + j] = M(i, j) for i > j (so W = L^-T D^-1 has W(i, j) =
// M(j, i) / D(j) above the diagonal and 1 / D(j) on it), dvec[j] = D(j).
// `col` is kNB doubles of scratch. Returns, in wave 0, true when a pivot
// is exactly zero. Ends with a barrier.
template <int kLd>
__device__ __forceinline__ bool FactorInvertBlock(double* lt, double* dvec, double* col, int nb) {
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int np = (nb + 15) >> 4;  // panels with a column inside the block
  bool zero = false;
  for (int idx = tid; idx < kNB * kNB; idx += kThreads) {
    const int i = idx & (kNB - 1);
    const int j = idx / kNB;
    if (j >= 16 * np) lt[j * kLd + i] = i == j ? 1.0 : 0.0;
  }
  if (tid >= 16 * np && tid < kNB) dvec[tid] = 1.0;
  if (wave == 0) FactorPanel<kLd, 0>(lt, dvec, col, nb, lane, zero);
  __syncthreads();
  if (np > 1) {
    UpdatePanel<kLd, 0>(lt, dvec, np, wave, lane);
    __syncthreads();
    if (wave == 0) {
      FactorPanel<kLd, 1>(lt, dvec, col, nb, lane, zero);
    } else if (wave == 1 && lane < 16) {
      InvertDiag16<kLd>(lt, 0, lane);
    }
    __syncthreads();
  }
  if (np > 2) {
    UpdatePanel<kLd, 1>(lt, dvec, np, wave, lane);
    __syncthreads();
    if (wave == 0) {
      FactorPanel<kLd, 2>(lt, dvec, col, nb, lane, zero);
    } else if (wave == 1 && lane < 16) {
      InvertDiag16<kLd>(lt, 16, lane);
    }
    __syncthreads();
  }
  if (np > 3) {
    UpdatePanel<kLd, 2>(lt, dvec, np, wave, lane);
    __syncthreads();
    if (wave == 0) {
      FactorPanel<kLd, 3>(lt, dvec, col, nb, lane, zero);
    } else if (wave == 1 && lane < 16) {
      InvertDiag16<kLd>(lt, 32, lane);
    } else if (wave == 2) {
      InvertOffDiag<kLd>(lt, 1, 0, lane);
    }
    __syncthreads();
  }
  const int last = np - 1;
  if (wave == 0 && lane < 16) {
    InvertDiag16<kLd>(lt, 16 * last, lane);
  } else if (wave == 1 && last >= 2) {
    for (int q = 0; q < last - 1; ++q) InvertOffDiag<kLd>(lt, last - 1, q, lane);
  }
  __syncthreads();
  if (wave < last) InvertOffDiag<kLd>(lt, last, wave, lane);
  __syncthreads();
  return zero;
}

// Stores a block left by FactorInvertBlock, kN threads cooperating (thread
// tid takes entries tid, tid + kN, ... in column-major order): the staging
// slot (L below the diagonal, D on it, W above; identity padding beyond nb
// in both), D alone into `d`, and, with `inv`, M itself (unit diagonal,
// zeros above).
template <int kLd, int kN>
__device__ __forceinline__ void StoreFactored(const double* lt, const double* dvec, double* stage, double* d,
                                              double* inv, int tid) {
  if (tid < kNB) d[tid] = dvec[tid];
#pragma unroll 4
  for (int e = 0; e < kNB * kNB / kN; ++e) {
    const int idx = tid + kN * e;
    const int i = idx % kNB;
    const int j = idx / kNB;
    const double v = lt[j * kLd + i];
    stage[idx] = i > j ? v : (i < j ? v * lt[j * kLd + j] : dvec[j]);
    if (inv != nullptr) {
      const double m = lt[i * kLd + j];
      inv[idx] = i > j ? m : (i == j ? 1.0 : 0.0);
    }
  }
}

// The kb-th diagonal block of front s into the LDS tile lt (column-major,
// kLd apart; entries past nb hold copies of the block's last row or
// column, which FactorInvertBlock ignores), the whole workgroup; every load
// is issued (clamped into the block), so they are in flight together.
template <int kLd>
__device__ __forceinline__ void LoadDiagBlock(const double* F, int f, int c0, int nb, double* lt) {
  constexpr int kPer = kNB * kNB / kThreads;
  double v[kPer];
#pragma unroll
  for (int e = 0; e < kPer; ++e) {
    const int idx = threadIdx.x + kThreads * e;
    const int i = min(idx & (kNB - 1), nb - 1);
    const int j = min(idx / kNB, nb - 1);
    v[e] = F[static_cast<int64_t>(c0 + j) * f + c0 + i];
  }
#pragma unroll
  for (int e = 0; e < kPer; ++e) {
    const int idx = threadIdx.x + kThreads * e;
    lt[(idx / kNB) * kLd + (idx & (kNB - 1))] = v[e];
  }
}

// The kb-th diagonal block of every front at the level, one workgroup per
// front: factored and inverted in the LDS (FactorInvertBlock) and stored in
// the front's staging slot (with the zero-pivot flag), for KPanelStep's
// staged mode.
// A single problem's kernel has its registers capped at two workgroups per
// compute unit (its leaf level has hundreds of fronts); the slot-table form
// keeps the larger budget it needs without spilling.
template <class Slots>
constexpr int kDiagBlockOccupancy = std::is_same<Slots, OneSlot>::value ? 2 : 1;

template <class Slots, class Lists>
__global__ void __launch_bounds__(kThreads, kDiagBlockOccupancy<Slots>) KDiagBlock(Slots slots, Lists plan, int kb) {
  BIND_SLOT(blockIdx.y);
  if (Halted(W.state)) return;
  const int s = PlanItem(plan, blockIdx.y, blockIdx.x);
  if (s < 0) return;
  const int ns = A.sn_first[s + 1] - A.sn_first[s];
  const int c0 = kb * kNB;
  if (c0 >= ns) return;
  const int nb = min(kNB, ns - c0);
  const int f = A.sn_f[s];
  const double* F = W.panel + A.panel_off[s];
  const int tid = threadIdx.x;
  constexpr int kLd = kNB + 1;
  __shared__ double lt[kNB * kLd + 2 * kNB];  // the tile, D, scratch
  double* dvec = lt + kNB * kLd;
  LoadDiagBlock<kLd>(F, f, c0, nb, lt);
  __syncthreads();
  const bool zero = FactorInvertBlock<kLd>(lt, dvec, dvec + kNB, nb);
  StoreFactored<kLd, kThreads>(lt, dvec, StageSlot(W, s, kb), DiagD(A, W, s, kb), DiagInv(A, W, s, kb), tid);
  if (tid == 0 && zero) {
    W.state->not_spd = 1;
    W.state->halt = 1;
  }
}

// The first 64-column step of the blocked panel factorization of every
// front at the level (later steps are fused into the previous step's
// trailing update, KGemmAhead): the panel rows below the diagonal block, 64
// per workgroup, become L21 = A21 W with W = L_kk^-T D_k^-1, a
// 64 x 64 x 64 FP64 MFMA product. W is read from the staging slot
// (KDiagBlock filled it) or, with kFused, formed by every workgroup,
// which factors and inverts the block itself; workgroup 0 then stores the
// staging slot. The trailing-update kernel moves the factored
// block into the panel, so no workgroup of this launch reads a partly
// written block.
template <bool kFused, class Slots, class Lists>
__global__ void __launch_bounds__(kThreads) KPanelStep(Slots slots, Lists plan, int kb) {
  BIND_SLOT(blockIdx.z);
  if (Halted(W.state)) return;
  const int s = PlanItem(plan, blockIdx.z, blockIdx.y);
  if (s < 0) return;
  const int ns = A.sn_first[s + 1] - A.sn_first[s];
  const int c0 = kb * kNB;
  if (c0 >= ns) return;
  const int nb = min(kNB, ns - c0);
  const int f = A.sn_f[s];
  const int r0 = c0 + nb + static_cast<int>(blockIdx.x) * kTile;  // first row of this workgroup
  constexpr bool fused = kFused;
  const bool owner = fused && blockIdx.x == 0;  // stores the staging slot
  const bool rows = r0 < f;
  if (!rows && !owner) return;
  double* F = W.panel + A.panel_off[s];
  __shared__ double wt[kNB][kLdT];  // wt[k][j] = W(k, j)
  __shared__ double at[kNB][kLdT];  // at[k][r] = F(r0 + r, c0 + k)
  __shared__ double dsh[kNB];  // D, with kFused
  __shared__ double colscr[kFused ? kNB : 1];  // the factorization's scratch
  __shared__ int zero_pivot;
  const int tid = threadIdx.x;
  // A21: rows r0 .. r0 + 63 of the block column, zero beyond the front and
  // beyond nb (W's padding is identity). Loaded at clamped addresses into
  // registers first, so the loads are in flight together (with the staged
  // W's), then selected into the LDS.
  constexpr int kPer = kNB * kTile / kThreads;
  double av[kPer];
  auto load_rows = [&]() {
#pragma unroll
    for (int e = 0; e < kPer; ++e) {
      const int idx = tid + kThreads * e;
      const int r = idx % kTile;
      const int k = idx / kTile;
      av[e] = F[static_cast<int64_t>(c0 + min(k, nb - 1)) * f + min(r0 + r, f - 1)];
    }
  };
  auto store_rows = [&]() {
#pragma unroll
    for (int e = 0; e < kPer; ++e) {
      const int idx = tid + kThreads * e;
      const int r = idx % kTile;
      const int k = idx / kTile;
      const double x = Opaque(av[e]);
      at[k][r] = (k < nb && r0 + r < f) ? x : 0.0;
    }
  };
  if constexpr (!fused) {
    load_rows();  // every workgroup left here has rows
    // W from the staging slot: this thread's entries are (i, j) = (tid % 64,
    // tid / 64 + 4 e); at most one of them is diagonal, so one division.
    const double* stage = StageSlot(W, s, kb);
    double v[kPer];
#pragma unroll
    for (int e = 0; e < kPer; ++e) v[e] = stage[tid + kThreads * e];
    store_rows();
    const int i = tid % kNB;
    double dv = 1.0;
#pragma unroll
    for (int e = 0; e < kPer; ++e) dv = tid / kNB + 4 * e == i ? v[e] : dv;
    const double rd = 1.0 / dv;
#pragma unroll
    for (int e = 0; e < kPer; ++e) {
      const int j = tid / kNB + 4 * e;
      wt[i][j] = i < j ? v[e] : (i == j ? rd : 0.0);
    }
  } else {
    // at is the factorization's tile until the rows are loaded.
    LoadDiagBlock<kLdT>(F, f, c0, nb, &at[0][0]);
    __syncthreads();
    const bool zero = FactorInvertBlock<kLdT>(&at[0][0], dsh, colscr, nb);
    if (tid == 0) zero_pivot = zero ? 1 : 0;
    // W from the tile: M(j, i) above the diagonal, 1 / D(j) on it.
#pragma unroll 4
    for (int e = 0; e < kNB * kNB / kThreads; ++e) {
      const int idx = tid + kThreads * e;
      const int i = idx % kNB;
      const int j = idx / kNB;
      const double m = at[j][i];
      wt[i][j] = i < j ? m * at[j][j] : (i == j ? m : 0.0);
    }
    if (owner) {
      StoreFactored<kLdT, kThreads>(&at[0][0], dsh, StageSlot(W, s, kb), DiagD(A, W, s, kb), DiagInv(A, W, s, kb),
                                    tid);
    }
    __syncthreads();  // at is free again
    if (rows) {
      load_rows();
      store_rows();
    }
  }
  __syncthreads();
  if (owner && tid == 0 && zero_pivot != 0) {
    W.state->not_spd = 1;
    W.state->halt = 1;
  }
  if (!rows) return;
  // Each wavefront owns a 32 x 32 quadrant: four 16 x 16 FP64 MFMA blocks
  // (operand layout as in the trailing update below: lane l's result e is
  // row (l & 15), column 4 e + (l >> 4) of its block).
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wr = 32 * (wave & 1);
  const int wc = 32 * (wave >> 1);
  v4d acc[2][2];
#pragma unroll
  for (int cb = 0; cb < 2; ++cb)
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) acc[cb][rb] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k = 0; k < kNB; k += 4) {
    const int kr = k + (lane >> 4);
    double av[2], bv[2];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) av[rb] = at[kr][wr + 16 * rb + (lane & 15)];
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) bv[cb] = wt[kr][wc + 16 * cb + (lane & 15)];
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
#pragma unroll
      for (int rb = 0; rb < 2; ++rb) {
        acc[cb][rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[cb], av[rb], acc[cb][rb], 0, 0, 0);
      }
  }
#pragma unroll
  for (int cb = 0; cb < 2; ++cb) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int c = wc + 16 * cb + 4 * e + (lane >> 4);
      if (c >= nb) continue;
#pragma unroll
      for (int rb = 0; rb < 2; ++rb) {
        const int r = r0 + wr + 16 * rb + (lane & 15);
        if (r < f) F[static_cast<int64_t>(c0 + c) * f + r] = acc[cb][rb][e];
      }
    }
  }
}

// acc[t] = A_t (B D)^T, k = 0 .. K - 1, on FP64 MFMA for kTiles 64x64 tiles
// that share their B rows: A_t holds panel rows rbase + r0[t] .. + 63 and B
// rows rbase + cb0 .. + 63 of panel columns kbase .. kbase + K - 1 (rows
// from M, resp. N, on and k from K on are zero); D(k) = dk[k * dstride].
// Each wavefront owns a 32x32 quadrant of every tile, four 16x16 blocks. The
// MFMA's A operand is the B chunk (its 16 rows are tile columns) and its B
// operand the A chunk (its 16 columns are tile rows), so that lane l's
// result e is tile row (l & 15), column 4 e + (l >> 4) of its block: a
// wavefront's C accesses are 16-row runs of 4 columns (coalesced in the
// column-major panel / update matrix). K chunks of kTk go through a ring of
// kDepth register sets, so the loads of up to kDepth chunks are in flight
// together (a chunk's products take far less time than a load's round
// trip), and through two LDS buffers (2 (kTiles + 1) kTk (kTile + 2)
// doubles; rows padded by two doubles, so the two k rows an MFMA operand
// read touches fall in different banks). The chunks' order, and so every
// sum's, does not depend on kDepth. With kRelease the last barrier comes
// after the last product, so the LDS is free on return.
template <int kTiles, bool kRelease, int kDepth, bool kShortK>
__device__ __forceinline__ void TileProducts(const double* F, int f, int kbase, int K, int rbase, int M, int N,
                                             const int (&r0)[kTiles], int cb0, const double* dk,
                                             int64_t dstride, double* lds, v4d (&acc)[kTiles][2][2]) {
  auto As = reinterpret_cast<double (*)[kTiles][kTk][kTile + 2]>(lds);
  auto Bs = reinterpret_cast<double (*)[kTk][kTile + 2]>(lds + 2 * kTiles * kTk * (kTile + 2));
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wr = 32 * (wave & 1);
  const int wc = 32 * (wave >> 1);
#pragma unroll
  for (int t = 0; t < kTiles; ++t)
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
#pragma unroll
      for (int rb = 0; rb < 2; ++rb) acc[t][cb][rb] = v4d{0.0, 0.0, 0.0, 0.0};
  // Staging of one K chunk: each thread loads 4 entries of every operand
  // (rows rr = idx & 63 of k row kk = idx >> 6) and one D_k (k = tid %
  // kTk). Every load is issued at an address clamped into the operands, and
  // the padding is selected only when the chunk is stored to the LDS,
  // through Opaque: a load feeding a select directly would become a branch
  // that waits for its data before the next load is issued. D_k goes into
  // the first padding column of B's k row and scales B when it is read as
  // an MFMA operand (the same product b D_k as a scaled copy would hold).
  // Thread (wave w, lane l) takes row l of k rows w, w + 4, w + 8, w + 12
  // of a chunk (entry idx = tid + kThreads e: rr = l, k = idx >> 6): a
  // wave-uniform column and a per-lane row offset (fixed over the chunks).
  double ra[kDepth][kTiles][4], rb[kDepth][4], rd[kDepth];
  const int wq = WaveIndex();
  uint32_t aoff[kTiles];
#pragma unroll
  for (int t = 0; t < kTiles; ++t) aoff[t] = 8u * static_cast<uint32_t>(rbase + min(r0[t] + lane, M - 1));
  const uint32_t boff = 8u * static_cast<uint32_t>(rbase + min(cb0 + lane, N - 1));
  const int dkk = tid & (kTk - 1);
  auto load_chunk = [&](int q, int k0) {
    rd[q] = dk[min(k0 + dkk, K - 1) * dstride];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int kq = min(k0 + wq + 4 * e, K - 1);
      const double* col = F + static_cast<int64_t>(kbase + kq) * f;
#pragma unroll
      for (int t = 0; t < kTiles; ++t) ra[q][t][e] = LoadAt(col, aoff[t]);
      rb[q][e] = LoadAt(col, boff);
    }
  };
  auto store_chunk = [&](int q, int buf, int k0) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int idx = tid + kThreads * e;
      const int rr = idx & 63;
      const bool kin = k0 + (idx >> 6) < K;
#pragma unroll
      for (int t = 0; t < kTiles; ++t) {
        const double a = Opaque(ra[q][t][e]);
        As[buf][t][idx >> 6][rr] = (kin && r0[t] + rr < M) ? a : 0.0;
      }
      const double b = Opaque(rb[q][e]);
      Bs[buf][idx >> 6][rr] = (kin && cb0 + rr < N) ? b : 0.0;
    }
    if (tid < kTk) {
      const double d = Opaque(rd[q]);
      Bs[buf][tid][kTile] = k0 + tid < K ? d : 0.0;
    }
  };
  auto products = [&](int buf) {
#pragma unroll
    for (int kk = 0; kk < kTk; kk += 4) {
      const int kr = kk + (lane >> 4);
      double av[kTiles][2], bv[2];
#pragma unroll
      for (int t = 0; t < kTiles; ++t)
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) av[t][rb] = As[buf][t][kr][wr + 16 * rb + (lane & 15)];
      const double dkr = Bs[buf][kr][kTile];
#pragma unroll
      for (int cb = 0; cb < 2; ++cb) bv[cb] = Bs[buf][kr][wc + 16 * cb + (lane & 15)] * dkr;
#pragma unroll
      for (int t = 0; t < kTiles; ++t)
#pragma unroll
        for (int cb = 0; cb < 2; ++cb)
#pragma unroll
          for (int rb = 0; rb < 2; ++rb) {
            acc[t][cb][rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[cb], av[t][rb], acc[t][cb][rb], 0, 0, 0);
          }
    }
  };
  const int chunks = (K + kTk - 1) / kTk;
#pragma unroll
  for (int q = 0; q < kDepth; ++q) load_chunk(q, q * kTk);
  store_chunk(0, 0, 0);
  __syncthreads();
  auto step = [&](int q, int c) {
    if (c < chunks) {
      products(c & 1);
      if (c + 1 < chunks) {
        store_chunk((q + 1) % kDepth, (c + 1) & 1, (c + 1) * kTk);
        __syncthreads();
      }
    }
  };
  int g = 0;
  if (!kShortK) {
    for (; g + 2 * kDepth <= chunks; g += kDepth) {
#pragma unroll
      for (int q = 0; q < kDepth; ++q) {
        const int c = g + q;
        load_chunk(q, (c + kDepth) * kTk);  // set q's chunk c is in the LDS
        products(c & 1);
        store_chunk((q + 1) % kDepth, (c + 1) & 1, (c + 1) * kTk);
        __syncthreads();
      }
    }
#pragma unroll
    for (int q = 0; q < kDepth; ++q) {
      const int c = g + q;
      if (c + kDepth < chunks) load_chunk(q, (c + kDepth) * kTk);
      step(q, c);
    }
    g += kDepth;
  }
#pragma unroll
  for (int q = 0; q < kDepth; ++q) step(q, g + q);
  if (kRelease) __syncthreads();
}

// The factored diagonal block of step kb (staging slot: L below the
// diagonal, D on it) goes into the panel, by the workgroup of tile (0, 0).
// Split in two: each thread's loads (thread t takes entries t, t + kThreads,
// ...), then its stores, which can come after other work of the workgroup.
constexpr int kMovePer = kNB * kNB / kThreads;
__device__ __forceinline__ void LoadStagedBlock(const double* stage, double (&v)[kMovePer]) {
#pragma unroll
  for (int e = 0; e < kMovePer; ++e) v[e] = stage[threadIdx.x + kThreads * e];
}
__device__ __forceinline__ void StoreStagedBlock(double* F, int f, int c0, int nb, const double (&v)[kMovePer]) {
#pragma unroll
  for (int e = 0; e < kMovePer; ++e) {
    const int idx = threadIdx.x + kThreads * e;
    const int i = idx & (kNB - 1);
    const int j = idx / kNB;
    if (i >= j && i < nb) F[static_cast<int64_t>(c0 + j) * f + c0 + i] = Opaque(v[e]);
  }
}
__device__ __forceinline__ void MoveStagedBlock(double* F, int f, int c0, int nb, const double* stage) {
  double v[kMovePer];
  LoadStagedBlock(stage, v);
  StoreStagedBlock(F, f, c0, nb, v);
}

// C -= acc over the tile's lower entries (c <= r, inside M x N): the 16
// entries of a lane are loaded together first (clamped addresses), then
// the lower ones stored.
__device__ __forceinline__ void SubtractLower(double* C, int64_t ldc, int R0, int C0, int M, int N,
                                              const v4d (&acc)[2][2]) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wr = 32 * (wave & 1);
  const int wc = 32 * (wave >> 1);
  double prev[2][4][2];
#pragma unroll
  for (int cb = 0; cb < 2; ++cb) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int c = min(C0 + wc + 16 * cb + 4 * e + (lane >> 4), N - 1);
#pragma unroll
      for (int rb = 0; rb < 2; ++rb) {
        const int r = min(R0 + wr + 16 * rb + (lane & 15), M - 1);
        prev[cb][e][rb] = C[static_cast<int64_t>(c) * ldc + r];
      }
    }
  }
#pragma unroll
  for (int cb = 0; cb < 2; ++cb) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int c = C0 + wc + 16 * cb + 4 * e + (lane >> 4);
#pragma unroll
      for (int rb = 0; rb < 2; ++rb) {
        const int r = R0 + wr + 16 * rb + (lane & 15);
        if (c < N && r < M && c <= r) C[static_cast<int64_t>(c) * ldc + r] = prev[cb][e][rb] - acc[cb][rb][e];
      }
    }
  }
}

// C -= A_blk (L_blk D)^T over the lower triangle, 64x64 tiles.
// mode 0: trailing panel update after block column kb (K = nb), whose
// factored diagonal block tile (0, 0)'s workgroup first moves into the
// panel;
// mode 1: update matrix U = -L21 D L21^T (K = n_s), written, not accumulated;
// D comes from the copy of every block's D (the front's last factored
// diagonal block may still be in its staging slot), and in single-slot
// launches tile (0, 0) then moves that block into the panel (no other tile
// reads or writes it; where a launch of mode 0 moved it already, the same
// values again).
// Tile (ti, tj) (tj <= ti) of front s in slot `slot`.
template <int kMode, class Slots>
__device__ __forceinline__ void GemmLowerBody(const Slots& slots, int slot, int s, int ti, int tj, int kb) {
  constexpr int mode = kMode;
  BIND_SLOT(slot);
  if (Halted(W.state)) return;
  const int ns = A.sn_first[s + 1] - A.sn_first[s];
  const int f = A.sn_f[s];
  double* Fw = W.panel + A.panel_off[s];
  int kbase, K, rbase, M, N;
  double* C;
  int64_t ldc;
  const double* dk;
  int64_t dstride;
  if constexpr (mode == 0) {
    const int c0 = kb * kNB;
    if (c0 >= ns) return;
    const int nb = min(kNB, ns - c0);
    const double* stage = StageSlot(W, s, kb);
    if (ti == 0 && tj == 0) MoveStagedBlock(Fw, f, c0, nb, stage);
    kbase = c0;
    K = nb;
    rbase = c0 + nb;  // panel row of tile row 0 and panel column of tile col 0
    M = f - rbase;
    N = ns - rbase;
    C = Fw + static_cast<int64_t>(rbase) * f + rbase;
    ldc = f;
    dk = stage;
    dstride = kNB + 1;
  } else {
    const int ms = A.sn_m[s];
    kbase = 0;
    K = ns;
    rbase = ns;
    M = ms;
    N = ms;
    C = W.upd + A.upd_off[s];
    ldc = ms;
    dk = DiagD(A, W, s, 0);
    dstride = 1;
  }
  const int R0 = ti * kTile;
  const int C0 = tj * kTile;
  if (R0 >= M || C0 >= N) return;
  __shared__ double lds[2 * 2 * kTk * (kTile + 2)];
  v4d acc[1][2][2];
  const int r0[1] = {R0};
  TileProducts<1, false, 2, false>(Fw, f, kbase, K, rbase, M, N, r0, C0, dk, dstride, lds, acc);
  if constexpr (mode == 0) {
    SubtractLower(C, ldc, R0, C0, M, N, acc[0]);
  } else {
    // U = -L21 D L21^T plus, child by child in ascending order, the entries
    // of the children's update matrices that land in it (their columns past
    // rel_split: the extend-add of the update part, gathered through each
    // child's map from the parent's update rows to its own, -1 where it
    // has none).
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wr = 32 * (wave & 1);
    const int wc = 32 * (wave >> 1);
    double v[2][4][2];
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
#pragma unroll
      for (int e = 0; e < 4; ++e)
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) v[cb][e][rb] = 0.0 - acc[0][cb][rb][e];
    for (int64_t ci = A.children_ptr[s]; ci < A.children_ptr[s + 1]; ++ci) {
      const int ch = A.children[ci];
      const int64_t uoff = A.uinv_off[ch];
      if (uoff < 0) continue;
      const int32_t* inv = A.uinv + uoff;
      const int mc = A.sn_m[ch];
      const double* uc = W.upd + A.upd_off[ch];
      int crow[2], ccol[2][4];
#pragma unroll
      for (int rb = 0; rb < 2; ++rb) crow[rb] = inv[min(R0 + wr + 16 * rb + (lane & 15), M - 1)];
#pragma unroll
      for (int cb = 0; cb < 2; ++cb)
#pragma unroll
        for (int e = 0; e < 4; ++e) ccol[cb][e] = inv[min(C0 + wc + 16 * cb + 4 * e + (lane >> 4), N - 1)];
      double u[2][4][2];
#pragma unroll
      for (int cb = 0; cb < 2; ++cb)
#pragma unroll
        for (int e = 0; e < 4; ++e)
#pragma unroll
          for (int rb = 0; rb < 2; ++rb) {
            u[cb][e][rb] = uc[static_cast<int64_t>(max(ccol[cb][e], 0)) * mc + max(crow[rb], 0)];
          }
#pragma unroll
      for (int cb = 0; cb < 2; ++cb)
#pragma unroll
        for (int e = 0; e < 4; ++e)
#pragma unroll
          for (int rb = 0; rb < 2; ++rb) {
            const double x = Opaque(u[cb][e][rb]);
            if (crow[rb] >= 0 && ccol[cb][e] >= 0) v[cb][e][rb] = v[cb][e][rb] + x;
          }
    }
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int c = C0 + wc + 16 * cb + 4 * e + (lane >> 4);
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) {
          const int r = R0 + wr + 16 * rb + (lane & 15);
          if (c < N && r < M && c <= r) C[static_cast<int64_t>(c) * ldc + r] = v[cb][e][rb];
        }
      }
    }
    // Tile (0, 0) then moves the front's last factored diagonal block from
    // its staging slot into the panel (single-slot launches; the slot-table
    // form keeps the launch of mode 0 for it, within its register budget).
    if (std::is_same<Slots, OneSlot>::value && ti == 0 && tj == 0) {
      const int c_last = ((ns - 1) / kNB) * kNB;
      MoveStagedBlock(Fw, f, c_last, ns - c_last, StageSlot(W, s, c_last / kNB));
    }
  }
}

// LDS of the fused step: the K chunks of two tiles, then the diagonal tile
// (column-major, kNB + 1 apart), D, the factorization's scratch and the
// TRSM rows.
constexpr int kAheadLd = kNB + 1;
constexpr int kAheadLds = kNB * kAheadLd + 2 * kNB + kNB * kLdT;
static_assert(kAheadLds >= 2 * 3 * kTk * (kTile + 2), "fused step staging");
static_assert(4 * kTk >= kNB, "a block column's products load every chunk up front");

// Trailing update of step kb fused with step kb + 1's panel step. Tile
// (ti, tj), tj <= ti, of front s. The tiles of column 0 are the next block
// column. Each of their workgroups also updates the next diagonal block
// (tile (0, 0)), factors and inverts it (FactorInvertBlock) -- every copy
// computes the same values from the same inputs, bit for bit -- and multiplies its
// own updated rows by W = L^-T D^-1 (in tile (0, 0) the rows below the
// block's nb1 columns), so step kb + 1 has no launch of its own. Nothing
// here writes the diagonal tile into the panel (the copies read its values
// before the update); the workgroup of tile (0, 0) stages the factored
// block (and L^-1 for the solves) and sets the zero-pivot flag, and the
// next launch moves the block into the panel. The other tiles are plain
// trailing updates. A tile past the block column's rows or columns exits,
// so from the products on N > 0: block kb is full and block kb + 1 starts
// at panel column rbase.
template <class Slots>
__device__ __forceinline__ void AheadBody(const Slots& slots, int slot, int s, int ti, int tj, int kb) {
  BIND_SLOT(slot);
  if (Halted(W.state)) return;
  const int ns = A.sn_first[s + 1] - A.sn_first[s];
  const int f = A.sn_f[s];
  double* Fw = W.panel + A.panel_off[s];
  const int c0 = kb * kNB;
  if (c0 >= ns) return;
  const int nb = min(kNB, ns - c0);
  const double* stage = StageSlot(W, s, kb);
  // Tile (0, 0)'s workgroup moves the block; its stores come last (the
  // block is disjoint from everything the step reads or writes).
  const bool mover = ti == 0 && tj == 0;
  double mv[kMovePer];
  if (mover) LoadStagedBlock(stage, mv);
  const int rbase = c0 + nb;
  const int M = f - rbase;
  const int N = ns - rbase;
  const int R0 = ti * kTile;
  const int C0 = tj * kTile;
  if (R0 >= M || C0 >= N) {
    if (mover) StoreStagedBlock(Fw, f, c0, nb, mv);
    return;
  }
  double* C = Fw + static_cast<int64_t>(rbase) * f + rbase;
  __shared__ double lds[kAheadLds];
  if (tj > 0) {
    v4d acc[1][2][2];
    const int r0[1] = {R0};
    TileProducts<1, false, 4, true>(Fw, f, c0, nb, rbase, M, N, r0, C0, stage, kNB + 1, lds, acc);
    SubtractLower(C, f, R0, C0, M, N, acc[0]);
    return;
  }
  const int nb1 = min(kNB, N);
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wr = 32 * (wave & 1);
  const int wc = 32 * (wave >> 1);
  double* lt = lds;                         // the diagonal tile, then its factor
  double* dvec = lds + kNB * kAheadLd;      // D
  double* colscr = dvec + kNB;              // the factorization's scratch
  auto at = reinterpret_cast<double (*)[kLdT]>(colscr + kNB);  // at[k][r]: TRSM rows
  // The updated tile (0, 0) goes to lt (rows below nb1: to at, in tile
  // (0, 0)'s own workgroup), tile (ti, 0) to at; everything else in at is
  // zero (rows past M, columns past nb1). The tiles' current values are
  // loaded before the products, so that their loads are in flight with the
  // products' operands.
  auto load_prev = [&](double (&prev)[2][4][2], int t_r0) {
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int c = min(wc + 16 * cb + 4 * e + (lane >> 4), N - 1);
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) {
          const int r = min(t_r0 + wr + 16 * rb + (lane & 15), M - 1);
          prev[cb][e][rb] = C[static_cast<int64_t>(c) * f + r];
        }
      }
    }
  };
  auto put = [&](const v4d (&acc)[2][2], const double (&prev)[2][4][2], int t_r0, bool diag, bool rows) {
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int c = wc + 16 * cb + 4 * e + (lane >> 4);
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) {
          const int r = wr + 16 * rb + (lane & 15);
          const double v = prev[cb][e][rb] - acc[cb][rb][e];
          if (diag && r < nb1) {
            if (c <= r && c < N && r < M) lt[c * kAheadLd + r] = v;
            if (rows) at[c][r] = 0.0;
          } else if (rows) {
            at[c][r] = (c < nb1 && t_r0 + r < M) ? v : 0.0;
          }
        }
      }
    }
  };
  if (ti == 0) {
    v4d acc[1][2][2];
    const int r0[1] = {0};
    double prev[2][4][2];
    load_prev(prev, 0);
    TileProducts<1, true, 4, true>(Fw, f, c0, nb, rbase, M, N, r0, 0, stage, kNB + 1, lds, acc);
    put(acc[0], prev, 0, true, true);
  } else {
    v4d acc[2][2][2];
    const int r0[2] = {0, R0};
    double prev[2][2][4][2];
    load_prev(prev[0], 0);
    load_prev(prev[1], R0);
    TileProducts<2, true, 4, true>(Fw, f, c0, nb, rbase, M, N, r0, 0, stage, kNB + 1, lds, acc);
    put(acc[0], prev[0], 0, true, false);
    put(acc[1], prev[1], R0, false, true);
  }
  __syncthreads();
  const bool zero = FactorInvertBlock<kAheadLd>(lt, dvec, colscr, nb1);
  if (ti == 0 && tid == 0 && zero) {
    W.state->not_spd = 1;
    W.state->halt = 1;
  }
  if (ti == 0) {
    StoreFactored<kAheadLd, kThreads>(lt, dvec, StageSlot(W, s, kb + 1), DiagD(A, W, s, kb + 1),
                                      DiagInv(A, W, s, kb + 1), tid);
  }
  // L21 = A21 W, W(k, j) = M(j, k) / D(j) above the diagonal and 1 / D(j)
  // on it (lt holds M(j, k) at j * kAheadLd + k and 1 / D(j) at
  // j * kAheadLd + j), the same 64 x 64 x 64 FP64 MFMA product as the
  // panel step's.
  v4d acc[2][2];
#pragma unroll
  for (int cb = 0; cb < 2; ++cb)
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) acc[cb][rb] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k = 0; k < kNB; k += 4) {
    const int kr = k + (lane >> 4);
    double av[2], bv[2];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) av[rb] = at[kr][wr + 16 * rb + (lane & 15)];
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
      const int j = wc + 16 * cb + (lane & 15);
      const double m = lt[j * kAheadLd + kr];
      const double id = lt[j * kAheadLd + j];
      bv[cb] = kr < j ? m * id : (kr == j ? m : 0.0);
    }
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
#pragma unroll
      for (int rb = 0; rb < 2; ++rb) {
        acc[cb][rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[cb], av[rb], acc[cb][rb], 0, 0, 0);
      }
  }
#pragma unroll
  for (int cb = 0; cb < 2; ++cb) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int c = wc + 16 * cb + 4 * e + (lane >> 4);
      if (c >= nb1) continue;
#pragma unroll
      for (int rb = 0; rb < 2; ++rb) {
        const int r = wr + 16 * rb + (lane & 15);
        if (R0 + r < M && (ti > 0 || r >= nb1)) {
          Fw[static_cast<int64_t>(rbase + c) * f + rbase + R0 + r] = acc[cb][rb][e];
        }
      }
    }
  }
  if (mover) StoreStagedBlock(Fw, f, c0, nb, mv);
}

// Grid (tile rows, tile columns, fronts x slots) over the level's fronts.
template <class Slots, class Lists>
__global__ void __launch_bounds__(kThreads, 4) KGemmLower(Slots slots, Lists plan, int kb, int fronts) {
  if (blockIdx.y > blockIdx.x) return;
  const int slot = static_cast<int>(blockIdx.z) / fronts;
  const int s = PlanItem(plan, slot, static_cast<int>(blockIdx.z) % fronts);
  if (s < 0) return;
  GemmLowerBody<0>(slots, slot, s, blockIdx.x, blockIdx.y, kb);
}

// Trailing update of step kb with step kb + 1's diagonal block and panel
// rows (same grid).
template <class Slots, class Lists>
__global__ void __launch_bounds__(kThreads, 1) KGemmAhead(Slots slots, Lists plan, int kb, int fronts) {
  if (blockIdx.y > blockIdx.x) return;
  const int slot = static_cast<int>(blockIdx.z) / fronts;
  const int s = PlanItem(plan, slot, static_cast<int>(blockIdx.z) % fronts);
  if (s < 0) return;
  AheadBody(slots, slot, s, blockIdx.x, blockIdx.y, kb);
}

// Update matrices of the level's fronts over a flat list of their lower
// tiles (grid: tiles, slots): item t of a slot's list is the pair (front,
// ti << 16 | tj), so every workgroup has a tile to compute.
template <class Slots, class Lists>
__global__ void __launch_bounds__(kThreads, 4) KGemmUpdate(Slots slots, Lists tiles) {
  const int slot = blockIdx.y;
  const ListRef lr = tiles.get(slot);
  const int t = blockIdx.x;
  if (t >= lr.count) return;
  const int s = lr.items[2 * t];
  const int packed = lr.items[2 * t + 1];
  GemmLowerBody<1>(slots, slot, s, packed >> 16, packed & 0xffff, 0);
}

// ---------------------------------------------------------------- solves

// acc - sum_k F(k, i) v[k] over one block column k in [k0, k1): column k
// goes to partial sum (k - k0) % 4, in increasing k, whatever the batch of
// loads issued together (kBatch, a multiple of 4); KForwardBlocks forms the
// same partial sums with four threads per row.
template <int kBatch>
__device__ __forceinline__ double RowUpdate(const double* F, int f, int i, const double* v, int k0, int k1,
                                            double acc0) {
  double acc[4] = {acc0, 0.0, 0.0, 0.0};
  int k = k0;
  for (; k + kBatch <= k1; k += kBatch) {
    double fv[kBatch];
#pragma unroll
    for (int t = 0; t < kBatch; ++t) fv[t] = F[static_cast<int64_t>(k + t) * f + i];
#pragma unroll
    for (int t = 0; t < kBatch; ++t) acc[t & 3] = fma(-fv[t], v[k + t], acc[t & 3]);
  }
  for (; k + 4 <= k1; k += 4) {
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = fma(-F[static_cast<int64_t>(k + t) * f + i], v[k + t], acc[t]);
  }
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    if (k + t < k1) acc[t] = fma(-F[static_cast<int64_t>(k + t) * f + i], v[k + t], acc[t]);
  }
  return (acc[0] + acc[1]) + (acc[2] + acc[3]);
}

// Forward: gather the children's update vectors, then L y = b over the front
// (unit lower), leaving y in the own rows and the update for the parent in R.
// With `tri` the rows below the front's own columns (i >= n) are left as
// assembled and KForwardRest applies L21 to them across many workgroups.
template <class Slots, class Lists>
__global__ void KForward(Slots slots, Lists plan, int tri) {
  BIND_SLOT(blockIdx.y);
  if (Halted(W.state)) return;
  const int s = PlanItem(plan, blockIdx.y, blockIdx.x);
  if (s < 0) return;
  const int first = A.sn_first[s];
  const int n = A.sn_first[s + 1] - first;
  const int f = A.sn_f[s];
  const double* F = W.panel + A.panel_off[s];
  extern __shared__ double v[];
  const int tid = threadIdx.x;
  for (int i = tid; i < f; i += blockDim.x) {
    v[i] = i < n ? W.rhs[A.iperm[first + i]] : 0.0;
  }
  __syncthreads();
  for (int64_t ci = A.children_ptr[s]; ci < A.children_ptr[s + 1]; ++ci) {
    const int c = A.children[ci];
    const int mc = A.sn_m[c];
    const int nc = A.sn_first[c + 1] - A.sn_first[c];
    const int32_t* rel = A.rel + A.rel_off[c];
    const double* u = W.vec + A.sn_idx_ptr[c] + nc;
    for (int a = tid; a < mc; a += blockDim.x) v[rel[a]] += u[a];
    __syncthreads();
  }
  for (int b0 = 0; b0 < n; b0 += 64) {
    const int b1 = min(b0 + 64, n);
    if (tid < 64) {
      // Diagonal block: y = M v with M = L_kk^-1 (identity padded, so the
      // entries v[j] past the block meet zeros); lane tid forms row tid,
      // four partial sums. Every lane reads v before any writes it.
      const double* Mb = DiagInv(A, W, s, b0 / kNB);
      double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
      for (int h = 0; h < kNB; h += 32) {  // 32 loads in flight (the register budget of 1024 threads)
#pragma unroll
        for (int t = 0; t < 32; ++t) {
          acc[t & 3] = fma(Mb[(h + t) * kNB + tid], v[min(b0 + h + t, f - 1)], acc[t & 3]);
        }
      }
      const double y = (acc[0] + acc[1]) + (acc[2] + acc[3]);
      if (b0 + tid < b1) v[b0 + tid] = y;
    }
    __syncthreads();
    // Rows below the block. The load batch does not change the sums; the
    // slot-table form (which keeps more of its slot in registers) takes
    // half of it to stay within the 1024-thread register budget.
    constexpr int kBatch = std::is_same<Slots, OneSlot>::value ? 16 : 8;
    const int rows_end = tri != 0 ? n : f;
    for (int i = b1 + tid; i < rows_end; i += blockDim.x) v[i] = RowUpdate<kBatch>(F, f, i, v, b0, b1, v[i]);
    __syncthreads();
  }
  double* out = W.vec + A.sn_idx_ptr[s];
  for (int i = tid; i < f; i += blockDim.x) out[i] = v[i];
}

// v2 -= L21 v1 for the rows below the own columns (KForwardBlocks leaves
// them as assembled): one thread per row, block column by block column
// (RowUpdate).
template <class Slots, class Lists>
__global__ void __launch_bounds__(64) KForwardRest(Slots slots,
                                                   Lists plan) {
  BIND_SLOT(blockIdx.z);
  if (Halted(W.state)) return;
  const int s = PlanItem(plan, blockIdx.z, blockIdx.y);
  if (s < 0) return;
  const int n = A.sn_first[s + 1] - A.sn_first[s];
  const int f = A.sn_f[s];
  const int i = n + blockIdx.x * 64 + threadIdx.x;
  if (i >= f) return;
  const double* F = W.panel + A.panel_off[s];
  double* v = W.vec + A.sn_idx_ptr[s];
  double acc = v[i];
  for (int k0 = 0; k0 < n; k0 += kNB) acc = RowUpdate<32>(F, f, i, v, k0, min(n, k0 + kNB), acc);
  v[i] = acc;
}

// L21^T x2 per column of the fronts at the level, one wavefront per column
// (lane-strided sum, wave reduction), into tmp at the column's elimination
// index.
template <class Slots, class Lists>
__global__ void KBackwardRest(Slots slots, Lists plan) {
  BIND_SLOT(blockIdx.z);
  if (Halted(W.state)) return;
  const int s = PlanItem(plan, blockIdx.z, blockIdx.y);
  if (s < 0) return;
  const int first = A.sn_first[s];
  const int n = A.sn_first[s + 1] - first;
  const int j = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (j >= n) return;
  const int m = A.sn_m[s];
  const int f = n + m;
  const int lane = threadIdx.x & 63;
  const int32_t* idx = A.sn_idx + A.sn_idx_ptr[s];
  const double* col = W.panel + A.panel_off[s] + static_cast<int64_t>(j) * f + n;
  double acc = 0.0;
  int a = lane;
  for (; a + 192 < m; a += 256) {
    const double f0 = col[a], f1 = col[a + 64], f2 = col[a + 128], f3 = col[a + 192];
    acc = fma(f0, W.xnew[idx[n + a]], acc);
    acc = fma(f1, W.xnew[idx[n + a + 64]], acc);
    acc = fma(f2, W.xnew[idx[n + a + 128]], acc);
    acc = fma(f3, W.xnew[idx[n + a + 192]], acc);
  }
  for (; a < m; a += 64) acc = fma(col[a], W.xnew[idx[n + a]], acc);
  acc = WaveSum(acc);
  if (lane == 0) W.solve_tmp[first + j] = acc;
}

// Backward: x1 = L11^-T (D^-1 y1 - L21^T x2), x2 from the ancestors. With
// `tri`, L21^T x2 comes from KBackwardRest.
template <class Slots, class Lists>
__global__ void KBackward(Slots slots, Lists plan, int tri) {
  BIND_SLOT(blockIdx.y);
  if (Halted(W.state)) return;
  const int s = PlanItem(plan, blockIdx.y, blockIdx.x);
  if (s < 0) return;
  const int first = A.sn_first[s];
  const int n = A.sn_first[s + 1] - first;
  const int m = A.sn_m[s];
  const int f = n + m;
  const double* F = W.panel + A.panel_off[s];
  const double* y = W.vec + A.sn_idx_ptr[s];
  const int32_t* idx = A.sn_idx + A.sn_idx_ptr[s];
  extern __shared__ double t[];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int waves = blockDim.x >> 6;
  for (int j = tid; j < n; j += blockDim.x) {
    const double yd = y[j] / F[static_cast<int64_t>(j) * f + j];
    t[j] = tri != 0 ? yd - W.solve_tmp[first + j] : yd;
  }
  for (int a = tid; a < m; a += blockDim.x) t[n + a] = W.xnew[idx[n + a]];
  __syncthreads();
  for (int j = wave; tri == 0 && j < n; j += waves) {
    const double* col = F + static_cast<int64_t>(j) * f + n;
    double acc = 0.0;
    int a = lane;
    for (; a + 192 < m; a += 256) {
      const double f0 = col[a], f1 = col[a + 64], f2 = col[a + 128], f3 = col[a + 192];
      acc = fma(f0, t[n + a], acc);
      acc = fma(f1, t[n + a + 64], acc);
      acc = fma(f2, t[n + a + 128], acc);
      acc = fma(f3, t[n + a + 192], acc);
    }
    for (; a < m; a += 64) acc = fma(col[a], t[n + a], acc);
    acc = WaveSum(acc);
    if (lane == 0) t[j] -= acc;
  }
  __syncthreads();
  for (int kb = (n + kNB - 1) / kNB - 1; kb >= 0; --kb) {
    // The factorization's block columns, last first.
    const int b0 = kb * kNB;
    const int b1 = min(n, b0 + kNB);
    // Columns j0 and j0 + waves of this wave side by side (their loads in
    // flight together; two keep the 1024-thread kernel within its 128
    // registers), each summed in the same order as on its own; a column past
    // the block reads column b1 - 1 and is discarded.
    constexpr int kCols = 2;
    for (int j0 = b0 + wave; j0 < b1; j0 += kCols * waves) {
      const double* col[kCols];
#pragma unroll
      for (int q = 0; q < kCols; ++q) col[q] = F + static_cast<int64_t>(min(j0 + q * waves, b1 - 1)) * f;
      double acc[kCols];
#pragma unroll
      for (int q = 0; q < kCols; ++q) acc[q] = 0.0;
      int i = b1 + lane;
      for (; i + 192 < n; i += 256) {
        double fv[kCols][4];
#pragma unroll
        for (int q = 0; q < kCols; ++q) {
#pragma unroll
          for (int u = 0; u < 4; ++u) fv[q][u] = col[q][i + 64 * u];
        }
#pragma unroll
        for (int q = 0; q < kCols; ++q) {
#pragma unroll
          for (int u = 0; u < 4; ++u) acc[q] = fma(fv[q][u], t[i + 64 * u], acc[q]);
        }
      }
      for (; i < n; i += 64) {
#pragma unroll
        for (int q = 0; q < kCols; ++q) acc[q] = fma(col[q][i], t[i], acc[q]);
      }
#pragma unroll
      for (int q = 0; q < kCols; ++q) {
        const double sum = WaveSum(acc[q]);
        if (lane == 0 && j0 + q * waves < b1) t[j0 + q * waves] -= sum;
      }
    }
    __syncthreads();
    if (tid < 64) {
      // Within the block: x = M^T t with M = L_kk^-1 (identity padded, so
      // the entries t[i] past the block meet zeros); lane tid forms x_tid
      // from column tid of M, two doubles a load, four partial sums. Every
      // lane reads t before any writes it.
      const double2* Mc = reinterpret_cast<const double2*>(DiagInv(A, W, s, kb) + tid * kNB);
      double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int i = 0; i < kNB / 2; ++i) {
        const double2 m = Mc[i];
        acc[(2 * i) & 3] = fma(m.x, t[min(b0 + 2 * i, f - 1)], acc[(2 * i) & 3]);
        acc[(2 * i + 1) & 3] = fma(m.y, t[min(b0 + 2 * i + 1, f - 1)], acc[(2 * i + 1) & 3]);
      }
      const double x = (acc[0] + acc[1]) + (acc[2] + acc[3]);
      if (b0 + tid < b1) t[b0 + tid] = x;
    }
    __syncthreads();
  }
  for (int j = tid; j < n; j += blockDim.x) {
    W.xnew[first + j] = t[j];
    W.step[A.iperm[first + j]] = t[j];
  }
}

// ------------------------------------ multi-workgroup solves of wide fronts

// Synchronization words of supernode s's multi-workgroup solves: a ticket
// counter, a finished counter, an abort word, then one ready flag per
// 64-column block. All are zero between launches (the last workgroup of a
// front to finish clears them).
__device__ __forceinline__ int32_t* SolveSync(const DeviceAnalysis& A, const DeviceWork& W, int s) {
  return W.solve_sync + A.inv_off[s] / (kNB * kNB) + 3 * static_cast<int64_t>(s);
}

// Polls of a ready flag (with a short sleep between) before a wait gives up:
// about a second, against microseconds in a working solve.
constexpr int kSpinLimit = 1 << 20;

// Thread 0 waits for a block's ready flag (acquire, device scope); false
// when the front's solve aborted or the wait gave up (then it aborts it).
__device__ bool WaitReady(int32_t* flag, int32_t* abort_word) {
  for (int n = 0; n < kSpinLimit; ++n) {
    if (__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != 0) return true;
    if (__hip_atomic_load(abort_word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return false;
    __builtin_amdgcn_s_sleep(1);
  }
  __hip_atomic_store(abort_word, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return false;
}

// The workgroup's ticket (the order in which the front's workgroups
// started): a workgroup only waits for blocks of smaller tickets, whose
// workgroups are already running, so the waits always end.
__device__ __forceinline__ int TakeTicket(int32_t* sync, int* shared_ticket) {
  if (threadIdx.x == 0) *shared_ticket = atomicAdd(&sync[0], 1);
  __syncthreads();
  return *shared_ticket;
}

// Publishes this workgroup's block (its stores are done by every thread
// before the barrier), then the last of the front's nbk workgroups clears
// the words for the next launch and reports an aborted solve.
__device__ void PublishBlock(int32_t* sync, int kb, int nbk, LmState* state) {
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    __hip_atomic_store(&sync[3 + kb], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    if (atomicAdd(&sync[1], 1) == nbk - 1) {
      if (__hip_atomic_load(&sync[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
        state->not_spd = kSolveSyncFault;
        state->halt = 1;
      }
      for (int q = 0; q < nbk; ++q) sync[3 + q] = 0;
      sync[0] = 0;
      sync[1] = 0;
      sync[2] = 0;
    }
  }
}

// Forward solve L y = v over the own columns of the level's fronts, one
// workgroup per 64-column block (taken in ticket order): the block's rows
// start from the right-hand side and the children's update vectors, take
// the row update of every earlier block as soon as its y is ready (the
// same partial sums as RowUpdate: thread p of a row sums the columns
// k0 + p, k0 + p + 4, ...), then y_b = M_b v_b with the kept inverse (four
// partial sums over t = p, p + 4, ...). The rows below the own columns are
// left as assembled (ticket 0 writes them) for KForwardRest.
template <class Slots, class Lists>
__global__ void __launch_bounds__(kThreads) KForwardBlocks(Slots slots, Lists plan) {
  BIND_SLOT(blockIdx.z);
  if (Halted(W.state)) return;
  const int s = PlanItem(plan, blockIdx.z, blockIdx.y);
  if (s < 0) return;
  const int first = A.sn_first[s];
  const int n = A.sn_first[s + 1] - first;
  const int nbk = (n + kNB - 1) / kNB;
  if (static_cast<int>(blockIdx.x) >= nbk) return;
  const int f = A.sn_f[s];
  const double* F = W.panel + A.panel_off[s];
  double* out = W.vec + A.sn_idx_ptr[s];
  int32_t* sync = SolveSync(A, W, s);
  __shared__ int ticket;
  __shared__ double v[kNB];   // the block's rows (to vend)
  __shared__ double yq[kNB];  // y of the block being applied
  const int tid = threadIdx.x;
  const int kb = TakeTicket(sync, &ticket);
  if (kb >= nbk) {  // words left over from an aborted launch: abort this one too
    if (tid == 0) __hip_atomic_store(&sync[2], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  const int b0 = kb * kNB;
  const int b1 = min(n, b0 + kNB);
  const int vend = min(b0 + kNB, f);
  const int r = tid >> 2;  // row b0 + r
  const int p = tid & 3;   // columns k0 + p + 4 j
  // The diagonal block's inverse, this thread's entries (t = p + 4 j of row r).
  const double* Mb = DiagInv(A, W, s, kb);
  double mv[kNB / 4];
#pragma unroll
  for (int j = 0; j < kNB / 4; ++j) mv[j] = Mb[(p + 4 * j) * kNB + r];
  for (int i = tid; i < kNB; i += kThreads) v[i] = b0 + i < n ? W.rhs[A.iperm[first + b0 + i]] : 0.0;
  __syncthreads();
  for (int64_t ci = A.children_ptr[s]; ci < A.children_ptr[s + 1]; ++ci) {
    const int c = A.children[ci];
    const int mc = A.sn_m[c];
    const int nc = A.sn_first[c + 1] - A.sn_first[c];
    const int32_t* rel = A.rel + A.rel_off[c];
    const double* u = W.vec + A.sn_idx_ptr[c] + nc;
    for (int a = tid; a < mc; a += kThreads) {
      const int i = rel[a];
      if (i >= b0 && i < vend) v[i - b0] += u[a];
    }
    __syncthreads();
  }
  if (kb == 0) {
    // The rows below the own columns, as assembled (for KForwardRest).
    for (int i = n + tid; i < f; i += kThreads) out[i] = 0.0;
    __syncthreads();
    for (int64_t ci = A.children_ptr[s]; ci < A.children_ptr[s + 1]; ++ci) {
      const int c = A.children[ci];
      const int mc = A.sn_m[c];
      const int nc = A.sn_first[c + 1] - A.sn_first[c];
      const int32_t* rel = A.rel + A.rel_off[c];
      const double* u = W.vec + A.sn_idx_ptr[c] + nc;
      for (int a = tid; a < mc; a += kThreads) {
        const int i = rel[a];
        if (i >= n) out[i] += u[a];
      }
      __threadfence_block();
      __syncthreads();
    }
  }
  // Row updates of every earlier block, in block order.
  const bool own_row = b0 + r < b1;
  const int row = min(b0 + r, f - 1);
  double run = p == 0 ? v[r] : 0.0;  // the row's running value (thread p = 0)
  double fv[kNB / 4];
  auto load_block = [&](int q) {
#pragma unroll
    for (int j = 0; j < kNB / 4; ++j) fv[j] = F[static_cast<int64_t>(q * kNB + p + 4 * j) * f + row];
  };
  if (kb > 0) load_block(0);
  for (int q = 0; q < kb; ++q) {
    if (tid == 0) (void)WaitReady(&sync[3 + q], &sync[2]);
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (tid < kNB) yq[tid] = out[q * kNB + tid];
    __syncthreads();
    double acc = p == 0 ? run : 0.0;
#pragma unroll
    for (int j = 0; j < kNB / 4; ++j) acc = fma(-fv[j], yq[p + 4 * j], acc);
    if (q + 1 < kb) load_block(q + 1);  // in flight during the next wait
    const double a1 = __shfl_down(acc, 1, 4);
    const double a2 = __shfl_down(acc, 2, 4);
    const double a3 = __shfl_down(acc, 3, 4);
    if (p == 0 && own_row) run = (acc + a1) + (a2 + a3);
    __syncthreads();  // yq is rewritten next
  }
  if (p == 0 && own_row) v[r] = run;
  __syncthreads();
  // y_b = M_b v_b (identity padded; rows past vend read the last one).
  double acc = 0.0;
#pragma unroll
  for (int j = 0; j < kNB / 4; ++j) acc = fma(mv[j], v[min(p + 4 * j, vend - 1 - b0)], acc);
  const double a1 = __shfl_down(acc, 1, 4);
  const double a2 = __shfl_down(acc, 2, 4);
  const double a3 = __shfl_down(acc, 3, 4);
  if (p == 0 && own_row) out[b0 + r] = (acc + a1) + (a2 + a3);
  PublishBlock(sync, kb, nbk, W.state);
}

// Backward solve over the own columns of the level's fronts, one workgroup
// per 64-column block, the last block first (ticket order): t = D^-1 y
// (less L21^T x2 from KBackwardRest when the front has rows below), less
// L_qb^T x_q for every later block q as soon as its x is ready (thread p of
// a column sums rows 16 p .. 16 p + 15 of the block), then x_b = M_b^T t_b
// with the kept inverse.
template <class Slots, class Lists>
__global__ void __launch_bounds__(kThreads) KBackwardBlocks(Slots slots, Lists plan, int tri) {
  BIND_SLOT(blockIdx.z);
  if (Halted(W.state)) return;
  const int s = PlanItem(plan, blockIdx.z, blockIdx.y);
  if (s < 0) return;
  const int first = A.sn_first[s];
  const int n = A.sn_first[s + 1] - first;
  const int nbk = (n + kNB - 1) / kNB;
  if (static_cast<int>(blockIdx.x) >= nbk) return;
  const int f = A.sn_f[s];
  const double* F = W.panel + A.panel_off[s];
  const double* y = W.vec + A.sn_idx_ptr[s];
  const int32_t* idx = A.sn_idx + A.sn_idx_ptr[s];
  int32_t* sync = SolveSync(A, W, s);
  __shared__ int ticket;
  __shared__ double t[kNB];   // the block's t, then x
  __shared__ double xq[kNB];  // x of the block being applied
  const int tid = threadIdx.x;
  const int kb = nbk - 1 - TakeTicket(sync, &ticket);
  if (kb < 0) {  // words left over from an aborted launch: abort this one too
    if (tid == 0) __hip_atomic_store(&sync[2], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  const int b0 = kb * kNB;
  const int b1 = min(n, b0 + kNB);
  const int c = tid >> 2;  // column b0 + c
  const int p = tid & 3;   // rows 16 p .. 16 p + 15 of a block, or i = p + 4 j of the diagonal block
  const int col = min(b0 + c, n - 1);
  const double* Mc = DiagInv(A, W, s, kb) + c * kNB;
  double mv[kNB / 4];
#pragma unroll
  for (int j = 0; j < kNB / 4; ++j) mv[j] = Mc[p + 4 * j];
  for (int i = tid; i < kNB; i += kThreads) {
    const int j = b0 + i;
    if (j < b1) {
      const double yd = y[j] / F[static_cast<int64_t>(j) * f + j];
      t[i] = tri != 0 ? yd - W.solve_tmp[first + j] : yd;
    } else {
      t[i] = j < f ? W.xnew[idx[j]] : 0.0;  // past the own columns (meets the inverse's zero padding)
    }
  }
  __syncthreads();
  double run = p == 0 ? t[c] : 0.0;  // the column's running value (thread p = 0)
  double fv[16];
  auto load_block = [&](int q) {
#pragma unroll
    for (int j = 0; j < 16; ++j) fv[j] = F[static_cast<int64_t>(col) * f + min(q * kNB + 16 * p + j, f - 1)];
  };
  if (kb + 1 < nbk) load_block(nbk - 1);
  for (int q = nbk - 1; q > kb; --q) {
    if (tid == 0) (void)WaitReady(&sync[3 + q], &sync[2]);
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    const int q1 = min(n, q * kNB + kNB);
    if (tid < kNB) xq[tid] = q * kNB + tid < q1 ? W.xnew[first + q * kNB + tid] : 0.0;
    __syncthreads();
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < 16; ++j) acc = fma(fv[j], xq[16 * p + j], acc);
    if (q - 1 > kb) load_block(q - 1);  // in flight during the next wait
    const double a1 = __shfl_down(acc, 1, 4);
    const double a2 = __shfl_down(acc, 2, 4);
    const double a3 = __shfl_down(acc, 3, 4);
    if (p == 0) run -= (acc + a1) + (a2 + a3);
    __syncthreads();  // xq is rewritten next
  }
  if (p == 0 && b0 + c < b1) t[c] = run;
  __syncthreads();
  // x_b = M_b^T t_b: column c of M (identity padded).
  double acc = 0.0;
#pragma unroll
  for (int j = 0; j < kNB / 4; ++j) acc = fma(mv[j], t[p + 4 * j], acc);
  const double a1 = __shfl_down(acc, 1, 4);
  const double a2 = __shfl_down(acc, 2, 4);
  const double a3 = __shfl_down(acc, 3, 4);
  if (p == 0 && b0 + c < b1) {
    const double x = (acc + a1) + (a2 + a3);
    W.xnew[first + b0 + c] = x;
    W.step[A.iperm[first + b0 + c]] = x;
  }
  PublishBlock(sync, kb, nbk, W.state);
}

// ---------------------------------------------------------------- LM control

__device__ double BlockSum(double v, double* red) {
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
  for (int stride = blockDim.x / 2; stride > 0; stride >>= 1) {
    if (tid < stride) red[tid] += red[tid + stride];
    __syncthreads();
  }
  const double out = red[0];
  __syncthreads();
  return out;
}

// Problem::Cost terms (problem.cpp:132-156) into per-block partial sums.
template <class Slots>
__global__ void __launch_bounds__(kThreads) KCost(Slots slots, int ignore_halt) {
  BIND_SLOT(blockIdx.y);
  const double* map = W.map;
  const double* poses = W.poses;
  double* partials = W.partials;
  if (ignore_halt == 0 && Halted(W.state)) return;
  __shared__ double red[kThreads];
  const GridParams g = P.grid;
  const int64_t total = P.num_points + P.num_eikonal + P.num_odom;
  double acc = 0.0;
  for (int64_t idx = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; idx < total;
       idx += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    if (idx < P.num_points) {
      const int64_t s = idx;
      const int f = P.pt_frame[s];
      const double px = P.pt_x[s];
      const double py = P.pt_y[s];
      const double c = cos(poses[3 * f + 2]);
      const double sn = sin(poses[3 * f + 2]);
      const double gx = c * px + (-sn) * py + poses[3 * f];
      const double gy = sn * px + c * py + poses[3 * f + 1];
      const double fx = (gx - g.min_x) / g.dx;
      const double fy = (gy - g.min_y) / g.dy;
      double residual = -P.pt_delta[s];
      if (!(fx < 0.0 || fy < 0.0 || fx > g.nx - 1 || fy > g.ny - 1)) {
        CellW cell;
        cell.w = min(static_cast<int>(fx), g.nx - 2);
        cell.h = min(static_cast<int>(fy), g.ny - 2);
        cell.a = fx - cell.w;
        cell.b = fy - cell.h;
        FillWeights(cell);
        residual = InterpolateCell(g, map, cell) - P.pt_delta[s];
      }
      const double r = P.pt_sw[s] * residual;
      if (P.huber_delta > 0.0 && fabs(r) > P.huber_delta) {
        acc += P.huber_delta * (2.0 * fabs(r) - P.huber_delta);
      } else {
        acc += r * r;
      }
    } else if (idx < P.num_points + P.num_eikonal) {
      const int64_t e = idx - P.num_points;
      const int w = P.eik_w[e];
      const int h = P.eik_h[e];
      const double d00 = map[h * g.nx + w];
      const double d10 = map[h * g.nx + w + 1];
      const double d01 = map[(h + 1) * g.nx + w];
      const double grx = (d10 - d00) / g.dx;
      const double gry = (d01 - d00) / g.dy;
      const double r = P.eik_sw[e] * (1.0 - (P.eik_nx[e] * grx + P.eik_ny[e] * gry));
      acc += r * r;
    } else {
      const int64_t o = idx - P.num_points - P.num_eikonal;
      const double* pi = poses + 3 * P.od_i[o];
      const double* pj = poses + 3 * P.od_j[o];
      const double c = cos(pi[2]);
      const double s = sin(pi[2]);
      const double dtx = pj[0] - pi[0];
      const double dty = pj[1] - pi[1];
      const double sw = P.od_sw[o];
      const double a = sw * ((c * dtx + s * dty) - P.od_dx[o]);
      const double b = sw * (((-s) * dtx + c * dty) - P.od_dy[o]);
      const double cc = sw * WrapAngle(pj[2] - pi[2] - P.od_dth[o]);
      acc += a * a + b * b + cc * cc;
    }
  }
  const double total_sum = BlockSum(acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = total_sum;
}

template <class Slots>
__global__ void KLmInit(Slots slots, double lambda_init) {
  BIND_SLOT(blockIdx.y);
  __shared__ double red[kThreads];
  double acc = 0.0;
  for (int i = threadIdx.x; i < kPartialBlocks; i += blockDim.x) acc += W.partials[i];
  const double cost = BlockSum(acc, red);
  if (threadIdx.x == 0) {
    LmState* st = W.state;
    st->lambda = lambda_init;
    st->previous_cost = cost;
    st->nu = 2.0;
    st->initial_cost = cost;
    st->final_cost = cost;
    st->last_cost = cost;
    st->step_norm = 0.0;
    st->iter = 0;
    st->halt = 0;
    st->converged = 0;
    st->overflow = 0;
    st->not_spd = 0;
    st->revert = 0;
    W.cost_history[0] = cost;
  }
}

// Step norm and the predicted reduction step . (lambda step + rhs).
template <class Slots>
__global__ void KStepReduce(Slots slots) {
  BIND_SLOT(blockIdx.y);
  const int n = P.n;
  const double* step = W.step;
  const double* rhs = W.rhs;
  double* partials = W.partials;
  const LmState* st = W.state;
  if (Halted(st)) return;
  __shared__ double red[kThreads];
  const double lambda = st->lambda;
  double a = 0.0;
  double b = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double s = step[i];
    a += s * s;
    b += s * (lambda * s + rhs[i]);
  }
  const double sa = BlockSum(a, red);
  const double sb = BlockSum(b, red);
  if (threadIdx.x == 0) {
    partials[kPartialBlocks + blockIdx.x] = sa;
    partials[2 * kPartialBlocks + blockIdx.x] = sb;
  }
}

template <class Slots>
__global__ void KBackup(Slots slots) {
  BIND_SLOT(blockIdx.y);
  const int nmap = S.nmap;
  const int npose = 3 * P.num_poses;
  if (Halted(W.state)) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nmap + npose; i += gridDim.x * blockDim.x) {
    if (i < nmap) W.map_backup[i] = W.map[i];
    else W.poses_backup[i - nmap] = W.poses[i - nmap];
  }
}

// Problem::ApplyStep (problem.cpp:117-130)
template <class Slots>
__global__ void KApplyStep(Slots slots) {
  BIND_SLOT(blockIdx.y);
  if (Halted(W.state)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < P.num_active) {
    W.map[P.active_to_node[i]] += W.step[i];
  } else if (i < P.num_active + P.num_poses - 1) {
    const int pose = i - P.num_active + 1;
    const int col = P.num_active + 3 * (pose - 1);
    W.poses[3 * pose] += W.step[col];
    W.poses[3 * pose + 1] += W.step[col + 1];
    W.poses[3 * pose + 2] = WrapAngle(W.poses[3 * pose + 2] + W.step[col + 2]);
  }
}

// Convergence test and damping update of SolveLm (lm_solver.cpp:376-425).
template <class Slots>
__global__ void KControl(Slots slots, LmOptions o) {
  BIND_SLOT(blockIdx.y);
  if (Halted(W.state)) return;
  __shared__ double red[kThreads];
  double c = 0.0, a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < kPartialBlocks; i += blockDim.x) {
    c += W.partials[i];
    a += W.partials[kPartialBlocks + i];
    b += W.partials[2 * kPartialBlocks + i];
  }
  const double cost = BlockSum(c, red);
  const double norm_sq = BlockSum(a, red);
  const double predicted = BlockSum(b, red);
  if (threadIdx.x != 0) return;
  LmState* st = W.state;
  st->revert = 0;
  st->last_cost = cost;
  st->iter += 1;
  const int iter = st->iter;
  st->step_norm = sqrt(norm_sq);
  if (sqrt(norm_sq) < o.step_tolerance) {
    if (iter < o.history_capacity) W.cost_history[iter] = cost;
    st->final_cost = cost;
    st->converged = 1;
    st->halt = 1;
    return;
  }
  double current = cost;
  if (o.trust_region != 0) {
    const double rho = predicted > 0.0 ? (st->previous_cost - cost) / predicted : -1.0;
    if (rho > 0.0) {
      st->previous_cost = cost;
      const double shrink = 1.0 - pow(2.0 * rho - 1.0, 3.0);
      st->lambda = fmax(st->lambda * fmax(1.0 / 3.0, shrink), 1e-7);
      st->nu = 2.0;
    } else {
      st->revert = 1;
      current = st->previous_cost;
      st->lambda *= st->nu;
      st->nu *= 2.0;
    }
  } else if (cost < st->previous_cost) {
    st->lambda /= o.lambda_factor;
    st->previous_cost = cost;
  } else {
    st->lambda *= o.lambda_factor;
    if (o.reject_worse_steps != 0) {
      st->revert = 1;
      current = st->previous_cost;
    } else {
      st->previous_cost = cost;
    }
  }
  if (iter < o.history_capacity) W.cost_history[iter] = current;
  st->final_cost = current;
  if (iter >= o.max_iterations) st->halt = 1;
}

/// Copies the LM state of every slot into out[slot] (one thread per slot).
template <class Slots>
__global__ void KGatherStates(Slots slots, int count, LmState* out) {
  const int b = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (b >= count) return;
  out[b] = *slots.get(b).W.state;
}

template <class Slots>
__global__ void KSumPartials(Slots slots, double* out) {
  BIND_SLOT(0);
  const double* partials = W.partials;
  __shared__ double red[kThreads];
  double acc = 0.0;
  for (int i = threadIdx.x; i < kPartialBlocks; i += blockDim.x) acc += partials[i];
  const double total = BlockSum(acc, red);
  if (threadIdx.x == 0) *out = total;
}

template <class Slots>
__global__ void KRevert(Slots slots) {
  BIND_SLOT(blockIdx.y);
  const int nmap = S.nmap;
  const int npose = 3 * P.num_poses;
  if (W.state->revert == 0) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nmap + npose; i += gridDim.x * blockDim.x) {
    if (i < nmap) W.map[i] = W.map_backup[i];
    else W.poses[i - nmap] = W.poses_backup[i - nmap];
  }
}

}  // namespace

void LaunchShape::Merge(const LaunchShape& o) {
  slots += o.slots;
  points = std::max(points, o.points);
  eikonal = std::max(eikonal, o.eikonal);
  odom = std::max(odom, o.odom);
  cells = std::max(cells, o.cells);
  nodes = std::max(nodes, o.nodes);
  n = std::max(n, o.n);
  pairs = std::max(pairs, o.pairs);
  state = std::max(state, o.state);
  poses = std::max(poses, o.poses);
  step_items = std::max(step_items, o.step_items);
  max_points_per_pose = std::max(max_points_per_pose, o.max_points_per_pose);
  panel_total = std::max(panel_total, o.panel_total);
  if (levels.size() < o.levels.size()) levels.resize(o.levels.size());
  for (size_t L = 0; L < o.levels.size(); ++L) {
    Level& a = levels[L];
    const Level& b = o.levels[L];
    a.count = std::max(a.count, b.count);
    a.all_update = a.all_update && b.all_update;
    a.max_nb = std::max(a.max_nb, b.max_nb);
    a.max_f = std::max(a.max_f, b.max_f);
    a.max_m = std::max(a.max_m, b.max_m);
    a.tiles = std::max(a.tiles, b.tiles);
    a.has_children = a.has_children || b.has_children;
  }
}

namespace {

template <bool kOne>
auto SlotsOf(const SlotArgs& a) {
  if constexpr (kOne) {
    return OneSlot{*a.one};
  } else {
    return SlotTable{a.table};
  }
}

template <bool kOne>
auto LevelLists(const LaunchShape& sh, int L) {
  if constexpr (kOne) {
    return OneList{sh.level_plan_h[static_cast<size_t>(L)]};
  } else {
    return ListTable{sh.level_plan + static_cast<size_t>(L) * static_cast<size_t>(sh.slots)};
  }
}

template <bool kOne>
auto TileLists(const LaunchShape& sh, int L) {
  if constexpr (kOne) {
    return OneList{sh.tile_plan_h[static_cast<size_t>(L)]};
  } else {
    return ListTable{sh.tile_plan + static_cast<size_t>(L) * static_cast<size_t>(sh.slots)};
  }
}

bool OneOf(const SlotArgs& a, const LaunchShape& sh) { return a.one != nullptr && sh.slots == 1; }


template <bool kOne>
void AssembleImpl(const SlotArgs& a, const LaunchShape& sh, bool check_footprint, hipStream_t s,
                  hipEvent_t before, hipEvent_t after) {
  const auto slots = SlotsOf<kOne>(a);
  const unsigned k = static_cast<unsigned>(sh.slots);
  if (before != nullptr) (void)hipEventRecord(before, s);
  if (sh.points > 0) {
    KAssemblePoints<<<dim3(Blocks(sh.points), k), kThreads, 0, s>>>(slots, check_footprint ? 1 : 0);
  }
  if (after != nullptr) (void)hipEventRecord(after, s);
  if (sh.eikonal > 0) {
    KAssembleEikonal<<<dim3(Blocks(sh.eikonal), k), kThreads, 0, s>>>(slots, nullptr);
  }
  if (sh.odom > 0) {
    KAssembleOdom<<<dim3(Blocks(sh.odom), k), kThreads, 0, s>>>(slots);
  }
}

template <bool kOne>
void NormalSlotsImpl(const SlotArgs& a, const LaunchShape& sh, const LmOptions& o, hipStream_t s) {
  const auto slots = SlotsOf<kOne>(a);
  using Slots = decltype(slots);
  const unsigned k = static_cast<unsigned>(sh.slots);
  if (sh.panel_total > 0) {
    KZeroPanels<<<dim3(Blocks(sh.panel_total, kThreads, 4096), k), kThreads, 0, s>>>(slots);
  }
  int sort_size = 1;
  while (sort_size < sh.max_points_per_pose) sort_size <<= 1;
  const size_t lds = static_cast<size_t>(sort_size) * (8 + 4 + 4 + 4) + 4;
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&KPoseBlock<Slots>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
  KPoseBlock<Slots><<<dim3(sh.poses, k), kThreads, lds, s>>>(slots, o.marquardt_scaling, sort_size,
                                                             g_lds_forms.load() ? 1 : 0);
  KCellGather<<<dim3(Blocks(sh.cells, kThreads / 16), k), kThreads, 0, s>>>(slots);
  KNodeSlots<<<dim3(Blocks(sh.nodes), k), kThreads, 0, s>>>(slots, o.marquardt_scaling);
  if (sh.pairs > 0) {
    KOdomPairs<<<dim3(Blocks(sh.pairs), k), kThreads, 0, s>>>(slots);
  }
}

// Above this many TRSM workgroups a level factors its diagonal blocks in a
// launch of their own instead of once per workgroup.
constexpr int kStagedMinBlocks = 256;

// `panels_done(L)` runs once level L's fronts have their final panel rows
// and kept inverses (before their update matrices, whose launch also moves
// each front's last diagonal block into its panel: the forward solve reads
// the kept inverses, not those blocks), where a forward solve of the level
// may start.
template <bool kOne, class PanelsDone>
void FactorImpl(const SlotArgs& a, const LaunchShape& sh, hipStream_t s, StepEvents* step_events,
                PanelsDone&& panels_done) {
  const auto slots = SlotsOf<kOne>(a);
  const unsigned k = static_cast<unsigned>(sh.slots);
  for (int L = 0; L < static_cast<int>(sh.levels.size()); ++L) {
    const LaunchShape::Level& lv = sh.levels[static_cast<size_t>(L)];
    const int count = lv.count;
    if (count > 0 && lv.has_children) {
      const dim3 grid((lv.max_nb + kWaves - 1) / kWaves, count, k);
      if (g_lds_forms.load() && lv.max_f <= kExtendLdsRows) {
        const int ldf = lv.max_f;  // at most 64 KiB of LDS: no attribute needed
        const size_t lds = static_cast<size_t>(kWaves) * ldf * sizeof(double);
        KExtendAddPanelLds<<<grid, kThreads, lds, s>>>(slots, LevelLists<kOne>(sh, L), ldf);
      } else {
        KExtendAddPanel<<<grid, kThreads, 0, s>>>(slots, LevelLists<kOne>(sh, L));
      }
    }
    if (count > 0) {
      const auto plan = LevelLists<kOne>(sh, L);
      const int nblocks = (lv.max_nb + kNB - 1) / kNB;
      for (int kb = 0; kb < nblocks; ++kb) {
        // The first step's panel: with few fronts each workgroup factors
        // the diagonal block itself (one launch on the critical path), with
        // many the block is factored once per front first. Every later
        // step's diagonal block and panel rows are done by the previous
        // step's trailing update (KGemmAhead).
        if (kb == 0) {
          // Rows below the block: at most f - 1 (a front's block has at
          // least one column); one 64-row workgroup each, at least one.
          const int rows = lv.max_f - 1;
          const int row_blocks = rows > 0 ? (rows + kTile - 1) / kTile : 1;
          const bool staged = static_cast<int64_t>(row_blocks) * count * sh.slots > kStagedMinBlocks;
          const bool timed = step_events != nullptr && step_events->used < step_events->capacity;
          if (timed) (void)hipEventRecord(step_events->events[2 * step_events->used], s);
          if (staged) {
            KDiagBlock<<<dim3(count, k), kThreads, 0, s>>>(slots, plan, kb);
            KPanelStep<false><<<dim3(row_blocks, count, k), kThreads, 0, s>>>(slots, plan, kb);
          } else {
            KPanelStep<true><<<dim3(row_blocks, count, k), kThreads, 0, s>>>(slots, plan, kb);
          }
          if (timed) {
            (void)hipEventRecord(step_events->events[2 * step_events->used + 1], s);
            ++step_events->used;
          }
        }
        // The last step has no trailing columns: it only moves the factored
        // diagonal blocks into the panels, which a single-slot update-matrix
        // launch does where every front has an update matrix.
        if (kOne && kb + 1 == nblocks && lv.all_update && lv.tiles > 0) break;
        const int trail_rows = lv.max_f - (kb + 1) * kNB;
        const int trail_cols = lv.max_nb - (kb + 1) * kNB;
        // With no columns beyond the block in any front only tile (0, 0)
        // has work: moving the factored block into the panel.
        const int tr = trail_rows > 0 && trail_cols > 0 ? (trail_rows + kTile - 1) / kTile : 1;
        const int tc = trail_cols > 0 ? (trail_cols + kTile - 1) / kTile : 1;
        const dim3 grid(tr, tc, static_cast<unsigned>(count) * k);
        const bool gt = step_events != nullptr && step_events->gemm_used < step_events->gemm_capacity;
        if (gt) (void)hipEventRecord(step_events->gemm_events[2 * step_events->gemm_used], s);
        if (kb + 1 < nblocks) {
          KGemmAhead<<<grid, kThreads, 0, s>>>(slots, plan, kb, count);
        } else {
          KGemmLower<<<grid, kThreads, 0, s>>>(slots, plan, kb, count);
        }
        if (gt) (void)hipEventRecord(step_events->gemm_events[2 * step_events->gemm_used++ + 1], s);
      }
      panels_done(L);
      if (lv.tiles > 0) {
        const bool gt = step_events != nullptr && step_events->gemm_used < step_events->gemm_capacity;
        if (gt) (void)hipEventRecord(step_events->gemm_events[2 * step_events->gemm_used], s);
        KGemmUpdate<<<dim3(lv.tiles, k), kThreads, 0, s>>>(slots, TileLists<kOne>(sh, L));
        if (gt) (void)hipEventRecord(step_events->gemm_events[2 * step_events->gemm_used++ + 1], s);
      }
    }
  }
}


// Forward solve of level L's fronts (after their children's). Levels with
// a front of more than one 64-column block: one workgroup per block of each
// front for the own columns, then the rows below them spread over many
// workgroups; other levels: one workgroup per front. For fronts of one
// block both forms give the same values, and the choice follows the widest
// front of the launch, so a front's result does not depend on the other
// fronts or slots of its launch.
template <bool kOne>
void ForwardLevel(const SlotArgs& a, const LaunchShape& sh, hipStream_t s, int L) {
  const auto slots = SlotsOf<kOne>(a);
  const unsigned k = static_cast<unsigned>(sh.slots);
  const LaunchShape::Level& lv = sh.levels[static_cast<size_t>(L)];
  if (lv.count <= 0) return;
  const auto plan = LevelLists<kOne>(sh, L);
  if (lv.max_nb > kNB) {
    const int blocks = (lv.max_nb + kNB - 1) / kNB;
    KForwardBlocks<<<dim3(blocks, lv.count, k), kThreads, 0, s>>>(slots, plan);
    if (lv.max_m > 0) KForwardRest<<<dim3((lv.max_m + 63) / 64, lv.count, k), 64, 0, s>>>(slots, plan);
    return;
  }
  // Few large fronts: one 1024-thread workgroup each for more loads in flight.
  const int threads = lv.count <= 32 ? 1024 : kThreads;
  KForward<<<dim3(lv.count, k), threads, static_cast<size_t>(lv.max_f) * sizeof(double), s>>>(slots, plan, 0);
}

// Backward solves, top level first. Levels with a front of more than one
// block: L21^T x2 over many workgroups, then one workgroup per 64-column
// block of each front (its own summation order for the columns of the
// front, which a front of one block does not have); other levels: one
// workgroup per front.
template <bool kOne>
void BackwardImpl(const SlotArgs& a, const LaunchShape& sh, hipStream_t s) {
  const auto slots = SlotsOf<kOne>(a);
  const unsigned k = static_cast<unsigned>(sh.slots);
  for (int L = static_cast<int>(sh.levels.size()) - 1; L >= 0; --L) {
    const LaunchShape::Level& lv = sh.levels[static_cast<size_t>(L)];
    if (lv.count <= 0) continue;
    const auto plan = LevelLists<kOne>(sh, L);
    if (lv.max_nb > kNB) {
      const int tri = lv.max_m > 0 ? 1 : 0;
      if (tri != 0) KBackwardRest<<<dim3((lv.max_nb + 3) / 4, lv.count, k), kThreads, 0, s>>>(slots, plan);
      const int blocks = (lv.max_nb + kNB - 1) / kNB;
      KBackwardBlocks<<<dim3(blocks, lv.count, k), kThreads, 0, s>>>(slots, plan, tri);
      continue;
    }
    const int threads = lv.count <= 32 ? 1024 : kThreads;
    KBackward<<<dim3(lv.count, k), threads, static_cast<size_t>(lv.max_f) * sizeof(double), s>>>(slots, plan, 0);
  }
}

template <bool kOne>
void SolveImpl(const SlotArgs& a, const LaunchShape& sh, hipStream_t s) {
  for (int L = 0; L < static_cast<int>(sh.levels.size()); ++L) ForwardLevel<kOne>(a, sh, s, L);
  BackwardImpl<kOne>(a, sh, s);
}

// Factorization and solves with every level's forward solve on `aux`, as
// soon as the level's panels are final, beside the factorization of the
// levels above it on `s` (a few large fronts' dependent steps leave most of
// the device free); the backward solves follow on `s` once the forward ones
// are done. The same kernels as the sequential form, in the same order per
// level, so the results are the same.
template <bool kOne>
hipError_t FactorSolveOverlapped(const SlotArgs& a, const LaunchShape& sh, hipStream_t s, hipStream_t aux,
                                 const hipEvent_t* events) {
  hipError_t err = hipSuccess;
  auto check = [&](hipError_t e) {
    if (err == hipSuccess) err = e;
  };
  FactorImpl<kOne>(a, sh, s, nullptr, [&](int L) {
    check(hipEventRecord(events[L], s));
    check(hipStreamWaitEvent(aux, events[L], 0));
    ForwardLevel<kOne>(a, sh, aux, L);
  });
  const hipEvent_t done = events[sh.levels.size()];
  check(hipEventRecord(done, aux));
  check(hipStreamWaitEvent(s, done, 0));
  BackwardImpl<kOne>(a, sh, s);
  return err;
}

template <bool kOne>
void StepAndControlImpl(const SlotArgs& a, const LaunchShape& sh, const LmOptions& o, hipStream_t s) {
  const auto slots = SlotsOf<kOne>(a);
  const unsigned k = static_cast<unsigned>(sh.slots);
  KStepReduce<<<dim3(kPartialBlocks, k), kThreads, 0, s>>>(slots);
  const bool revertible = o.reject_worse_steps != 0 || o.trust_region != 0;
  if (revertible) {
    KBackup<<<dim3(Blocks(sh.state, kThreads, 1024), k), kThreads, 0, s>>>(slots);
  }
  KApplyStep<<<dim3(Blocks(sh.step_items), k), kThreads, 0, s>>>(slots);
  KCost<<<dim3(kPartialBlocks, k), kThreads, 0, s>>>(slots, 0);
  KControl<<<dim3(1, k), kThreads, 0, s>>>(slots, o);
  if (revertible) {
    KRevert<<<dim3(Blocks(sh.state, kThreads, 1024), k), kThreads, 0, s>>>(slots);
  }
}

}  // namespace

void LaunchAssemble(const SlotArgs& slots, const LaunchShape& sh, bool check_footprint, hipStream_t s,
                    hipEvent_t before, hipEvent_t after) {
  if (OneOf(slots, sh)) {
    AssembleImpl<true>(slots, sh, check_footprint, s, before, after);
  } else {
    AssembleImpl<false>(slots, sh, check_footprint, s, before, after);
  }
}

void LaunchAssembleEikonalJ(const SlotArgs& slots, const LaunchShape& sh, double* j_eik, hipStream_t s) {
  if (sh.eikonal > 0) {
    KAssembleEikonal<<<dim3(Blocks(sh.eikonal), 1), kThreads, 0, s>>>(SlotTable{slots.table}, j_eik);
  }
}

void SetLdsForms(bool enabled) { g_lds_forms.store(enabled); }

void LaunchNormalSlots(const SlotArgs& slots, const LaunchShape& sh, const LmOptions& o, hipStream_t s) {
  if (OneOf(slots, sh)) {
    NormalSlotsImpl<true>(slots, sh, o, s);
  } else {
    NormalSlotsImpl<false>(slots, sh, o, s);
  }
}

void LaunchFactor(const SlotArgs& slots, const LaunchShape& sh, hipStream_t s, StepEvents* step_events) {
  if (OneOf(slots, sh)) {
    FactorImpl<true>(slots, sh, s, step_events, [](int) {});
  } else {
    FactorImpl<false>(slots, sh, s, step_events, [](int) {});
  }
}

hipError_t LaunchFactorSolveOverlapped(const SlotArgs& slots, const LaunchShape& sh, hipStream_t s,
                                       hipStream_t aux, const hipEvent_t* events) {
  if (OneOf(slots, sh)) return FactorSolveOverlapped<true>(slots, sh, s, aux, events);
  return FactorSolveOverlapped<false>(slots, sh, s, aux, events);
}

void LaunchSolve(const SlotArgs& slots, const LaunchShape& sh, hipStream_t s) {
  if (OneOf(slots, sh)) {
    SolveImpl<true>(slots, sh, s);
  } else {
    SolveImpl<false>(slots, sh, s);
  }
}

void LaunchCost(const SlotArgs& slots, const LaunchShape& sh, hipStream_t s) {
  KCost<<<dim3(kPartialBlocks, static_cast<unsigned>(sh.slots)), kThreads, 0, s>>>(SlotTable{slots.table}, 1);
}

void LaunchSumPartials(const SlotArgs& slots, double* out, hipStream_t s) {
  KSumPartials<<<1, kThreads, 0, s>>>(SlotTable{slots.table}, out);
}

void LaunchLmInit(const SlotArgs& slots, const LaunchShape& sh, double lambda_init, hipStream_t s) {
  KLmInit<<<dim3(1, static_cast<unsigned>(sh.slots)), kThreads, 0, s>>>(SlotTable{slots.table}, lambda_init);
}

void LaunchGatherStates(const SlotArgs& slots, const LaunchShape& sh, LmState* out, hipStream_t s) {
  const int k = sh.slots;
  KGatherStates<<<(k + 63) / 64, 64, 0, s>>>(SlotTable{slots.table}, k, out);
}

void LaunchStepAndControl(const SlotArgs& slots, const LaunchShape& sh, const LmOptions& o, hipStream_t s) {
  if (OneOf(slots, sh)) {
    StepAndControlImpl<true>(slots, sh, o, s);
  } else {
    StepAndControlImpl<false>(slots, sh, o, s);
  }
}

}  // namespace sdfslam
