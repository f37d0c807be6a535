// Symbolic analysis of the joint-LM normal equations for the device backend.
//
// The reference rebuilds J from triplets every iteration and re-analyses the
// J^T J pattern whenever it changes (src/solvers/lm_solver.cpp:298-350). Here
// the pattern is a fixed superset, valid for a whole "pattern epoch":
//   node-node  9-point stencil over active nodes (every cell couples its 4
//              nodes; eikonal rows couple (w,h),(w+1,h),(w,h+1));
//   pose-node  each pose's footprint: the nodes of the cells its points fall
//              in at the epoch's state, dilated by `margin` cells;
//   pose-pose  3x3 diagonal blocks and the odometry/relation pairs.
// A point that leaves its pose's dilated cells ends the epoch (the device
// raises a flag and the host re-runs this analysis).
//
// Ordering: geometric nested dissection of the node grid; each pose is
// placed in the deepest dissection region that contains its footprint's
// bounding box and is eliminated with that region's separator. Supernodes
// are the dissection tree nodes (leaf blocks and separators with their
// poses). Front structure R_s follows block elimination: R_s = A-pattern of
// the supernode's columns beyond it, united with the children's R_c; the
// parent is the supernode holding min(R_s).
#pragma once

#include <cstdint>
#include <vector>

#include "sdfslam_hip.h"

namespace sdfslam {

struct AnalysisOptions {
  int margin{6};       // footprint dilation in cells
  int leaf_nodes{64};  // dissection stops at regions of at most this many nodes
  // Per-pose dilation (one entry per pose); empty = `margin` for every pose.
  std::vector<int32_t> pose_margin;
};

struct Analysis {
  // Sizes.
  int nx{0}, ny{0}, num_cells{0};
  int num_active{0}, num_poses{0}, n{0};
  // Ordering: new index of each original column, and its inverse.
  std::vector<int32_t> perm, iperm;
  // Supernodes in elimination (post)order.
  int num_sn{0};
  std::vector<int32_t> sn_first;  // num_sn + 1, ranges of new indices
  std::vector<int32_t> sn_m;      // |R_s|
  std::vector<int32_t> sn_f;      // n_s + m_s
  std::vector<int64_t> sn_idx_ptr;
  std::vector<int32_t> sn_idx;  // front index lists: own columns, then R_s (new indices)
  std::vector<int64_t> panel_off;  // (f_s x n_s) column-major panels
  std::vector<int64_t> upd_off;    // (m_s x m_s) column-major update matrices
  int64_t panel_total{0}, upd_total{0};
  std::vector<int32_t> sn_parent, sn_level, snode_of;
  int num_levels{0};
  std::vector<std::vector<int32_t>> children;  // ascending
  // Child -> parent position map of R_c inside the parent's front, and the
  // number of R_c entries that land in the parent's own columns (a prefix).
  std::vector<int64_t> rel_off;  // per supernode (as child), offset into rel
  std::vector<int32_t> rel;
  std::vector<int32_t> rel_split;
  // Node-node slots: 5 per grid node (self, right, up-left, up, up-right);
  // panel offset or -1.
  std::vector<int64_t> nn_off;
  // Eikonal spec at each grid node, or -1.
  std::vector<int32_t> eik_at;
  // Pose footprints: nodes (grid ids, ascending) and the panel offset of the
  // (node, pose x) entry; the y/theta entries follow at +stride, +2*stride.
  std::vector<int64_t> fp_ptr;
  std::vector<int32_t> fp_node;
  std::vector<int64_t> fp_off;
  std::vector<int32_t> fp_stride;
  // Dilated cell set of each pose as a bitmap (num_poses x words_per_pose).
  int words_per_pose{0};
  std::vector<uint32_t> fp_cellbits;
  // Frames whose dilated cells contain each cell (ascending), and the
  // cell's rank among that frame's dilated cells (ascending cell ids).
  std::vector<int64_t> cp_ptr;
  std::vector<int32_t> cp_pose;
  std::vector<int32_t> cp_rank;
  // Each frame's dilated cells as a range of ranks (dl_ptr[P] .. dl_ptr[P + 1])
  // and, per word of its cell bitmap, the number of dilated cells before it.
  std::vector<int64_t> dl_ptr;
  std::vector<int32_t> fp_cellrank;
  // Point rows of each pose (ascending).
  std::vector<int64_t> pr_ptr;
  std::vector<int32_t> pr_row;
  int max_points_per_pose{0};
  // Pose diagonal blocks: 6 offsets per pose (xx, yx, yy, tx, ty, tt).
  std::vector<int64_t> pd_off;
  // Odometry/relation specs touching each pose (spec order), and distinct
  // off-diagonal pose pairs with their specs and 9 offsets each
  // (P component a, Q component b) at index 3a+b.
  std::vector<int64_t> po_ptr;
  std::vector<int32_t> po_spec;
  std::vector<int32_t> op_pose;  // 2 per pair (P < Q)
  std::vector<int64_t> op_ptr;
  std::vector<int32_t> op_spec;
  std::vector<int64_t> op_off;  // 9 per pair
  // Level schedule.
  std::vector<std::vector<int32_t>> level_sn;
  std::vector<int32_t> level_max_nb;  // max n_s per level
  std::vector<int32_t> level_max_f;
  std::vector<int32_t> level_max_m;
  std::vector<int32_t> level_max_children;
  // Statistics.
  int64_t pattern_nnz_lower{0};
  int64_t factor_nnz{0};
  double factor_flops{0.0};
  // Launches of the first 64-column panel step per factorization (one per
  // level) and their algorithmic FP64 flops: per front, the diagonal
  // block's LDL^T (2/3 nb^3) and the TRSM of the rows below it (nb^2 per
  // row).
  int32_t panel_step_launches{0};
  double panel_step_flops{0.0};
  // Trailing-update and update-matrix launches per factorization (one per
  // step and level but the last step of a level whose fronts all have
  // update matrices, one per level with update matrices) and their flops:
  // 2 K per lower-triangle output entry, plus the next step's diagonal
  // block and TRSM that a trailing update does (as above, counted once).
  int32_t gemm_launches{0};
  double gemm_flops{0.0};
};

/// Builds the analysis for the problem at the given host state.
Analysis Analyze(const sdfslam_problem_view& view, const double* map_values, const double* poses,
                 const AnalysisOptions& options);

/// Panel offset of lower entry (i, j), i >= j, in new indices; -1 if (i, j)
/// is outside the front of j's supernode.
int64_t PanelOffset(const Analysis& a, int i, int j);

/// All slots of the fixed lower pattern in original column order:
/// (row, col, panel offset), sorted by (col, row).
struct PatternEntry {
  int32_t row;
  int32_t col;
  int64_t off;
};
std::vector<PatternEntry> LowerPattern(const Analysis& a, const sdfslam_problem_view& view);

/// Symbolic LDL^T of the reference pattern under the analysis's order:
/// `jcols` holds 7 column slots per Jacobian row (-1 = empty; the device
/// assembly's layout), the pattern is every pair of columns sharing a row
/// plus the diagonal (lm_solver.cpp:113-172, 321-330), permuted by a.perm.
/// Outputs the nonzeros of L below the diagonal and sum over columns of
/// c^2 + 3c (c = a column's count below the diagonal): the factorization's
/// flops without the superset pattern's explicit zeros or supernode padding.
void ReferenceFactorCounts(const Analysis& a, int64_t rows, const int32_t* jcols, int64_t* nnz, double* flops);

}  // namespace sdfslam
