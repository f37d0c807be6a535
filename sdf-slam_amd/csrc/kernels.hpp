// Device kernels of the joint-LM backend and their launch wrappers.
// All wrappers enqueue on the given stream and never synchronize.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace sdfslam {

struct GridParams {
  int nx;
  int ny;
  double min_x;
  double min_y;
  double dx;
  double dy;
};

/// LM state kept on the device (SolveLm's locals, lm_solver.cpp:204-215).
struct LmState {
  double lambda;
  double previous_cost;
  double nu;
  double initial_cost;
  double final_cost;
  double last_cost;
  double step_norm;
  int iter;          // iterations finished (result.iterations)
  int halt;          // 1 once converged, out of iterations, or failed
  int converged;
  int overflow;      // a point left its pose's footprint: re-analyse
  int not_spd;       // zero pivot (kSolveSyncFault: a multi-workgroup solve gave up waiting)
  int revert;        // restore the backup after this iteration
};

// not_spd value of a multi-workgroup triangular solve whose workgroups gave
// up waiting for each other (an internal fault, not a pivot).
constexpr int kSolveSyncFault = 2;

struct LmOptions {
  int max_iterations;
  int history_capacity;  // entries of cost_history
  double step_tolerance;
  double lambda_factor;
  int reject_worse_steps;
  int marquardt_scaling;
  int trust_region;
};

/// Point specs, eikonal specs, odometry specs (SoA, device pointers).
struct DeviceProblem {
  GridParams grid;
  int num_cells;
  int num_active;
  int num_poses;
  int n;
  int64_t num_points;
  int64_t num_eikonal;
  int64_t num_odom;
  const int32_t* node_to_active;
  const int32_t* active_to_node;
  const int32_t* pt_frame;
  const double* pt_x;
  const double* pt_y;
  const double* pt_delta;
  const double* pt_sw;
  const int32_t* eik_w;
  const int32_t* eik_h;
  const double* eik_nx;
  const double* eik_ny;
  const double* eik_sw;
  const int32_t* od_i;
  const int32_t* od_j;
  const double* od_dx;
  const double* od_dy;
  const double* od_dth;
  const double* od_sw;
  double huber_delta;
  int smooth_gradient;
  double smooth_gradient_step;
};

/// Per-iteration work buffers.
struct DeviceWork {
  double* map;
  double* poses;
  double* map_backup;
  double* poses_backup;
  double* r_pt;       // point residuals (scaled)
  double* jw;         // 4 x num_points scaled node weights
  double* jp;         // 3 x num_points scaled pose partials
  uint32_t* cellkey;  // cell id or num_cells when the point has no Jacobian row
  double* r_eik;
  double* od_r;   // 3 per spec
  double* od_ji;  // 9 per spec, row-major (residual component, pose component)
  double* od_jj;
  int32_t* seg_cell;   // per frame, its occupied cells (ascending) at the frame's row offset
  double* seg_part;    // 14 x num_points: per occupied cell, 10 weight products + 4 weight x residual
  double* seg_jp;      // 12 x num_points: per occupied cell, the 3 pose partials x each of the 4 weights
  int32_t* seg_count;  // occupied cells per frame
  int32_t* seg_map;    // per frame and dilated-cell rank: the cell's segment (index in seg_cell) or -1
  double* nn;  // 10 x num_cells
  double* rn;  // 4 x num_cells
  double* rhs;  // -J^T r in original column order
  double* step;  // step in original column order
  double* panel;
  double* diag_stage;  // factored 64x64 diagonal blocks in flight, two per supernode
  double* diag_inv;    // L_kk^-1 of every 64x64 diagonal block (identity padded), for the solves
  double* diag_d;      // D of every 64x64 diagonal block (64 per block), for the update matrices
  double* upd;
  double* vec;  // forward-solve vectors, sum of front sizes
  double* xnew;  // solution in elimination order
  double* partials;
  double* cost_history;
  int32_t* pose_overflow;  // per frame: 1 when one of its points left the footprint
  double* solve_tmp;       // L21^T x2 per column (elimination order), backward solve
  int32_t* solve_sync;     // per supernode: ticket, finished, abort, then a flag per 64-column block
  LmState* state;
};

/// Device copies of the symbolic analysis.
struct DeviceAnalysis {
  const int32_t* perm;
  const int32_t* iperm;
  const int32_t* sn_first;
  const int32_t* sn_m;
  const int32_t* sn_f;
  const int64_t* sn_idx_ptr;
  const int32_t* sn_idx;
  const int64_t* panel_off;
  const int64_t* upd_off;
  const int32_t* sn_parent;
  const int64_t* rel_off;
  const int32_t* rel;
  const int32_t* rel_split;
  const int64_t* uinv_off;  // per child: its map in uinv (-1: no entries in the parent's update matrix)
  const int32_t* uinv;      // parent update row -> child update row, or -1
  const int64_t* pinv_off;  // per child: its map in pinv (-1: no entries in the parent's own columns)
  const int32_t* pinv;      // parent own column -> child column, or -1
  const int64_t* nn_off;
  const int32_t* eik_at;
  const int64_t* fp_ptr;
  const int32_t* fp_node;
  const int64_t* fp_off;
  const int32_t* fp_stride;
  int words_per_pose;
  const uint32_t* fp_cellbits;
  const int64_t* pr_ptr;
  const int32_t* pr_row;
  const int64_t* cp_ptr;  // candidate frames of each cell
  const int32_t* cp_pose;
  const int32_t* cp_rank;      // the cell's rank among the candidate frame's dilated cells
  const int64_t* dl_ptr;       // per frame: its range of dilated-cell ranks
  const int32_t* fp_cellrank;  // per frame and bitmap word: dilated cells before the word
  const int64_t* pd_off;
  const int64_t* po_ptr;
  const int32_t* po_spec;
  int num_pairs;
  const int32_t* op_pose;
  const int64_t* op_ptr;
  const int32_t* op_spec;
  const int64_t* op_off;
  const int64_t* children_ptr;  // children of each supernode, ascending
  const int32_t* children;
  const int64_t* inv_off;  // per supernode: its block columns' entries in diag_inv
};

constexpr int kPartialBlocks = 512;

/// One problem of a launch. Every kernel takes a table of slots in device
/// memory and works on the slot named by its last grid dimension, so one
/// launch serves many independent problems (the batched and served forms);
/// a single problem is a table of one slot. A slot's results do not depend
/// on the other slots of the launch or on the launch's grid sizes.
struct Slot {
  DeviceProblem P;
  DeviceAnalysis A;
  DeviceWork W;
  int64_t panel_total;
  int nmap;  // nx * ny
};

/// The supernodes a level kernel works on in one slot: the fronts of a level,
/// or its update-matrix tiles (device pointer and length).
struct ListRef {
  const int32_t* items;
  int32_t count;
};

/// Grid sizes of a launch over a slot table: the largest sizes over its
/// slots, level by level for the factorization and the solves.
struct LaunchShape {
  int slots{0};
  int64_t points{0}, eikonal{0}, odom{0}, cells{0}, nodes{0}, n{0}, pairs{0};
  int64_t state{0};  // nx * ny + 3 * num_poses
  int poses{0};
  int64_t step_items{0};  // num_active + num_poses
  int max_points_per_pose{1};
  int64_t panel_total{0};
  struct Level {
    int count{0}, max_nb{0}, max_f{0}, max_m{0};
    int tiles{0};  // 64 x 64 lower tiles of the fronts' update matrices
    // Every front has an update matrix (in single-slot launches its tile
    // (0, 0) then moves the front's last factored diagonal block into the
    // panel).
    bool all_update{true};
    bool has_children{false};
  };
  std::vector<Level> levels;
  // Launch plans in device memory: level_plan[L * slots + b] = the fronts of
  // level L in slot b (count 0 where a slot has none); tile_plan[L * slots + b] = the update-matrix tiles of level L in slot b
  // (count pairs (front, ti << 16 | tj) at items).
  const ListRef* level_plan{nullptr};
  const ListRef* tile_plan{nullptr};
  // Host copies of the plans, for launches over a single slot.
  std::vector<ListRef> level_plan_h, tile_plan_h;
  /// Grows this shape to cover `o` as well.
  void Merge(const LaunchShape& o);
};

/// The slots of a launch: the device table of sh.slots slots and, when the
/// launch has one slot, its host copy (the kernels then receive the slot and
/// their lists by value, with no memory indirection before they start).
struct SlotArgs {
  const Slot* table;
  const Slot* one;
};

/// Point, eikonal and odometry rows. When `before`/`after` are given they are
/// recorded around the point-row kernel alone.
void LaunchAssemble(const SlotArgs& slots, const LaunchShape& sh, bool check_footprint, hipStream_t s,
                    hipEvent_t before = nullptr, hipEvent_t after = nullptr);
/// Zeroes the factor panels and forms the damped J^T J and -J^T r in them.
void LaunchNormalSlots(const SlotArgs& slots, const LaunchShape& sh, const LmOptions& o, hipStream_t s);
/// Process-wide: the pose-block and panel extend-add kernels' LDS forms
/// (default) or their global-memory forms, which give the same results (for
/// tests); launches made or captured after the call follow it.
void SetLdsForms(bool enabled);

/// Optional timing of the panel steps: pairs (events[2k], events[2k+1])
/// bracket the diagonal-block and panel-step kernels of one 64-column step; pairs
/// of gemm_events bracket one trailing-update or update-matrix launch.
struct StepEvents {
  hipEvent_t* events;
  int capacity;  // pairs
  int used;
  hipEvent_t* gemm_events;
  int gemm_capacity;  // pairs
  int gemm_used;
};
/// Multifrontal LDL^T, level by level (extend-add, panel steps, trailing
/// updates, update matrices).
void LaunchFactor(const SlotArgs& slots, const LaunchShape& sh, hipStream_t s,
                  StepEvents* step_events = nullptr);
/// Forward and backward solves, level by level; the step in original order.
void LaunchSolve(const SlotArgs& slots, const LaunchShape& sh, hipStream_t s);
/// LaunchFactor then LaunchSolve, with each level's forward solve on `aux`
/// as soon as the level's panels are final, beside the factorization of
/// the levels above it (same kernels, same results). `events` holds
/// sh.levels.size() + 1 events (recorded and waited on here; they may be
/// reused once this work is queued). Returns the first event error.
hipError_t LaunchFactorSolveOverlapped(const SlotArgs& slots, const LaunchShape& sh, hipStream_t s,
                                       hipStream_t aux, const hipEvent_t* events);

/// Scaled eikonal Jacobian entries (3 per row) into j_eik (one slot).
void LaunchAssembleEikonalJ(const SlotArgs& slots, const LaunchShape& sh, double* j_eik, hipStream_t s);
/// Cost partials of the current state (ignores the halt flag).
void LaunchCost(const SlotArgs& slots, const LaunchShape& sh, hipStream_t s);
/// Sums the kPartialBlocks cost partials of slot 0 into *out (fixed order).
void LaunchSumPartials(const SlotArgs& slots, double* out, hipStream_t s);
/// LM state from the cost partials (one slot per workgroup).
void LaunchLmInit(const SlotArgs& slots, const LaunchShape& sh, double lambda_init, hipStream_t s);
/// LM state of every slot into out[0 .. sh.slots).
void LaunchGatherStates(const SlotArgs& slots, const LaunchShape& sh, LmState* out, hipStream_t s);
void LaunchStepAndControl(const SlotArgs& slots, const LaunchShape& sh, const LmOptions& o,
                          hipStream_t s);

}  // namespace sdfslam
