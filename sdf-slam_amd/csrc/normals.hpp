// Device normal estimation of the problem build (see normals.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "sdfslam_hip.h"

namespace sdfslam {

/// Node normals and eikonal scales (nx * ny each, node-major, the normal as
/// x, y pairs) from scan points in world coordinates and their sensor
/// origins (x, y pairs): kNN PCA normal per point over all points, then the
/// nearest point's normal per node, with the host builder's values bit for
/// bit. Returns an SDFSLAM_* status; NormalsLastError() names a failure.
int NodeNormalsOnDevice(const double* points_xy, const double* origins_xy, int64_t num_points, int32_t k,
                        const sdfslam_grid& grid, int32_t method, double corner_threshold,
                        double* node_normal_xy, double* node_scale);
/// Map values (nx * ny, node-major) seeded from scan points in world
/// coordinates and their sensor origins (x, y pairs): the distance to the
/// nearest point (ties to the lower index), negated behind the surface when
/// signed_distance != 0 (InitializeFromScans, map_init.cpp:31-69). With no
/// points the values are left unchanged.
int InitFromScansOnDevice(const double* points_xy, const double* origins_xy, int64_t num_points,
                          const sdfslam_grid& grid, int32_t signed_distance, double* map_values);
/// InitFromScansOnDevice without allocation: device pointers (points and
/// origins as x, y pairs, map values node-major), work enqueued on `stream`
/// with no host synchronisation; every intermediate buffer lives in one
/// file-scope device array per GPU. At most kNoAllocMaxPoints points.
constexpr int64_t kNoAllocMaxPoints = int64_t{1} << 19;
int InitFromScansNoAlloc(const double* points_xy, const double* origins_xy, int64_t num_points,
                         const sdfslam_grid& grid, int32_t signed_distance, double* map_values,
                         hipStream_t stream);
/// Inputs of the eikonal normals of a device-resident problem (device
/// pointers): its scan points are the point rows scan_rows[0 .. scan_points)
/// (sensor coordinates sensor_x/sensor_y, frame index per row), placed by
/// poses (3 per frame) with cos_sin (2 per frame, the host's std::cos and
/// std::sin of each pose angle); the eikonal specs (eik_w, eik_h) receive
/// their node's normal in eik_nx / eik_ny.
struct EikonalNormalsJob {
  const int32_t* scan_rows;
  int64_t scan_points;
  const int32_t* frame;
  const double* sensor_x;
  const double* sensor_y;
  const double* cos_sin;
  const double* poses;
  int nx, ny;
  double min_x, min_y, dx, dy;
  int32_t k;
  int32_t method;
  double corner_threshold;
  const int32_t* eik_w;
  const int32_t* eik_h;
  int64_t num_eikonal;
  double* eik_nx;
  double* eik_ny;
};
/// Device bytes of the work buffer EnqueueEikonalNormals needs.
size_t EikonalNormalsWorkBytes(int64_t scan_points, int64_t nodes);
/// Enqueues the eikonal normals on `stream` (pca or hybrid method: their
/// eikonal scale is 1, so the specs themselves do not change): the values
/// AssignNodeNormals(ComputeScanNormals(...)) gives on the host, bit for bit.
int EnqueueEikonalNormals(const EikonalNormalsJob& job, void* work, hipStream_t stream);
/// Point rows of frames appended to a device-resident problem (device
/// pointers): frame f (f = 0 .. frames-1, problem frame first_frame + f) has
/// raw scan points scan_xy[2 * scan_off[f] ..] (x, y pairs, scan_off[f + 1] -
/// scan_off[f] of them), its rows start at row_off[f] and its entries of the
/// scan-row list at list_off[f]. Rows per frame (Problem::Problem,
/// problem.cpp:26-37): the scan points in order (expected SDF 0, weight
/// sqrt_w_scan), then the hallucinated samples of every point longer than
/// 1e-9 (hallucination.cpp:5-28: for j = 1 .. per_side the free-space sample
/// at +j * step toward the sensor, then, with both directions, the occluded
/// one). At most 65,535 frames per job.
struct FrameRowsJob {
  int32_t frames;
  int32_t first_frame;
  const double* scan_xy;
  const int64_t* scan_off;
  const int64_t* row_off;
  const int64_t* list_off;
  int32_t per_side;
  int32_t both_directions;
  double step;
  double sqrt_w_scan;
  double sqrt_w_hallucination;
  int32_t* frame;
  double* x;
  double* y;
  double* expected_sdf;
  double* sqrt_weight;
  int32_t* scan_list;
};
int EnqueueFrameRows(const FrameRowsJob& job, hipStream_t stream);
/// Active-region marks (problem.cpp:39-70): marks[node] = 1 for every node
/// within `margin` cells (and the cell's own +1) of a cell holding a point
/// row placed at its frame's pose (cos_sin as in EikonalNormalsJob); the
/// marks are cleared first. Device pointers; num_rows point rows.
int EnqueueActiveMarks(const int32_t* frame, const double* x, const double* y, int64_t num_rows,
                       const double* cos_sin, const double* poses, const sdfslam_grid& grid, int32_t margin,
                       uint8_t* marks, hipStream_t stream);
const char* NormalsLastError();

}  // namespace sdfslam
