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
const char* NormalsLastError();

}  // namespace sdfslam
