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
const char* NormalsLastError();

}  // namespace sdfslam
