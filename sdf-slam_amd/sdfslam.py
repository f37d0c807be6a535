"""Python bindings (ctypes) for the MI355X joint-LM backend and its host library.

Mirrors the C ABIs in include/sdfslam_hip.h (libsdfslam_hip.so) and
include/sdfslam_host.h (libsdfslam_host.so). The libraries are built in-tree
into sdf-slam_amd/lib by `make -C sdf-slam_amd`; importing this module fails
loudly when they are missing. There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")

SDFSLAM_OK = 0
SDFSLAM_ERR_INVALID_ARGUMENT = 1
SDFSLAM_ERR_NOT_SPD = 2
SDFSLAM_ERR_HIP = 3
SDFSLAM_ERR_OOM = 4
SDFSLAM_ERR_NO_DEVICE = 5
SDFSLAM_ERR_INTERNAL = 6

_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_f64p = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("min_x", C.c_double),
                ("min_y", C.c_double), ("dx", C.c_double), ("dy", C.c_double)]


class ProblemView(C.Structure):
    _fields_ = [
        ("grid", Grid), ("num_poses", C.c_int32), ("num_active_nodes", C.c_int32),
        ("node_to_active", _i32p),
        ("num_points", C.c_int64), ("point_frame", _i32p), ("point_x", _f64p),
        ("point_y", _f64p), ("point_expected_sdf", _f64p), ("point_sqrt_weight", _f64p),
        ("num_eikonal", C.c_int64), ("eik_w", _i32p), ("eik_h", _i32p), ("eik_nx", _f64p),
        ("eik_ny", _f64p), ("eik_sqrt_weight", _f64p),
        ("num_odom", C.c_int64), ("odom_i", _i32p), ("odom_j", _i32p), ("odom_dx", _f64p),
        ("odom_dy", _f64p), ("odom_dtheta", _f64p), ("odom_sqrt_weight", _f64p),
        ("huber_delta", C.c_double), ("smooth_gradient", C.c_int32),
        ("smooth_gradient_step", C.c_double),
    ]


class SolverOptions(C.Structure):
    _fields_ = [("max_iterations", C.c_int32), ("step_tolerance", C.c_double),
                ("lambda_init", C.c_double), ("lambda_factor", C.c_double),
                ("reject_worse_steps", C.c_int32), ("marquardt_scaling", C.c_int32),
                ("trust_region", C.c_int32), ("num_threads", C.c_int32),
                ("linear_solver", C.c_int32)]


class SolveResult(C.Structure):
    _fields_ = [("converged", C.c_int32), ("iterations", C.c_int32),
                ("initial_cost", C.c_double), ("final_cost", C.c_double),
                ("duration_seconds", C.c_double), ("cost_history", _f64p),
                ("cost_history_capacity", C.c_int32), ("cost_history_length", C.c_int32)]


class AnalysisInfo(C.Structure):
    _fields_ = [("num_parameters", C.c_int64), ("num_residuals", C.c_int64),
                ("pattern_nnz_lower", C.c_int64), ("num_supernodes", C.c_int64),
                ("factor_nnz", C.c_int64), ("factor_flops", C.c_double),
                ("update_doubles", C.c_int64), ("tree_levels", C.c_int32),
                ("pattern_epochs", C.c_int32), ("analysis_seconds", C.c_double),
                ("device_bytes", C.c_int64), ("panel_step_launches", C.c_int32),
                ("panel_step_flops", C.c_double), ("speculative_epochs", C.c_int32),
                ("gemm_launches", C.c_int32), ("gemm_flops", C.c_double),
                ("diag_blocks", C.c_int64)]


class PhaseTimes(C.Structure):
    _fields_ = [("assembly_ms", C.c_double), ("normal_ms", C.c_double),
                ("factor_ms", C.c_double), ("solve_ms", C.c_double),
                ("update_ms", C.c_double), ("total_ms", C.c_double),
                ("point_assembly_kernel_ms", C.c_double), ("panel_step_kernel_ms", C.c_double),
                ("gemm_kernel_ms", C.c_double)]


class ProblemOptions(C.Structure):
    _fields_ = [
        ("weight_scan", C.c_double), ("weight_hallucination", C.c_double),
        ("weight_eikonal", C.c_double), ("weight_odometry", C.c_double),
        ("weight_relation", C.c_double),
        ("hallucination_points_per_scan_point", C.c_int32),
        ("hallucination_step_size", C.c_double), ("hallucination_both_directions", C.c_int32),
        ("normals_method", C.c_int32), ("normals_k_neighbors", C.c_int32),
        ("normals_per_scan", C.c_int32), ("normals_corner_threshold", C.c_double),
        ("huber_delta", C.c_double), ("smooth_gradient", C.c_int32),
        ("smooth_gradient_step", C.c_double), ("active_region", C.c_int32),
        ("active_margin", C.c_int32),
    ]


# Exported symbols of each library, as include/*.h declares them.
class IncrementRecord(C.Structure):
    _fields_ = [("frames", C.c_int32), ("iterations", C.c_int32), ("initial_cost", C.c_double),
                ("final_cost", C.c_double), ("build_seconds", C.c_double),
                ("solve_seconds", C.c_double)]


class ServerOptions(C.Structure):
    _fields_ = [("workers", C.c_int32), ("max_batch", C.c_int32), ("pipelines", C.c_int32),
                ("device_budget_bytes", C.c_int64), ("max_in_flight", C.c_int32),
                ("device_normals", C.c_int32)]


class RunOptions(C.Structure):
    _fields_ = [("incremental", C.c_int32), ("increment_size", C.c_int32), ("max_frames", C.c_int32),
                ("snapshot_every", C.c_int32), ("init_from_scans", C.c_int32), ("init_signed", C.c_int32)]


class ServedOptions(C.Structure):
    _fields_ = [("sessions", C.c_int32), ("rate_hz", C.c_double), ("duration_s", C.c_double),
                ("workers", C.c_int32), ("max_batch", C.c_int32), ("pipelines", C.c_int32),
                ("seed0", C.c_uint64), ("max_in_flight", C.c_int32),
                ("device_normals", C.c_int32), ("start_frame", C.c_int32)]


class ServedResult(C.Structure):
    _fields_ = [("increments", C.c_int64), ("lm_iterations", C.c_int64),
                ("wall_seconds", C.c_double), ("increments_per_second", C.c_double),
                ("latency_p50_ms", C.c_double), ("latency_p99_ms", C.c_double),
                ("latency_mean_ms", C.c_double), ("build_mean_ms", C.c_double),
                ("solve_mean_ms", C.c_double), ("device_bytes", C.c_int64),
                ("batches", C.c_int64), ("mean_batch", C.c_double),
                ("prepare_batch_ms", C.c_double), ("solve_batch_ms", C.c_double),
                ("collect_batch_ms", C.c_double), ("capture_batch_ms", C.c_double),
                ("analysis_batch_ms", C.c_double), ("epochs_rebuilt", C.c_int64),
                ("latency_p50_after_first_ms", C.c_double), ("latency_p99_after_first_ms", C.c_double),
                ("first_increment_mean_ms", C.c_double), ("queue_p99_ms", C.c_double),
                ("prepare_p99_ms", C.c_double), ("solve_wait_p99_ms", C.c_double), ("solve_p99_ms", C.c_double)]


class NormalsRequest(C.Structure):
    _fields_ = [("k_neighbors", C.c_int32), ("method", C.c_int32), ("corner_threshold", C.c_double),
                ("frame_cos_sin", C.POINTER(C.c_double))]


class FrameOptions(C.Structure):
    _fields_ = [("hallucination_points", C.c_int32), ("hallucination_step", C.c_double),
                ("hallucination_both_directions", C.c_int32), ("weight_scan", C.c_double),
                ("weight_hallucination", C.c_double), ("weight_eikonal", C.c_double),
                ("weight_odometry", C.c_double), ("active_region", C.c_int32),
                ("active_margin", C.c_int32), ("normals", NormalsRequest)]


def frame_options(po):
    """FrameOptions of a session with ProblemOptions po (pca or hybrid normals)."""
    return FrameOptions(po.hallucination_points_per_scan_point, po.hallucination_step_size,
                        po.hallucination_both_directions, po.weight_scan, po.weight_hallucination,
                        po.weight_eikonal, po.weight_odometry, po.active_region, po.active_margin,
                        NormalsRequest(po.normals_k_neighbors, po.normals_method,
                                       po.normals_corner_threshold, None))


class BatchInfo(C.Structure):
    _fields_ = [("graph_captures", C.c_int32), ("graph_launches", C.c_int32),
                ("epochs_rebuilt", C.c_int32), ("setup_seconds", C.c_double),
                ("capture_seconds", C.c_double), ("analysis_seconds", C.c_double),
                ("run_seconds", C.c_double)]


class BatchOptions(C.Structure):
    _fields_ = [("instances", C.c_int32), ("iterations", C.c_int32), ("max_batch", C.c_int32),
                ("workers", C.c_int32), ("seed0", C.c_uint64)]


class BatchShape(C.Structure):
    _fields_ = [("max_batch", C.c_int32), ("workers", C.c_int32), ("lm_iterations_per_second", C.c_double)]


class BatchResult(C.Structure):
    _fields_ = [("lm_iterations", C.c_int64), ("unknowns", C.c_int64),
                ("wall_seconds", C.c_double), ("lm_iterations_per_second", C.c_double)]


# Solve function of one increment (sdfslam_solve_fn): user, view, options,
# map values, poses, result -> status.
SOLVE_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(ProblemView), C.POINTER(SolverOptions),
                       C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(SolveResult))


HIP_SYMBOLS = [
    "sdfslam_hip_status_string", "sdfslam_hip_last_error", "sdfslam_hip_solve",
    "sdfslam_hip_context_create", "sdfslam_hip_context_destroy", "sdfslam_hip_context_trim",
    "sdfslam_hip_set_state",
    "sdfslam_hip_get_state", "sdfslam_hip_analysis", "sdfslam_hip_assemble",
    "sdfslam_hip_normal_equations", "sdfslam_hip_linear_step", "sdfslam_hip_cost",
    "sdfslam_hip_context_solve", "sdfslam_hip_context_iterate", "sdfslam_hip_phase_times_get",
    "sdfslam_hip_context_update", "sdfslam_hip_device_memory", "sdfslam_hip_batch_solve",
    "sdfslam_hip_pose_margins", "sdfslam_hip_node_normals", "sdfslam_hip_init_from_scans",
    "sdfslam_hip_init_from_scans_noalloc", "sdfslam_hip_reference_factor_counts",
    "sdfslam_hip_elimination_order", "sdfslam_hip_context_update_device_normals",
    "sdfslam_hip_context_append_frames", "sdfslam_hip_context_point_rows", "sdfslam_hip_set_graphs",
    "sdfslam_hip_set_lds_forms",
]
HOST_SYMBOLS = [
    "sdfslam_host_default_problem_options", "sdfslam_host_default_solver_options",
    "sdfslam_host_last_error", "sdfslam_host_problem_create", "sdfslam_host_workload_create",
    "sdfslam_host_problem_destroy", "sdfslam_host_problem_view", "sdfslam_host_problem_get_state",
    "sdfslam_host_problem_set_state", "sdfslam_host_workload_info", "sdfslam_host_solve",
    "sdfslam_host_parse_solver_yaml", "sdfslam_host_run_incremental",
    "sdfslam_host_dataset_create", "sdfslam_host_dataset_destroy", "sdfslam_host_dataset_info",
    "sdfslam_host_dataset_frame", "sdfslam_server_create", "sdfslam_server_destroy",
    "sdfslam_server_open_session", "sdfslam_server_submit", "sdfslam_server_wait",
    "sdfslam_server_state", "sdfslam_server_close_session", "sdfslam_server_session_info",
    "sdfslam_host_served_benchmark", "sdfslam_host_batch_benchmark",
    "sdfslam_host_node_normals", "sdfslam_host_run", "sdfslam_host_init_from_scans",
    "sdfslam_host_batch_autotune",
]


def _load(name):
    path = os.path.join(LIB_DIR, name)
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built; run `make -C sdf-slam_amd` (hipcc, gfx950)")
    return C.CDLL(path, mode=C.RTLD_GLOBAL)


hip = _load("libsdfslam_hip.so")
host = _load("libsdfslam_host.so")

hip.sdfslam_hip_status_string.restype = C.c_char_p
hip.sdfslam_hip_status_string.argtypes = [C.c_int]
hip.sdfslam_hip_last_error.restype = C.c_char_p
hip.sdfslam_hip_last_error.argtypes = []
hip.sdfslam_hip_solve.argtypes = [C.POINTER(ProblemView), C.POINTER(SolverOptions), _f64p, _f64p,
                                  C.POINTER(SolveResult)]
hip.sdfslam_hip_context_create.argtypes = [C.POINTER(ProblemView), C.POINTER(SolverOptions),
                                           _f64p, _f64p, C.POINTER(C.c_void_p)]
hip.sdfslam_hip_context_destroy.argtypes = [C.c_void_p]
hip.sdfslam_hip_context_trim.argtypes = [C.c_void_p]
hip.sdfslam_hip_reference_factor_counts.argtypes = [C.c_void_p, _i64p, _f64p]
hip.sdfslam_hip_elimination_order.argtypes = [C.c_void_p, _i32p]
hip.sdfslam_hip_context_update_device_normals.argtypes = [C.c_void_p, C.POINTER(ProblemView), _f64p, _f64p,
                                                             C.POINTER(NormalsRequest)]
hip.sdfslam_hip_context_append_frames.argtypes = [C.c_void_p, C.c_int32, _i64p, _f64p, _f64p,
                                                  C.POINTER(FrameOptions), _f64p, _f64p]
hip.sdfslam_hip_context_point_rows.argtypes = [C.c_void_p, C.c_int64, _i64p, _i32p, _f64p, _f64p, _f64p, _f64p]
hip.sdfslam_hip_set_state.argtypes = [C.c_void_p, _f64p, _f64p]
hip.sdfslam_hip_get_state.argtypes = [C.c_void_p, _f64p, _f64p]
hip.sdfslam_hip_analysis.argtypes = [C.c_void_p, C.POINTER(AnalysisInfo)]
hip.sdfslam_hip_assemble.argtypes = [C.c_void_p, _f64p, _i32p, _f64p]
hip.sdfslam_hip_normal_equations.argtypes = [C.c_void_p, C.c_double, C.c_int64, _i64p, _i32p,
                                             _i32p, _f64p, _u8p, _f64p]
hip.sdfslam_hip_linear_step.argtypes = [C.c_void_p, C.c_double, _f64p]
hip.sdfslam_hip_cost.argtypes = [C.c_void_p, _f64p]
hip.sdfslam_hip_context_solve.argtypes = [C.c_void_p, C.POINTER(SolverOptions),
                                          C.POINTER(SolveResult)]
hip.sdfslam_hip_context_iterate.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
hip.sdfslam_hip_phase_times_get.argtypes = [C.c_void_p, C.POINTER(PhaseTimes)]
hip.sdfslam_hip_context_update.argtypes = [C.c_void_p, C.POINTER(ProblemView), _f64p, _f64p]
hip.sdfslam_hip_pose_margins.argtypes = [C.c_void_p, _i32p, C.c_int32, _i32p]
hip.sdfslam_hip_node_normals.argtypes = [_f64p, _f64p, C.c_int64, C.c_int32, C.POINTER(Grid), C.c_int32,
                                         C.c_double, _f64p, _f64p]
hip.sdfslam_hip_init_from_scans.argtypes = [_f64p, _f64p, C.c_int64, C.POINTER(Grid), C.c_int32, _f64p]
hip.sdfslam_hip_init_from_scans_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(Grid),
                                                    C.c_int32, C.c_void_p, C.c_void_p]
hip.sdfslam_hip_batch_solve.argtypes = [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(SolverOptions),
                                        C.POINTER(SolveResult), _i32p, C.POINTER(BatchInfo)]

host.sdfslam_host_last_error.restype = C.c_char_p
host.sdfslam_host_default_problem_options.argtypes = [C.POINTER(ProblemOptions)]
host.sdfslam_host_default_solver_options.argtypes = [C.POINTER(SolverOptions)]
host.sdfslam_host_problem_create.argtypes = [
    C.POINTER(Grid), _f64p, C.c_int32, _f64p, _i64p, _f64p, _f64p, C.c_int32, _i32p, _f64p,
    _f64p, C.POINTER(ProblemOptions), C.POINTER(C.c_void_p)]
host.sdfslam_host_workload_create.argtypes = [C.c_char_p, C.c_uint64, C.POINTER(C.c_void_p)]
host.sdfslam_host_problem_destroy.argtypes = [C.c_void_p]
host.sdfslam_host_problem_view.argtypes = [C.c_void_p, C.POINTER(ProblemView)]
host.sdfslam_host_problem_get_state.argtypes = [C.c_void_p, _f64p, _f64p]
host.sdfslam_host_problem_set_state.argtypes = [C.c_void_p, _f64p, _f64p]
host.sdfslam_host_workload_info.argtypes = [C.c_void_p, C.POINTER(SolverOptions), _i32p, _i32p,
                                            _f64p]
host.sdfslam_host_solve.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(SolverOptions),
                                    C.POINTER(SolveResult)]
host.sdfslam_host_run_incremental.argtypes = [
    C.c_char_p, C.c_uint64, C.POINTER(SolverOptions), C.c_int32, C.c_int32, SOLVE_FN, C.c_void_p,
    C.POINTER(SolveResult), C.POINTER(IncrementRecord), C.c_int32, _i32p, _f64p, _f64p]
host.sdfslam_host_run.argtypes = [C.c_char_p, C.c_uint64, C.POINTER(RunOptions), C.POINTER(SolverOptions),
                                  SOLVE_FN, C.c_void_p, C.c_char_p, C.POINTER(SolveResult)]
host.sdfslam_host_init_from_scans.argtypes = [C.POINTER(Grid), _f64p, _f64p, _i64p, C.c_int32, _f64p,
                                              C.c_int32, C.c_int32]
host.sdfslam_host_dataset_create.argtypes = [C.c_char_p, C.c_uint64, C.POINTER(C.c_void_p)]
host.sdfslam_host_dataset_destroy.argtypes = [C.c_void_p]
host.sdfslam_host_dataset_destroy.restype = None
host.sdfslam_host_dataset_info.argtypes = [C.c_void_p, _i32p, C.POINTER(Grid),
                                           C.POINTER(ProblemOptions), C.POINTER(SolverOptions)]
host.sdfslam_host_dataset_frame.argtypes = [C.c_void_p, C.c_int32, _f64p, C.c_int64, _i64p,
                                            _f64p, _f64p]
host.sdfslam_server_create.argtypes = [C.POINTER(ServerOptions), C.POINTER(C.c_void_p)]
host.sdfslam_server_destroy.argtypes = [C.c_void_p]
host.sdfslam_server_destroy.restype = None
host.sdfslam_server_open_session.argtypes = [C.c_void_p, C.POINTER(Grid), _f64p,
                                             C.POINTER(ProblemOptions), C.POINTER(SolverOptions),
                                             _f64p, _f64p, C.c_int64, _i32p]
host.sdfslam_server_submit.argtypes = [C.c_void_p, C.c_int32, _f64p, C.c_int64, _f64p, _i64p]
host.sdfslam_server_wait.argtypes = [C.c_void_p, C.c_int32, C.c_int64, C.POINTER(IncrementRecord),
                                     _f64p]
host.sdfslam_server_state.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _i32p, _f64p, _f64p]
host.sdfslam_server_close_session.argtypes = [C.c_void_p, C.c_int32]
host.sdfslam_server_session_info.argtypes = [C.c_void_p, C.c_int32, _i32p, _i32p, _i64p, _i64p, _i64p]
host.sdfslam_host_served_benchmark.argtypes = [C.POINTER(ServedOptions), C.POINTER(ServedResult)]
host.sdfslam_host_batch_benchmark.argtypes = [C.POINTER(BatchOptions), C.POINTER(BatchResult)]
host.sdfslam_host_batch_autotune.argtypes = [C.POINTER(BatchOptions), _i32p, C.c_int32, _i32p, C.c_int32, C.c_int32,
                                             C.POINTER(BatchShape), C.POINTER(BatchShape)]
host.sdfslam_host_node_normals.argtypes = [_f64p, _f64p, C.c_int64, C.c_int32, C.POINTER(Grid), C.c_int32,
                                           C.c_double, _f64p, _f64p]
host.sdfslam_host_parse_solver_yaml.argtypes = [C.c_char_p, C.POINTER(SolverOptions), C.c_char_p,
                                                C.c_int32]


class SdfSlamError(RuntimeError):
    """A non-zero status from either library; `.status` holds the code."""

    def __init__(self, status, message):
        super().__init__(f"[{status}] {message}")
        self.status = status


class FactorizationError(SdfSlamError):
    """Zero pivot: the reference's "sparse Cholesky factorization failed"."""


def _check_hip(status):
    if status != SDFSLAM_OK:
        msg = hip.sdfslam_hip_last_error().decode()
        if status == SDFSLAM_ERR_NOT_SPD:
            raise FactorizationError(status, msg or "sparse Cholesky factorization failed")
        if status == SDFSLAM_ERR_INVALID_ARGUMENT:
            raise ValueError(msg)
        raise SdfSlamError(status, msg)


def _check_host(status):
    if status != SDFSLAM_OK:
        msg = host.sdfslam_host_last_error().decode()
        if status == SDFSLAM_ERR_NOT_SPD:
            raise FactorizationError(status, msg)
        if status == SDFSLAM_ERR_INVALID_ARGUMENT:
            raise ValueError(msg)
        raise SdfSlamError(status, msg)


def ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def default_solver_options(**kw):
    o = SolverOptions()
    host.sdfslam_host_default_solver_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def default_problem_options(**kw):
    o = ProblemOptions()
    host.sdfslam_host_default_problem_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def parse_solver_yaml(path):
    o = SolverOptions()
    buf = C.create_string_buffer(16)
    _check_host(host.sdfslam_host_parse_solver_yaml(path.encode(), C.byref(o), buf, 16))
    return o, buf.value.decode()


class Problem:
    """A Problem built by the host library (Problem::Problem, problem.cpp:11-115)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)
        self.view = ProblemView()
        _check_host(host.sdfslam_host_problem_view(self._h, C.byref(self.view)))

    @classmethod
    def workload(cls, name, seed=1):
        h = C.c_void_p()
        _check_host(host.sdfslam_host_workload_create(name.encode(), seed, C.byref(h)))
        return cls(h.value)

    @classmethod
    def build(cls, grid, poses, scans, odometry=None, relations=(), options=None, map_values=None):
        """grid: (nx, ny, min_x, min_y, dx, dy); poses: (N, 3); scans: list of (k, 2) arrays;
        relations: iterable of (i, j, (dx, dy, dtheta), sqrt_weight)."""
        g = Grid(*grid)
        poses = np.ascontiguousarray(poses, dtype=np.float64)
        offsets = np.zeros(len(scans) + 1, dtype=np.int64)
        offsets[1:] = np.cumsum([len(s) for s in scans])
        xy = (np.ascontiguousarray(np.concatenate([np.asarray(s, float).reshape(-1, 2) for s in scans]))
              if offsets[-1] > 0 else np.zeros((1, 2)))
        odo = None if odometry is None else np.ascontiguousarray(odometry, dtype=np.float64)
        rel = list(relations)
        rf = np.array([[r[0], r[1]] for r in rel], dtype=np.int32).reshape(-1, 2)
        rm = np.array([r[2] for r in rel], dtype=np.float64).reshape(-1, 3)
        rw = np.array([r[3] for r in rel], dtype=np.float64)
        opts = options if options is not None else default_problem_options()
        mv = None if map_values is None else np.ascontiguousarray(map_values, dtype=np.float64)
        h = C.c_void_p()
        _check_host(host.sdfslam_host_problem_create(
            C.byref(g), ptr(mv, _f64p), len(poses), ptr(poses, _f64p), ptr(offsets, _i64p),
            ptr(xy, _f64p), ptr(odo, _f64p), len(rel), ptr(rf, _i32p), ptr(rm, _f64p),
            ptr(rw, _f64p), C.byref(opts), C.byref(h)))
        return cls(h.value)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            host.sdfslam_host_problem_destroy(h)
            self._h = None

    @property
    def num_parameters(self):
        return self.view.num_active_nodes + 3 * (self.view.num_poses - 1)

    @property
    def num_residuals(self):
        v = self.view
        return v.num_points + v.num_eikonal + 3 * v.num_odom

    def state(self):
        g = self.view.grid
        m = np.zeros(g.nx * g.ny)
        p = np.zeros(3 * self.view.num_poses)
        _check_host(host.sdfslam_host_problem_get_state(self._h, ptr(m, _f64p), ptr(p, _f64p)))
        return m, p

    def set_state(self, map_values=None, poses=None):
        mv = None if map_values is None else np.ascontiguousarray(map_values, dtype=np.float64)
        pv = None if poses is None else np.ascontiguousarray(poses, dtype=np.float64)
        _check_host(host.sdfslam_host_problem_set_state(self._h, ptr(mv, _f64p), ptr(pv, _f64p)))

    def workload_info(self):
        o = SolverOptions()
        inc = C.c_int32()
        ipi = C.c_int32()
        gt = np.zeros(3 * self.view.num_poses)
        _check_host(host.sdfslam_host_workload_info(self._h, C.byref(o), C.byref(inc),
                                                    C.byref(ipi), ptr(gt, _f64p)))
        return o, bool(inc.value), ipi.value, gt.reshape(-1, 3)

    def solve(self, options, backend="hip"):
        """Solve(problem, options) through the host library's dispatch."""
        hist = np.zeros(max(options.max_iterations, 0) + 1)
        r = SolveResult()
        r.cost_history = ptr(hist, _f64p)
        r.cost_history_capacity = len(hist)
        _check_host(host.sdfslam_host_solve(self._h, backend.encode(), C.byref(options), C.byref(r)))
        return _result_dict(r, hist)

    def arrays(self):
        """Numpy copies of the spec arrays behind the view."""
        v = self.view
        def a(p, n, dt):
            return np.ctypeslib.as_array(p, shape=(n,)).astype(dt).copy() if n > 0 else np.zeros(0, dt)
        g = v.grid
        return dict(
            node_to_active=a(v.node_to_active, g.nx * g.ny, np.int32),
            point_frame=a(v.point_frame, v.num_points, np.int32),
            point_x=a(v.point_x, v.num_points, np.float64),
            point_y=a(v.point_y, v.num_points, np.float64),
            point_expected_sdf=a(v.point_expected_sdf, v.num_points, np.float64),
            point_sqrt_weight=a(v.point_sqrt_weight, v.num_points, np.float64),
            eik_w=a(v.eik_w, v.num_eikonal, np.int32), eik_h=a(v.eik_h, v.num_eikonal, np.int32),
            eik_nx=a(v.eik_nx, v.num_eikonal, np.float64), eik_ny=a(v.eik_ny, v.num_eikonal, np.float64),
            eik_sqrt_weight=a(v.eik_sqrt_weight, v.num_eikonal, np.float64),
            odom_i=a(v.odom_i, v.num_odom, np.int32), odom_j=a(v.odom_j, v.num_odom, np.int32),
            odom_dx=a(v.odom_dx, v.num_odom, np.float64), odom_dy=a(v.odom_dy, v.num_odom, np.float64),
            odom_dtheta=a(v.odom_dtheta, v.num_odom, np.float64),
            odom_sqrt_weight=a(v.odom_sqrt_weight, v.num_odom, np.float64),
        )


def _result_dict(r, hist):
    n = min(r.cost_history_length, len(hist))
    return dict(converged=bool(r.converged), iterations=r.iterations, initial_cost=r.initial_cost,
                final_cost=r.final_cost, duration_seconds=r.duration_seconds,
                cost_history=hist[:n].copy())


def solve(problem: Problem, options, map_values, poses):
    """sdfslam_hip_solve on a view and explicit state; returns (result, map, poses)."""
    m = np.ascontiguousarray(map_values, dtype=np.float64).copy()
    p = np.ascontiguousarray(poses, dtype=np.float64).copy()
    hist = np.zeros(max(options.max_iterations, 0) + 1)
    r = SolveResult()
    r.cost_history = ptr(hist, _f64p)
    r.cost_history_capacity = len(hist)
    _check_hip(hip.sdfslam_hip_solve(C.byref(problem.view), C.byref(options), ptr(m, _f64p),
                                     ptr(p, _f64p), C.byref(r)))
    return _result_dict(r, hist), m, p


class Context:
    """A device-resident context (sdfslam_hip_context_*)."""

    def __init__(self, problem: Problem, options, map_values=None, poses=None):
        self.problem = problem
        m0, p0 = problem.state()
        m = np.ascontiguousarray(m0 if map_values is None else map_values, dtype=np.float64)
        p = np.ascontiguousarray(p0 if poses is None else poses, dtype=np.float64)
        self.options = options
        h = C.c_void_p()
        _check_hip(hip.sdfslam_hip_context_create(C.byref(problem.view), C.byref(options),
                                                  ptr(m, _f64p), ptr(p, _f64p), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            hip.sdfslam_hip_context_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def n(self):
        return self.problem.num_parameters

    def analysis(self):
        info = AnalysisInfo()
        _check_hip(hip.sdfslam_hip_analysis(self._h, C.byref(info)))
        return {k: getattr(info, k) for k, _ in AnalysisInfo._fields_}

    def set_state(self, map_values=None, poses=None):
        mv = None if map_values is None else np.ascontiguousarray(map_values, dtype=np.float64)
        pv = None if poses is None else np.ascontiguousarray(poses, dtype=np.float64)
        _check_hip(hip.sdfslam_hip_set_state(self._h, ptr(mv, _f64p), ptr(pv, _f64p)))

    @property
    def num_poses(self):
        return getattr(self, "_num_poses", None) or self.problem.view.num_poses

    def state(self):
        g = self.problem.view.grid
        m = np.zeros(g.nx * g.ny)
        p = np.zeros(3 * self.num_poses)
        _check_hip(hip.sdfslam_hip_get_state(self._h, ptr(m, _f64p), ptr(p, _f64p)))
        return m, p

    def append_frames(self, scans, odometry, options, map_values, poses):
        """Extends the context's problem by frames built on the device
        (sdfslam_hip_context_append_frames). scans: list of (k, 2) sensor-frame
        points; odometry: (frames, 3); options: FrameOptions; map_values and
        poses: the state (every pose, the new frames' last)."""
        counts = np.array([len(s) for s in scans], dtype=np.int64)
        xy = (np.ascontiguousarray(np.concatenate([np.asarray(s, float).reshape(-1, 2) for s in scans]))
              if counts.sum() > 0 else np.zeros((1, 2)))
        odo = np.ascontiguousarray(odometry, dtype=np.float64).reshape(-1, 3)
        m = np.ascontiguousarray(map_values, dtype=np.float64)
        p = np.ascontiguousarray(poses, dtype=np.float64).reshape(-1)
        _check_hip(hip.sdfslam_hip_context_append_frames(self._h, len(scans), ptr(counts, _i64p), ptr(xy, _f64p),
                                                        ptr(odo, _f64p), C.byref(options), ptr(m, _f64p),
                                                        ptr(p, _f64p)))
        self._num_poses = len(p) // 3

    def point_rows(self):
        """(frame, x, y, expected_sdf, sqrt_weight) of the rows the device holds."""
        cnt = np.zeros(1, dtype=np.int64)
        _check_hip(hip.sdfslam_hip_context_point_rows(self._h, 0, ptr(cnt, _i64p), None, None, None, None, None))
        n = int(cnt[0])
        f = np.zeros(n, dtype=np.int32)
        x, y, d, w = (np.zeros(n) for _ in range(4))
        _check_hip(hip.sdfslam_hip_context_point_rows(self._h, n, None, ptr(f, _i32p), ptr(x, _f64p),
                                                     ptr(y, _f64p), ptr(d, _f64p), ptr(w, _f64p)))
        return f, x, y, d, w

    def assemble(self):
        rows = self.problem.num_residuals
        r = np.zeros(rows)
        cols = np.zeros((rows, 7), dtype=np.int32)
        vals = np.zeros((rows, 7))
        _check_hip(hip.sdfslam_hip_assemble(self._h, ptr(r, _f64p), ptr(cols, _i32p), ptr(vals, _f64p)))
        return r, cols, vals

    def normal_equations(self, lam):
        """Damped J^T J (lower, original column order, fixed pattern), the
        structural mask of the reference pattern, and -J^T r."""
        n = self.n
        need = C.c_int64(0)
        cap = self.analysis()["pattern_nnz_lower"]
        while True:
            cp = np.zeros(n + 1, dtype=np.int32)
            ri = np.zeros(max(cap, 1), dtype=np.int32)
            va = np.zeros(max(cap, 1))
            st = np.zeros(max(cap, 1), dtype=np.uint8)
            rhs = np.zeros(n)
            status = hip.sdfslam_hip_normal_equations(
                self._h, lam, cap, C.byref(need), ptr(cp, _i32p), ptr(ri, _i32p), ptr(va, _f64p),
                ptr(st, _u8p), ptr(rhs, _f64p))
            if status == SDFSLAM_ERR_INVALID_ARGUMENT and need.value > cap:
                cap = need.value  # the pattern grew (a new epoch); retry with room for it
                continue
            _check_hip(status)
            k = need.value
            return cp, ri[:k], va[:k], st[:k].astype(bool), rhs

    def linear_step(self, lam):
        step = np.zeros(self.n)
        _check_hip(hip.sdfslam_hip_linear_step(self._h, lam, ptr(step, _f64p)))
        return step

    def cost(self):
        c = C.c_double()
        _check_hip(hip.sdfslam_hip_cost(self._h, C.byref(c)))
        return c.value

    def solve(self, options=None):
        o = options if options is not None else self.options
        hist = np.zeros(max(o.max_iterations, 0) + 1)
        r = SolveResult()
        r.cost_history = ptr(hist, _f64p)
        r.cost_history_capacity = len(hist)
        _check_hip(hip.sdfslam_hip_context_solve(self._h, C.byref(o), C.byref(r)))
        return _result_dict(r, hist)

    def reference_factor_counts(self):
        """(nonzeros of L, sum of c^2 + 3c) of the reference pattern under the
        context's order (sdfslam_hip_reference_factor_counts)."""
        nnz, flops = C.c_int64(), C.c_double()
        _check_hip(hip.sdfslam_hip_reference_factor_counts(self._h, C.byref(nnz), C.byref(flops)))
        return nnz.value, flops.value

    def elimination_order(self):
        """order[k] = the column eliminated k-th (sdfslam_hip_elimination_order)."""
        o = np.zeros(self.n, dtype=np.int32)
        _check_hip(hip.sdfslam_hip_elimination_order(self._h, ptr(o, _i32p)))
        return o

    def pose_margins(self):
        """Footprint margin (cells) of every pose in the current pattern epoch."""
        n = np.zeros(1, dtype=np.int32)
        _check_hip(hip.sdfslam_hip_pose_margins(self._h, None, 0, ptr(n, _i32p)))
        out = np.zeros(int(n[0]), dtype=np.int32)
        _check_hip(hip.sdfslam_hip_pose_margins(self._h, ptr(out, _i32p), len(out), ptr(n, _i32p)))
        return out

    def update(self, problem, map_values=None, poses=None, device_normals=None):
        """Replaces the context's problem (sdfslam_hip_context_update). With
        device_normals = (k_neighbors, method name, corner_threshold) its
        eikonal normals are computed on the device instead
        (sdfslam_hip_context_update_device_normals)."""
        m0, p0 = problem.state()
        m = np.ascontiguousarray(m0 if map_values is None else map_values, dtype=np.float64)
        p = np.ascontiguousarray(p0 if poses is None else poses, dtype=np.float64)
        if device_normals is None:
            _check_hip(hip.sdfslam_hip_context_update(self._h, C.byref(problem.view), ptr(m, _f64p),
                                                      ptr(p, _f64p)))
        else:
            # cos/sin of the frame angles from ::sincos in the library, as the
            # host problem builder places the points.
            req = NormalsRequest(int(device_normals[0]), NORMAL_METHODS[device_normals[1]],
                                 float(device_normals[2]), None)
            _check_hip(hip.sdfslam_hip_context_update_device_normals(
                self._h, C.byref(problem.view), ptr(m, _f64p), ptr(p, _f64p), C.byref(req)))
        self.problem = problem

    def iterate(self, k, phase_times=False):
        _check_hip(hip.sdfslam_hip_context_iterate(self._h, k, 1 if phase_times else 0))

    def phase_times(self):
        t = PhaseTimes()
        _check_hip(hip.sdfslam_hip_phase_times_get(self._h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in PhaseTimes._fields_}


def run_incremental(name, seed, options=None, increment_size=1, max_frames=0, solve_fn=None):
    """Incremental run of a synthetic workload (sdfslam_host_run_incremental).

    solve_fn: None solves on the device; otherwise a Python callable
    (view_ptr, options_ptr, map_ptr, poses_ptr, result_ptr) -> status used
    for every increment. Returns dict(total, records, poses, map).
    """
    full = Problem.workload(name, seed)
    frames = full.view.num_poses if max_frames <= 0 else min(max_frames, full.view.num_poses)
    g = full.view.grid
    poses = np.zeros(3 * frames)
    map_values = np.zeros(g.nx * g.ny)
    recs = (IncrementRecord * frames)()
    count = C.c_int32()
    hist = np.zeros(frames * ((options.max_iterations if options is not None else 64) + 1) + 1)
    total = SolveResult()
    total.cost_history = ptr(hist, _f64p)
    total.cost_history_capacity = len(hist)
    cb = SOLVE_FN(0)
    if solve_fn is not None:
        cb = SOLVE_FN(lambda user, v, o, m, p, r: solve_fn(v, o, m, p, r))
    _check_host(host.sdfslam_host_run_incremental(
        name.encode(), seed, C.byref(options) if options is not None else None, increment_size,
        frames, cb, None, C.byref(total), recs, frames, C.byref(count), ptr(poses, _f64p),
        ptr(map_values, _f64p)))
    records = [{k: getattr(recs[i], k) for k, _ in IncrementRecord._fields_}
               for i in range(count.value)]
    return dict(total=_result_dict(total, hist), records=records, poses=poses.reshape(-1, 3),
                map=map_values)


def run(name, seed, output_dir=None, incremental=False, increment_size=1, max_frames=0, snapshot_every=0,
        init_from_scans=False, init_signed=True, options=None, solve_fn=None):
    """A whole run of a synthetic workload with the reference's output files
    (sdfslam_host_run; Run, src/pipeline/runner.cpp:196-303). solve_fn as in
    run_incremental. Returns the accumulated SolveResult as a dict."""
    ro = RunOptions(int(incremental), increment_size, max_frames, snapshot_every, int(init_from_scans),
                    int(init_signed))
    frames = Problem.workload(name, seed).view.num_poses
    iters = options.max_iterations if options is not None else 256
    hist = np.zeros(frames * (iters + 1) + 1)
    total = SolveResult()
    total.cost_history = ptr(hist, _f64p)
    total.cost_history_capacity = len(hist)
    cb = SOLVE_FN(0)
    if solve_fn is not None:
        cb = SOLVE_FN(lambda user, v, o, m, p, r: solve_fn(v, o, m, p, r))
    _check_host(host.sdfslam_host_run(name.encode(), seed, C.byref(ro),
                                      C.byref(options) if options is not None else None, cb, None,
                                      output_dir.encode() if output_dir is not None else None,
                                      C.byref(total)))
    return _result_dict(total, hist)


def init_from_scans(grid, map_values, scans, poses, signed_distance=True, device=True):
    """Map seeding from scans (InitializeFromScans, src/core/map_init.cpp:31-69).

    scans: list of (n_i, 2) sensor-frame points; poses: (N, 3). With device
    True the host mirror places the points and the device takes the distances
    (sdfslam_host_init_from_scans); with device False the points must already
    be in world coordinates as (points (n, 2), origins (n, 2)) and the device
    entry point is called directly (sdfslam_hip_init_from_scans). Returns the
    new map values (nx * ny)."""
    out = np.ascontiguousarray(map_values, dtype=np.float64).copy()
    if device:
        ptrs = np.zeros(len(scans) + 1, dtype=np.int64)
        ptrs[1:] = np.cumsum([len(np.reshape(sc, (-1, 2))) for sc in scans])
        xy = (np.ascontiguousarray(np.concatenate([np.reshape(sc, (-1, 2)) for sc in scans]), dtype=np.float64)
              if len(scans) else np.zeros((0, 2)))
        ps = np.ascontiguousarray(poses, dtype=np.float64).reshape(-1, 3)
        _check_host(host.sdfslam_host_init_from_scans(C.byref(grid), ptr(out, _f64p), ptr(xy, _f64p),
                                                      ptr(ptrs, _i64p), len(scans), ptr(ps, _f64p),
                                                      len(ps), int(signed_distance)))
    else:
        pts, origins = scans
        p = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 2)
        o = np.ascontiguousarray(origins, dtype=np.float64).reshape(-1, 2)
        _check_hip(hip.sdfslam_hip_init_from_scans(ptr(p, _f64p), ptr(o, _f64p), len(p), C.byref(grid),
                                                   int(signed_distance), ptr(out, _f64p)))
    return out


NO_ALLOC_MAX_POINTS = 524288  # SDFSLAM_NO_ALLOC_MAX_POINTS


def init_from_scans_noalloc(grid, points_dev, origins_dev, num_points, map_dev, signed_distance=True,
                            stream=None):
    """sdfslam_hip_init_from_scans_noalloc: device addresses (ints) of the
    points and origins (x, y pairs) and of the map values, work enqueued on
    `stream` (a hipStream_t as an int, None = the null stream) with no host
    synchronisation."""
    _check_hip(hip.sdfslam_hip_init_from_scans_noalloc(C.c_void_p(points_dev), C.c_void_p(origins_dev),
                                                       num_points, C.byref(grid), int(signed_distance),
                                                       C.c_void_p(map_dev), C.c_void_p(stream)))


class Dataset:
    """Frames of a synthetic workload (sdfslam_host_dataset_*)."""

    def __init__(self, name, seed):
        h = C.c_void_p()
        _check_host(host.sdfslam_host_dataset_create(name.encode(), seed, C.byref(h)))
        self._h = h
        n = C.c_int32()
        self.grid = Grid()
        self.problem_options = ProblemOptions()
        self.step_options = SolverOptions()
        _check_host(host.sdfslam_host_dataset_info(self._h, C.byref(n), C.byref(self.grid),
                                                   C.byref(self.problem_options),
                                                   C.byref(self.step_options)))
        self.num_frames = n.value

    def __del__(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            host.sdfslam_host_dataset_destroy(self._h)
            self._h = None

    def frame(self, k):
        """(scan points (n, 2), dataset pose (3), odometry from frame k-1 (3))."""
        n = C.c_int64()
        _check_host(host.sdfslam_host_dataset_frame(self._h, k, None, 0, C.byref(n), None, None))
        xy = np.zeros(2 * n.value)
        pose = np.zeros(3)
        odo = np.zeros(3)
        _check_host(host.sdfslam_host_dataset_frame(self._h, k, ptr(xy, _f64p), n.value, C.byref(n),
                                                     ptr(pose, _f64p), ptr(odo, _f64p)))
        return xy.reshape(-1, 2), pose, odo


class Server:
    """Served form: many incremental sessions on one GPU (sdfslam_server_*)."""

    def __init__(self, workers=12, device_budget_bytes=0, max_batch=64, pipelines=3, max_in_flight=0,
                 device_normals=False):
        h = C.c_void_p()
        o = ServerOptions(workers, max_batch, pipelines, device_budget_bytes, max_in_flight,
                          int(device_normals))
        _check_host(host.sdfslam_server_create(C.byref(o), C.byref(h)))
        self._h = h
        self._grids = {}

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            host.sdfslam_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def open_session(self, grid, problem_options, step_options, first_pose, scan_xy, map_values=None):
        sid = C.c_int32()
        pose = np.ascontiguousarray(first_pose, dtype=np.float64)
        xy = np.ascontiguousarray(scan_xy, dtype=np.float64).reshape(-1)
        mv = None if map_values is None else np.ascontiguousarray(map_values, dtype=np.float64)
        _check_host(host.sdfslam_server_open_session(
            self._h, C.byref(grid), ptr(mv, _f64p), C.byref(problem_options), C.byref(step_options),
            ptr(pose, _f64p), ptr(xy, _f64p), len(xy) // 2, C.byref(sid)))
        self._grids[sid.value] = (grid.nx, grid.ny)
        return sid.value

    def submit(self, session, scan_xy, odometry):
        xy = np.ascontiguousarray(scan_xy, dtype=np.float64).reshape(-1)
        odo = np.ascontiguousarray(odometry, dtype=np.float64)
        ticket = C.c_int64()
        _check_host(host.sdfslam_server_submit(self._h, session, ptr(xy, _f64p), len(xy) // 2,
                                               ptr(odo, _f64p), C.byref(ticket)))
        return ticket.value

    def wait(self, session, ticket):
        rec = IncrementRecord()
        lat = C.c_double()
        _check_host(host.sdfslam_server_wait(self._h, session, ticket, C.byref(rec), C.byref(lat)))
        out = {k: getattr(rec, k) for k, _ in IncrementRecord._fields_}
        out["latency_seconds"] = lat.value
        return out

    def close_session(self, session):
        """Waits for the session's increments, frees its device memory and
        admits queued sessions that now fit (sdfslam_server_close_session)."""
        _check_host(host.sdfslam_server_close_session(self._h, session))

    def session_info(self, session):
        a, c = C.c_int32(), C.c_int32()
        b, sub, fin = C.c_int64(), C.c_int64(), C.c_int64()
        _check_host(host.sdfslam_server_session_info(self._h, session, C.byref(a), C.byref(c), C.byref(b),
                                                     C.byref(sub), C.byref(fin)))
        return dict(admitted=bool(a.value), closed=bool(c.value), device_bytes=b.value,
                    submitted=sub.value, finished=fin.value)

    def state(self, session):
        frames = C.c_int32()
        _check_host(host.sdfslam_server_state(self._h, session, 0, C.byref(frames), None, None))
        nx, ny = self._grids[session]
        poses = np.zeros(3 * frames.value)
        m = np.zeros(nx * ny)
        _check_host(host.sdfslam_server_state(self._h, session, frames.value, C.byref(frames),
                                              ptr(poses, _f64p), ptr(m, _f64p)))
        return poses.reshape(-1, 3), m


def set_graphs(enabled):
    """LM iterations as captured HIP graphs (default) or direct launches
    (sdfslam_hip_set_graphs), process-wide."""
    _check_hip(hip.sdfslam_hip_set_graphs(1 if enabled else 0))


def set_lds_forms(enabled):
    """The pose-block and extend-add kernels' LDS forms (default) or their
    global-memory forms, the same results (sdfslam_hip_set_lds_forms),
    process-wide, for launches made or captured after the call."""
    _check_hip(hip.sdfslam_hip_set_lds_forms(1 if enabled else 0))


def served_benchmark(sessions=32, rate_hz=10.0, duration_s=10.0, workers=12, max_batch=64,
                     pipelines=3, seed0=2000, max_in_flight=0, device_normals=False, start_frame=1):
    o = ServedOptions(sessions, rate_hz, duration_s, workers, max_batch, pipelines, seed0, max_in_flight,
                      int(device_normals), start_frame)
    r = ServedResult()
    _check_host(host.sdfslam_host_served_benchmark(C.byref(o), C.byref(r)))
    return {k: getattr(r, k) for k, _ in ServedResult._fields_}


NORMAL_METHODS = {"pca": 0, "hybrid": 1, "weighted": 2}


def node_normals(points_xy, origins_xy, k_neighbors, grid, method="pca", corner_threshold=0.2,
                 device=True):
    """Node normals and eikonal scales of a problem build from scan points and
    their sensor origins in world coordinates ((n, 2) arrays): on the device
    (sdfslam_hip_node_normals) or on the host (sdfslam_host_node_normals).
    grid: a Grid. Returns (normals (nx*ny, 2), scales (nx*ny,))."""
    p = np.ascontiguousarray(points_xy, dtype=np.float64).reshape(-1, 2)
    o = np.ascontiguousarray(origins_xy, dtype=np.float64).reshape(-1, 2)
    nodes = grid.nx * grid.ny
    nrm = np.zeros((nodes, 2))
    scale = np.zeros(nodes)
    fn = hip.sdfslam_hip_node_normals if device else host.sdfslam_host_node_normals
    st = fn(ptr(p, _f64p), ptr(o, _f64p), len(p), k_neighbors, C.byref(grid), NORMAL_METHODS[method],
            corner_threshold, ptr(nrm, _f64p), ptr(scale, _f64p))
    (_check_hip if device else _check_host)(st)
    return nrm, scale


def batch_autotune(instances=64, max_batches=(4, 8, 16), workers=(4, 8), trial_iterations=10, seed0=1000):
    """Autotuner of the batched form's shapes (sdfslam_host_batch_autotune):
    returns (best (max_batch, workers, lm iters/s), trials as such tuples)."""
    o = BatchOptions(instances, trial_iterations, 1, 1, seed0)
    mb = np.ascontiguousarray(max_batches, dtype=np.int32)
    w = np.ascontiguousarray(workers, dtype=np.int32)
    trials = (BatchShape * (len(mb) * len(w)))()
    best = BatchShape()
    _check_host(host.sdfslam_host_batch_autotune(C.byref(o), ptr(mb, _i32p), len(mb), ptr(w, _i32p), len(w),
                                                 trial_iterations, trials, C.byref(best)))
    as_tuple = lambda s: (s.max_batch, s.workers, s.lm_iterations_per_second)
    return as_tuple(best), [as_tuple(t) for t in trials]


def batch_benchmark(instances=64, iterations=30, max_batch=8, workers=8, seed0=1000):
    o = BatchOptions(instances, iterations, max_batch, workers, seed0)
    r = BatchResult()
    _check_host(host.sdfslam_host_batch_benchmark(C.byref(o), C.byref(r)))
    return {k: getattr(r, k) for k, _ in BatchResult._fields_}


def batch_solve(contexts, options, info=None):
    """sdfslam_hip_batch_solve over Context objects; returns (results,
    statuses). When `info` is a dict it receives the call's timing
    breakdown (BatchInfo fields)."""
    k = len(contexts)
    handles = (C.c_void_p * k)(*[c._h.value for c in contexts])
    hists = [np.zeros(max(options.max_iterations, 0) + 1) for _ in range(k)]
    res = (SolveResult * k)()
    for i in range(k):
        res[i].cost_history = ptr(hists[i], _f64p)
        res[i].cost_history_capacity = len(hists[i])
    st = np.zeros(k, dtype=np.int32)
    bi = BatchInfo()
    _check_hip(hip.sdfslam_hip_batch_solve(handles, k, C.byref(options), res, ptr(st, _i32p), C.byref(bi)))
    if info is not None:
        info.update({f: getattr(bi, f) for f, _ in BatchInfo._fields_})
    return [_result_dict(res[i], hists[i]) for i in range(k)], st
