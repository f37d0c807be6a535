"""ctypes access to the CPU oracle (oracle/liboracle.so) - test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load the
oracle; the product (sdf-slam_amd/) never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np
import scipy.sparse as sp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB = os.path.join(ORACLE_DIR, "liboracle.so")


def _load():
    if not os.path.exists(LIB):
        subprocess.run(["make", "-C", ORACLE_DIR], check=True, capture_output=True)
    return C.CDLL(LIB)


lib = _load()

_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_f64p = C.POINTER(C.c_double)
_vp = C.c_void_p


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def _setup(view_type, grid_type, solver_options_type, solve_result_type):
    V = C.POINTER(view_type)
    G = C.POINTER(grid_type)
    lib.oracle_wrap_angle.restype = C.c_double
    lib.oracle_wrap_angle.argtypes = [C.c_double]
    lib.oracle_compose.argtypes = [_f64p, _f64p, _f64p]
    lib.oracle_relative.argtypes = [_f64p, _f64p, _f64p]
    lib.oracle_locate.argtypes = [G, C.c_double, C.c_double, C.c_int, _i32p, _i32p, _f64p, _f64p]
    lib.oracle_interpolate.restype = C.c_double
    lib.oracle_interpolate.argtypes = [G, _f64p, C.c_double, C.c_double, C.c_int]
    lib.oracle_eval_point.argtypes = [G, _f64p, _f64p, C.c_double, C.c_double, C.c_double, C.c_int,
                                      C.c_double, _f64p, _i32p, _f64p, _f64p]
    lib.oracle_eval_eikonal.argtypes = [G, _f64p, C.c_int32, C.c_int32, C.c_double, C.c_double,
                                        _f64p, _i32p, _f64p]
    lib.oracle_eval_odom.argtypes = [_f64p, _f64p, _f64p, _f64p, _f64p, _f64p]
    lib.oracle_matrix_from_csc.restype = _vp
    lib.oracle_matrix_from_csc.argtypes = [C.c_int64, C.c_int64, _i64p, _i32p, _f64p]
    lib.oracle_matrix_free.argtypes = [_vp]
    lib.oracle_matrix_shape.argtypes = [_vp, _i64p, _i64p, _i64p]
    lib.oracle_matrix_copy.argtypes = [_vp, _i64p, _i32p, _f64p]
    lib.oracle_assemble.argtypes = [V, _f64p, _f64p, _f64p, C.POINTER(_vp)]
    lib.oracle_normal_matrix.restype = _vp
    lib.oracle_normal_matrix.argtypes = [_vp]
    lib.oracle_damp.argtypes = [_vp, C.c_double, C.c_int]
    lib.oracle_neg_atb.argtypes = [_vp, _f64p, _f64p]
    lib.oracle_order.argtypes = [V, _i32p]
    lib.oracle_ldlt_solve.argtypes = [_vp, _i32p, _f64p, _f64p, _i64p]
    lib.oracle_cost.restype = C.c_double
    lib.oracle_cost.argtypes = [V, _f64p, _f64p]
    lib.oracle_apply_step.argtypes = [V, _f64p, _f64p, _f64p]
    lib.oracle_linear_step.argtypes = [V, _f64p, _f64p, C.c_double, C.c_int, _f64p]
    lib.oracle_solve.argtypes = [V, C.POINTER(solver_options_type), _f64p, _f64p,
                                 C.POINTER(solve_result_type)]
    lib.oracle_solve_recorded.argtypes = [V, C.POINTER(solver_options_type), _f64p, _f64p,
                                          C.POINTER(solve_result_type), C.c_int32, _i32p, _f64p, _f64p]
    lib.oracle_last_phase_seconds.argtypes = [_f64p]
    lib.oracle_set_threads.argtypes = [C.c_int]
    lib.oracle_threads.restype = C.c_int
    lib.oracle_last_factor_stats.argtypes = [_i64p, _f64p]
    lib.oracle_amd_order.argtypes = [_vp, _i32p]
    lib.oracle_eval_points_frozen.argtypes = [G, _f64p, _f64p, C.c_int64, _i32p, _f64p, _f64p, _f64p,
                                              _i32p, _i32p, _f64p, _f64p, _f64p]
    lib.oracle_init_from_scans.argtypes = [G, _f64p, _f64p, _i64p, C.c_int32, _f64p, C.c_int32, C.c_int]


_configured = False


def configure(mod):
    """Binds argument types using the struct definitions of the sdfslam module."""
    global _configured
    if not _configured:
        _setup(mod.ProblemView, mod.Grid, mod.SolverOptions, mod.SolveResult)
        _configured = True


class OracleMatrix:
    def __init__(self, handle):
        self.h = _vp(handle)

    def __del__(self):
        if self.h:
            lib.oracle_matrix_free(self.h)
            self.h = None

    def to_scipy(self):
        rows, cols, nnz = C.c_int64(), C.c_int64(), C.c_int64()
        lib.oracle_matrix_shape(self.h, C.byref(rows), C.byref(cols), C.byref(nnz))
        cp = np.zeros(cols.value + 1, dtype=np.int64)
        ri = np.zeros(max(nnz.value, 1), dtype=np.int32)
        va = np.zeros(max(nnz.value, 1))
        lib.oracle_matrix_copy(self.h, _p(cp, _i64p), _p(ri, _i32p), _p(va, _f64p))
        return sp.csc_matrix((va[:nnz.value], ri[:nnz.value], cp), shape=(rows.value, cols.value))


def assemble(view, map_values, poses):
    rows = view.num_points + view.num_eikonal + 3 * view.num_odom
    r = np.zeros(rows)
    h = _vp()
    m = np.ascontiguousarray(map_values, dtype=np.float64)
    p = np.ascontiguousarray(poses, dtype=np.float64)
    lib.oracle_assemble(C.byref(view), _p(m, _f64p), _p(p, _f64p), _p(r, _f64p), C.byref(h))
    return r, OracleMatrix(h.value)


def from_csc(a):
    """An oracle matrix holding a copy of a scipy matrix (sorted CSC)."""
    a = sp.csc_matrix(a)
    a.sort_indices()
    cp = np.ascontiguousarray(a.indptr, dtype=np.int64)
    ri = np.ascontiguousarray(a.indices, dtype=np.int32)
    va = np.ascontiguousarray(a.data, dtype=np.float64)
    return OracleMatrix(lib.oracle_matrix_from_csc(a.shape[0], a.shape[1], _p(cp, _i64p),
                                                   _p(ri, _i32p), _p(va, _f64p)))


def normal_matrix(jac: OracleMatrix, lam=None, marquardt=False):
    a = OracleMatrix(lib.oracle_normal_matrix(jac.h))
    if lam is not None:
        lib.oracle_damp(a.h, lam, 1 if marquardt else 0)
    return a


def neg_atb(jac: OracleMatrix, r, n):
    rhs = np.zeros(n)
    rr = np.ascontiguousarray(r, dtype=np.float64)
    lib.oracle_neg_atb(jac.h, _p(rr, _f64p), _p(rhs, _f64p))
    return rhs


def order(view, n):
    o = np.zeros(n, dtype=np.int32)
    lib.oracle_order(C.byref(view), _p(o, _i32p))
    return o


def ldlt_solve(a: OracleMatrix, b, order_=None):
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b)
    lnz = C.c_int64()
    st = lib.oracle_ldlt_solve(a.h, _p(order_, _i32p), _p(b, _f64p), _p(x, _f64p), C.byref(lnz))
    return st, x


def cost(view, map_values, poses):
    m = np.ascontiguousarray(map_values, dtype=np.float64)
    p = np.ascontiguousarray(poses, dtype=np.float64)
    return lib.oracle_cost(C.byref(view), _p(m, _f64p), _p(p, _f64p))


def apply_step(view, step, map_values, poses):
    """Problem::ApplyStep restated; returns the new (map, poses)."""
    m = np.ascontiguousarray(map_values, dtype=np.float64).copy()
    p = np.ascontiguousarray(poses, dtype=np.float64).copy()
    st = np.ascontiguousarray(step, dtype=np.float64)
    lib.oracle_apply_step(C.byref(view), _p(st, _f64p), _p(m, _f64p), _p(p, _f64p))
    return m, p


def linear_step(view, map_values, poses, lam, marquardt=False):
    n = view.num_active_nodes + 3 * (view.num_poses - 1)
    step = np.zeros(n)
    m = np.ascontiguousarray(map_values, dtype=np.float64)
    p = np.ascontiguousarray(poses, dtype=np.float64)
    st = lib.oracle_linear_step(C.byref(view), _p(m, _f64p), _p(p, _f64p), lam,
                                1 if marquardt else 0, _p(step, _f64p))
    return st, step


def eval_points_frozen(grid, values, poses, frame, px, py, delta, cell_w, cell_h):
    """EvalPointResidualFrozenCell over arrays of rows; returns (residual (n,),
    d_nodes (n, 4), d_pose (n, 3))."""
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    frame, cell_w, cell_h = i32(frame), i32(cell_w), i32(cell_h)
    px, py, delta, v, ps = f64(px), f64(py), f64(delta), f64(values), f64(poses)
    n = len(px)
    r, dn, dp = np.zeros(n), np.zeros((n, 4)), np.zeros((n, 3))
    lib.oracle_eval_points_frozen(C.byref(grid), _p(v, _f64p), _p(ps, _f64p), n, _p(frame, _i32p),
                                  _p(px, _f64p), _p(py, _f64p), _p(delta, _f64p), _p(cell_w, _i32p),
                                  _p(cell_h, _i32p), _p(r, _f64p), _p(dn, _f64p), _p(dp, _f64p))
    return r, dn, dp


def init_from_scans(grid, map_values, scans, poses, signed_distance=True):
    """InitializeFromScans restated (exhaustive nearest point); returns the new map values."""
    out = np.ascontiguousarray(map_values, dtype=np.float64).copy()
    ptrs = np.zeros(len(scans) + 1, dtype=np.int64)
    ptrs[1:] = np.cumsum([len(np.reshape(sc, (-1, 2))) for sc in scans])
    xy = (np.ascontiguousarray(np.concatenate([np.reshape(sc, (-1, 2)) for sc in scans]), dtype=np.float64)
          if len(scans) else np.zeros((0, 2)))
    ps = np.ascontiguousarray(poses, dtype=np.float64).reshape(-1, 3)
    lib.oracle_init_from_scans(C.byref(grid), _p(out, _f64p), _p(xy, _f64p), _p(ptrs, _i64p), len(scans),
                               _p(ps, _f64p), len(ps), int(signed_distance))
    return out


def solve(mod, view, options, map_values, poses):
    """SolveLm restated; returns (status, result dict, map, poses)."""
    m = np.ascontiguousarray(map_values, dtype=np.float64).copy()
    p = np.ascontiguousarray(poses, dtype=np.float64).copy()
    hist = np.zeros(max(options.max_iterations, 0) + 1)
    r = mod.SolveResult()
    r.cost_history = _p(hist, _f64p)
    r.cost_history_capacity = len(hist)
    st = lib.oracle_solve(C.byref(view), C.byref(options), _p(m, _f64p), _p(p, _f64p), C.byref(r))
    n = min(r.cost_history_length, len(hist))
    res = dict(converged=bool(r.converged), iterations=r.iterations, initial_cost=r.initial_cost,
               final_cost=r.final_cost, duration_seconds=r.duration_seconds,
               cost_history=hist[:n].copy())
    return st, res, m, p


def solve_recorded(mod, view, options, map_values, poses, record_iterations):
    """solve() that also returns the state after each of record_iterations:
    (status, result, map, poses, maps (k, nx*ny), poses (k, 3*N))."""
    m = np.ascontiguousarray(map_values, dtype=np.float64).copy()
    p = np.ascontiguousarray(poses, dtype=np.float64).copy()
    its = np.ascontiguousarray(record_iterations, dtype=np.int32)
    rm = np.full((len(its), len(m)), np.nan)
    rp = np.full((len(its), len(p)), np.nan)
    hist = np.zeros(max(options.max_iterations, 0) + 1)
    r = mod.SolveResult()
    r.cost_history = _p(hist, _f64p)
    r.cost_history_capacity = len(hist)
    st = lib.oracle_solve_recorded(C.byref(view), C.byref(options), _p(m, _f64p), _p(p, _f64p),
                                   C.byref(r), len(its), _p(its, _i32p), _p(rm, _f64p), _p(rp, _f64p))
    n = min(r.cost_history_length, len(hist))
    res = dict(converged=bool(r.converged), iterations=r.iterations, initial_cost=r.initial_cost,
               final_cost=r.final_cost, duration_seconds=r.duration_seconds,
               cost_history=hist[:n].copy())
    return st, res, m, p, rm, rp


def set_threads(n):
    lib.oracle_set_threads(int(n))


def threads():
    return lib.oracle_threads()


def last_factor_stats():
    """(nonzeros of L below the diagonal, sum of c^2 + 3c) of the last LDL^T."""
    nnz, flops = C.c_int64(), C.c_double()
    lib.oracle_last_factor_stats(C.byref(nnz), C.byref(flops))
    return nnz.value, flops.value


def amd_order(a: OracleMatrix, n):
    o = np.zeros(n, dtype=np.int32)
    lib.oracle_amd_order(a.h, _p(o, _i32p))
    return o


def last_phase_seconds():
    out = np.zeros(4)
    lib.oracle_last_phase_seconds(_p(out, _f64p))
    return out


def eval_point(grid, values, pose, px, py, delta, smooth=0, smooth_step=0.0):
    r = C.c_double()
    ids = np.zeros(4, dtype=np.int32)
    dn = np.zeros(4)
    dp = np.zeros(3)
    v = np.ascontiguousarray(values, dtype=np.float64)
    pz = np.ascontiguousarray(pose, dtype=np.float64)
    valid = lib.oracle_eval_point(C.byref(grid), _p(v, _f64p), _p(pz, _f64p), px, py, delta,
                                  smooth, smooth_step, C.byref(r), _p(ids, _i32p), _p(dn, _f64p),
                                  _p(dp, _f64p))
    return bool(valid), r.value, ids, dn, dp


def eval_eikonal(grid, values, w, h, nx, ny):
    r = C.c_double()
    ids = np.zeros(3, dtype=np.int32)
    d = np.zeros(3)
    v = np.ascontiguousarray(values, dtype=np.float64)
    lib.oracle_eval_eikonal(C.byref(grid), _p(v, _f64p), w, h, nx, ny, C.byref(r), _p(ids, _i32p),
                            _p(d, _f64p))
    return r.value, ids, d


def eval_odom(pi, pj, meas):
    r = np.zeros(3)
    ji = np.zeros(9)
    jj = np.zeros(9)
    a = np.ascontiguousarray(pi, dtype=np.float64)
    b = np.ascontiguousarray(pj, dtype=np.float64)
    m = np.ascontiguousarray(meas, dtype=np.float64)
    lib.oracle_eval_odom(_p(a, _f64p), _p(b, _f64p), _p(m, _f64p), _p(r, _f64p), _p(ji, _f64p),
                         _p(jj, _f64p))
    return r, ji.reshape(3, 3), jj.reshape(3, 3)


def locate(grid, x, y, clamped=0):
    w, h = C.c_int32(), C.c_int32()
    a, b = C.c_double(), C.c_double()
    ok = lib.oracle_locate(C.byref(grid), x, y, clamped, C.byref(w), C.byref(h), C.byref(a),
                           C.byref(b))
    return bool(ok), w.value, h.value, a.value, b.value


def interpolate(grid, values, x, y, clamped=0):
    v = np.ascontiguousarray(values, dtype=np.float64)
    return lib.oracle_interpolate(C.byref(grid), _p(v, _f64p), x, y, clamped)


def wrap(a):
    return lib.oracle_wrap_angle(a)


def compose(a, rel):
    out = np.zeros(3)
    aa = np.ascontiguousarray(a, dtype=np.float64)
    rr = np.ascontiguousarray(rel, dtype=np.float64)
    lib.oracle_compose(_p(aa, _f64p), _p(rr, _f64p), _p(out, _f64p))
    return out


def relative(a, b):
    out = np.zeros(3)
    aa = np.ascontiguousarray(a, dtype=np.float64)
    bb = np.ascontiguousarray(b, dtype=np.float64)
    lib.oracle_relative(_p(aa, _f64p), _p(bb, _f64p), _p(out, _f64p))
    return out
