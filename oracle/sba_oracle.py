"""TEST INFRASTRUCTURE -- ctypes wrapper of the CPU oracle (oracle/sba_oracle.cpp).

Only tests/, the smoke check and the benchmark's CPU-baseline leg use this module, as the checker or as the
timed CPU baseline. ``solve_f64`` is the FP64 plain method (Cholesky RCS solve, the GPU bundler's control
flow); ``solve_f32`` restates the reference CPU path (schur_complement_bundler_cpu.cpp: FP32, dense Schur
product, Jacobi pseudo-inverse of the RCS with threshold 1e-6).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle_sba.so")


class OrcOptions(C.Structure):
    _fields_ = [("loss", C.c_int32), ("rcs_solver", C.c_int32), ("dense_schur", C.c_int32),
                ("fixed_iterations", C.c_int32), ("reorthonormalize", C.c_int32), ("round_reduced_f32", C.c_int32),
                ("flip_iteration", C.c_int32), ("pad", C.c_int32)]


class OrcResult(C.Structure):
    _fields_ = [("points", C.POINTER(C.c_double)), ("poses", C.POINTER(C.c_double)), ("iterations", C.c_int32),
                ("status", C.c_int32), ("accepted", C.c_int32), ("pad", C.c_int32), ("initial_cost", C.c_double),
                ("last_cost", C.c_double), ("lambda_", C.c_double), ("trace_capacity", C.c_int32),
                ("pad2", C.c_int32), ("trace_lambda", C.POINTER(C.c_double)), ("trace_cost", C.POINTER(C.c_double)),
                ("trace_rho", C.POINTER(C.c_double)), ("trace_flags", C.POINTER(C.c_int32))]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        from cuvslam_amd._lib import ProblemView
        PV = C.POINTER(ProblemView)
        dp = C.POINTER(C.c_double)
        for name in ("orc_solve_f64", "orc_solve_f32"):
            getattr(L, name).argtypes = [PV, C.POINTER(OrcOptions), C.POINTER(OrcResult)]
            getattr(L, name).restype = C.c_int
        L.orc_linearize_f64.argtypes = [PV, C.POINTER(OrcOptions), dp, dp, dp, dp]
        L.orc_step_f64.argtypes = [PV, C.POINTER(OrcOptions), C.c_double, dp, dp, dp, dp, dp]
        L.orc_step_f64.restype = C.c_int
        L.orc_predicted_f64.argtypes = [PV, C.POINTER(OrcOptions), C.c_double, C.c_double]
        L.orc_predicted_f64.restype = C.c_double
        L.orc_cost_f64.argtypes = [PV, C.POINTER(OrcOptions), dp, dp, C.POINTER(C.c_int32)]
        L.orc_cost_f64.restype = C.c_double
        L.orc_chol_solve_f64.argtypes = [C.c_int32, dp, dp, dp]
        L.orc_sym_pinv_f64.argtypes = [C.c_int32, dp, C.c_double, dp]
        L.orc_point_pinv_f64.argtypes = [dp, dp]
        L.orc_exp_f64.argtypes = [dp, dp]
        L.orc_exp_f64.restype = None
        L.orc_nearest_rotation_f64.argtypes = [dp, dp]
        L.orc_nearest_rotation_f64.restype = None
        L.orc_implicit_grad_f64.argtypes = [PV, C.POINTER(OrcOptions), dp, dp, dp, dp, dp, dp, dp, dp, C.c_int32,
                                            C.c_double]
        L.orc_implicit_grad_f64.restype = C.c_int
        L.orc_hvp_f64.argtypes = [PV, C.POINTER(OrcOptions), dp, dp, dp, dp]
        L.orc_hvp_f64.restype = None
        L.orc_grad_f64.argtypes = [PV, C.POINTER(OrcOptions), dp, dp, dp, dp, dp]
        L.orc_grad_f64.restype = None
        L.orc_solve_state_f64.argtypes = [PV, C.POINTER(OrcOptions), dp, dp, dp, dp, C.POINTER(OrcResult)]
        L.orc_winsorize_f64.argtypes = [PV, C.c_int32, dp, dp, C.POINTER(C.c_uint8)]
        L.orc_solve_state_f64.restype = C.c_int
        L.orc_loss_f64.argtypes = [C.c_int32, C.c_double, C.c_double, C.c_int32]
        L.orc_loss_f64.restype = C.c_double
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def options(loss=0, rcs_solver=0, dense_schur=0, fixed_iterations=0, reorthonormalize=1, round_reduced_f32=0,
            flip_iteration=0):
    o = OrcOptions()
    o.loss, o.rcs_solver, o.dense_schur = loss, rcs_solver, dense_schur
    o.fixed_iterations, o.reorthonormalize = fixed_iterations, reorthonormalize
    o.round_reduced_f32, o.flip_iteration = round_reduced_f32, flip_iteration
    return o


def solve(problem, precision="f64", trace=64, **opt):
    """Runs the oracle; returns a dict with final points (N,3), poses (N,4,4) [row, col], statistics and trace."""
    L = lib()
    v = problem.view()
    o = options(**opt)
    n_p, n_c = problem.points.shape[0], problem.rig_from_world.shape[0]
    pts = np.zeros((n_p, 3))
    poses = np.zeros((n_c, 16))
    tl, tc, tr = np.zeros(trace), np.zeros(trace), np.zeros(trace)
    tf = np.zeros(trace, np.int32)
    r = OrcResult()
    r.points, r.poses = _d(pts), _d(poses)
    r.trace_capacity = trace
    r.trace_lambda, r.trace_cost, r.trace_rho = _d(tl), _d(tc), _d(tr)
    r.trace_flags = tf.ctypes.data_as(C.POINTER(C.c_int32))
    fn = L.orc_solve_f64 if precision == "f64" else L.orc_solve_f32
    rc = fn(C.byref(v), C.byref(o), C.byref(r))
    it = r.iterations
    return dict(status=rc, points=pts, poses=poses.reshape(n_c, 4, 4).transpose(0, 2, 1), iterations=it,
                accepted=r.accepted, initial_cost=r.initial_cost, last_cost=r.last_cost, lam=r.lambda_,
                trace_lambda=tl[:it], trace_cost=tc[:it], trace_rho=tr[:it], trace_flags=tf[:it])


def linearize(problem, **opt):
    L = lib()
    v = problem.view()
    n = problem.num_observations
    r, Jc, Jp, w = np.zeros((n, 2)), np.zeros((n, 2, 6)), np.zeros((n, 2, 3)), np.zeros(n)
    L.orc_linearize_f64(C.byref(v), C.byref(options(**opt)), _d(r), _d(Jc), _d(Jp), _d(w))
    return r, Jc, Jp, w


def step(problem, lam, **opt):
    """One LM step at the problem's state: (S, e, dc, dp, scalars, rc). scalars: trial cost, skipped, d^T H d,
    d^T diag(H) d, max|dp|, max|dc|, current cost, factorization ok, predicted relative reduction."""
    L = lib()
    v = problem.view()
    ncf = problem.rig_from_world.shape[0] - problem.num_fixed_key_frames
    npf = problem.points.shape[0] - problem.num_fixed_points
    n = 6 * ncf
    S, e, dc, dp, sc = np.zeros((n, n)), np.zeros(n), np.zeros(n), np.zeros(3 * npf), np.zeros(9)
    rc = L.orc_step_f64(C.byref(v), C.byref(options(**opt)), lam, _d(S), _d(e), _d(dc), _d(dp), _d(sc))
    return S, e, dc, dp, sc, rc


def predicted(problem, lam, current, **opt):
    """Predicted relative reduction of the step with damping lam at the problem's state, over `current`."""
    return lib().orc_predicted_f64(C.byref(problem.view()), C.byref(options(**opt)), lam, current)


def cost(problem, dc=None, dp=None, **opt):
    L = lib()
    v = problem.view()
    sk = C.c_int32()
    dcp = _d(np.ascontiguousarray(dc, dtype=np.float64)) if dc is not None else None
    dpp = _d(np.ascontiguousarray(dp, dtype=np.float64)) if dp is not None else None
    c = L.orc_cost_f64(C.byref(v), C.byref(options(**opt)), dcp, dpp, C.byref(sk))
    return c, sk.value


def chol_solve(A, b):
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b)
    rc = lib().orc_chol_solve_f64(A.shape[0], _d(A), _d(b), _d(x))
    return x, rc


def sym_pinv(A, thr=1e-6):
    A = np.ascontiguousarray(A, dtype=np.float64)
    out = np.zeros_like(A)
    rank = lib().orc_sym_pinv_f64(A.shape[0], _d(A), thr, _d(out))
    return out, rank


def point_pinv(A):
    A = np.ascontiguousarray(A, dtype=np.float64)
    out = np.zeros((3, 3))
    rank = lib().orc_point_pinv_f64(_d(A), _d(out))
    return out, rank


def exp(tw):
    tw = np.ascontiguousarray(tw, dtype=np.float64)
    m = np.zeros(16)
    lib().orc_exp_f64(_d(tw), _d(m))
    return m.reshape(4, 4).T.copy()


def nearest_rotation(M):
    M = np.ascontiguousarray(M, dtype=np.float64)
    R = np.zeros((3, 3))
    lib().orc_nearest_rotation_f64(_d(M), _d(R))
    return R


def loss(kind, s, d, derivative=False):
    return lib().orc_loss_f64(kind, s, d, int(derivative))


def _nx(problem):
    ncf = problem.rig_from_world.shape[0] - problem.num_fixed_key_frames
    npf = problem.points.shape[0] - problem.num_fixed_points
    return 6 * ncf + 3 * npf


def _state(points, poses):
    """FP64 state overrides: points (N,3), poses (N,4,4) [row, col] -> C arrays (column-major poses)."""
    if points is None:
        return None, None, []
    P = np.ascontiguousarray(points, dtype=np.float64)
    Q = np.ascontiguousarray(np.asarray(poses, dtype=np.float64).transpose(0, 2, 1))
    return _d(P), _d(Q), [P, Q]


def _obs(xy, info):
    """FP64 observation overrides: xy (N,2), info (N,2,2) [row, col] -> C arrays (column-major info)."""
    keep = [None, None]
    if xy is not None:
        keep[0] = np.ascontiguousarray(xy, dtype=np.float64)
    if info is not None:
        keep[1] = np.ascontiguousarray(np.asarray(info, dtype=np.float64).transpose(0, 2, 1))
    return (_d(keep[0]) if xy is not None else None), (_d(keep[1]) if info is not None else None), keep


def grad(problem, points=None, poses=None, xy=None, info=None, **opt):
    """Gradient of the cost in tangent coordinates (left twists of free key frames, then free points)."""
    g = np.zeros(_nx(problem))
    sp, sq, keep = _state(points, poses)
    ox, oi, keep2 = _obs(xy, info)
    lib().orc_grad_f64(C.byref(problem.view()), C.byref(options(**opt)), sp, sq, ox, oi, _d(g))
    return g


def hvp(problem, x, points=None, poses=None, **opt):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros_like(x)
    sp, sq, keep = _state(points, poses)
    lib().orc_hvp_f64(C.byref(problem.view()), C.byref(options(**opt)), sp, sq, _d(x), _d(y))
    return y


def implicit_grad(problem, g, points=None, poses=None, xy=None, info=None, max_it=200, tol=1e-12, **opt):
    """(a, dL/dxy (N_o, 2), dL/dinfo (N_o, 3) [L00, L01 shared, L11], cg iterations) at the given state."""
    g = np.ascontiguousarray(g, dtype=np.float64)
    a = np.zeros_like(g)
    n = problem.num_observations
    dxy, dinfo = np.zeros((n, 2)), np.zeros((n, 3))
    sp, sq, keep = _state(points, poses)
    ox, oi, keep2 = _obs(xy, info)
    it = lib().orc_implicit_grad_f64(C.byref(problem.view()), C.byref(options(**opt)), sp, sq, ox, oi, _d(g), _d(a),
                                     _d(dxy), _d(dinfo), max_it, tol)
    return a, dxy, dinfo, it


def solve_state(problem, points, poses, xy=None, info=None, trace=0, **opt):
    """The FP64 solve from an FP64 initial state; returns final points and poses in FP64."""
    L = lib()
    v = problem.view()
    o = options(**opt)
    n_p, n_c = problem.points.shape[0], problem.rig_from_world.shape[0]
    pts = np.zeros((n_p, 3))
    pq = np.zeros((n_c, 16))
    r = OrcResult()
    r.points, r.poses = _d(pts), _d(pq)
    r.trace_capacity = 0
    sp, sq, keep = _state(points, poses)
    ox, oi, keep2 = _obs(xy, info)
    rc = L.orc_solve_state_f64(C.byref(v), C.byref(o), sp, sq, ox, oi, C.byref(r))
    return dict(status=rc, points=pts, poses=pq.reshape(n_c, 4, 4).transpose(0, 2, 1), iterations=r.iterations,
                last_cost=r.last_cost)


def winsorize(problem, key_frame, points=None, poses=None):
    """Landmark winsorizing at a state (points (N,3), poses (N,4,4); the problem's own when None): (reset mask
    (N_p,) bool, number reset)."""
    v = problem.view()
    sp, sq, keep = _state(points, poses)
    mask = np.zeros(problem.points.shape[0], np.uint8)
    n = lib().orc_winsorize_f64(C.byref(v), key_frame, sp, sq, mask.ctypes.data_as(C.POINTER(C.c_uint8)))
    return mask.astype(bool), n


# ---- the mono SBA solver (libs/sba/mono_sba_solver.cpp) -------------------------------------------------------------
class OrcMonoProblem(C.Structure):
    _fields_ = [("num_frames", C.c_int32), ("num_tracks", C.c_int32), ("num_obs", C.c_int32),
                ("constrained", C.POINTER(C.c_int32)), ("frame_poses", C.POINTER(C.c_float)),
                ("points", C.POINTER(C.c_float)), ("obs_frame", C.POINTER(C.c_int32)),
                ("obs_track", C.POINTER(C.c_int32)), ("obs_xy", C.POINTER(C.c_float)), ("max_iterations", C.c_int32),
                ("min_residual", C.c_float), ("min_residual_speed", C.c_float)]


class OrcMonoResult(C.Structure):
    _fields_ = [("frame_poses", C.POINTER(C.c_double)), ("points", C.POINTER(C.c_double)),
                ("covariance", C.c_double * 36), ("residual", C.c_double), ("residual_initial", C.c_double),
                ("lambda_", C.c_double), ("iterations", C.c_int32), ("failed", C.c_int32), ("feasible", C.c_int32),
                ("pad", C.c_int32)]


def mono_solve(window, precision="f64"):
    """The mono solver on a window dict (constrained (F,), frame_poses (F,4,4) world_from_camera, points (T,3),
    obs_frame, obs_track (O,), obs_xy (O,2), max_iterations, min_residual, min_residual_speed)."""
    L = lib()
    for name in ("orc_mono_solve_f64", "orc_mono_solve_f32"):
        getattr(L, name).argtypes = [C.POINTER(OrcMonoProblem), C.POINTER(OrcMonoResult)]
        getattr(L, name).restype = C.c_int
    con = np.ascontiguousarray(window["constrained"], np.int32)
    poses = np.ascontiguousarray(np.asarray(window["frame_poses"], np.float32).transpose(0, 2, 1))
    pts = np.ascontiguousarray(window["points"], np.float32)
    of = np.ascontiguousarray(window["obs_frame"], np.int32)
    ot = np.ascontiguousarray(window["obs_track"], np.int32)
    xy = np.ascontiguousarray(window["obs_xy"], np.float32)
    p = OrcMonoProblem(len(con), len(pts), len(of), con.ctypes.data_as(C.POINTER(C.c_int32)),
                       poses.ctypes.data_as(C.POINTER(C.c_float)), pts.ctypes.data_as(C.POINTER(C.c_float)),
                       of.ctypes.data_as(C.POINTER(C.c_int32)), ot.ctypes.data_as(C.POINTER(C.c_int32)),
                       xy.ctypes.data_as(C.POINTER(C.c_float)), int(window.get("max_iterations", 1000)),
                       float(window.get("min_residual", 0.001)), float(window.get("min_residual_speed", 0.001)))
    out_poses = np.zeros((len(con), 16))
    out_pts = np.zeros((len(pts), 3))
    r = OrcMonoResult()
    r.frame_poses, r.points = _d(out_poses), _d(out_pts)
    fn = L.orc_mono_solve_f64 if precision == "f64" else L.orc_mono_solve_f32
    rc = fn(C.byref(p), C.byref(r))
    return dict(status=rc, frame_poses=out_poses.reshape(-1, 4, 4).transpose(0, 2, 1), points=out_pts,
                covariance=np.array(r.covariance).reshape(6, 6), residual=r.residual,
                residual_initial=r.residual_initial, lam=r.lambda_, iterations=r.iterations, failed=bool(r.failed),
                feasible=bool(r.feasible))


def mono_derivs(a, b):
    """(A 2x6, B 2x3): the mono solver's projection derivatives of world point b through camera twist a."""
    L = lib()
    L.orc_mono_derivs_f64.argtypes = [C.POINTER(C.c_double)] * 4
    L.orc_mono_derivs_f64.restype = None
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    A, B = np.zeros(12), np.zeros(6)
    L.orc_mono_derivs_f64(_d(a), _d(b), _d(A), _d(B))
    return A.reshape(2, 6), B.reshape(2, 3)


def twist_log(M):
    L = lib()
    L.orc_twist_log_f64.argtypes = [C.POINTER(C.c_double)] * 2
    L.orc_twist_log_f64.restype = None
    m = np.ascontiguousarray(np.asarray(M, np.float64).T)
    tw = np.zeros(6)
    L.orc_twist_log_f64(_d(m), _d(tw))
    return tw


def twist_add(a, b):
    L = lib()
    L.orc_twist_add_f64.argtypes = [C.POINTER(C.c_double)] * 3
    L.orc_twist_add_f64.restype = None
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    out = np.zeros(6)
    L.orc_twist_add_f64(_d(a), _d(b), _d(out))
    return out
