"""TEST INFRASTRUCTURE -- ctypes wrapper of the inertial-SBA oracle (oracle/imu_oracle.cpp).

Only tests/ and the smoke check use this module, as the checker. ``solve`` is the FP64 restatement of the GPU
inertial bundler's solve (imu_sba_gpu_fixed_vel.cpp:75-363); ``integrate`` / ``integrate_f64`` the IMU
preintegration (imu_preintegration.cpp); ``model`` / ``system`` / ``cost`` expose its inner functions for the pins.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from cuvslam_amd import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle_imu.so")

_dp = C.POINTER(C.c_double)


class OrcImuResult(C.Structure):
    _fields_ = [("points", _dp), ("poses", _dp), ("vel_bias", _dp),
                ("preintegration", C.POINTER(_lib.ImuPreintegration)), ("status", C.c_int32),
                ("iterations", C.c_int32), ("accepted", C.c_int32), ("pad", C.c_int32),
                ("initial_cost", C.c_double), ("last_cost", C.c_double), ("lambda_", C.c_double),
                ("trace_capacity", C.c_int32), ("pad2", C.c_int32), ("trace_lambda", _dp), ("trace_cost", _dp),
                ("trace_rho", _dp), ("trace_flags", C.POINTER(C.c_int32))]


_lib_handle = None


def lib():
    global _lib_handle
    if _lib_handle is None:
        if not os.path.exists(LIB):
            subprocess.run(["make", "-s", "-C", HERE, "liboracle_imu.so"], check=True)
        L = C.CDLL(LIB)
        cal, pv = C.POINTER(_lib.ImuCalibration), C.POINTER(_lib.ImuProblemView)
        L.orc_imu_solve.argtypes = [cal, pv, C.POINTER(OrcImuResult)]
        L.orc_imu_solve.restype = C.c_int
        L.orc_imu_integrate.argtypes = [cal, C.POINTER(_lib.ImuMeasurement), C.c_int32, C.POINTER(C.c_float),
                                        C.POINTER(C.c_float), C.POINTER(_lib.ImuPreintegration)]
        L.orc_imu_integrate_f64.argtypes = [cal, C.POINTER(_lib.ImuMeasurement), C.c_int32, _dp, _dp, _dp]
        L.orc_imu_information.argtypes = [C.POINTER(_lib.ImuPreintegration), _dp, _dp, _dp]
        L.orc_imu_cost.argtypes = [cal, pv, _dp, _dp, _dp]
        L.orc_imu_cost.restype = C.c_double
        L.orc_imu_model.argtypes = [cal, pv, _dp, _dp, _dp, _dp, _dp]
        L.orc_imu_system.argtypes = [cal, pv, C.c_float, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_imu_system.restype = C.c_int
        L.orc_so3.argtypes = [C.c_int, _dp, _dp]
        _lib_handle = L
    return _lib_handle


def _d(a):
    return a.ctypes.data_as(_dp)


def integrate(calib, samples, gyro_bias=(0, 0, 0), acc_bias=(0, 0, 0)):
    """One preintegration record (the FP64 integration stored in FP32, as the ABI holds it)."""
    from cuvslam_amd.imu import MEAS_DTYPE, PREINT_DTYPE
    s = np.ascontiguousarray(samples, MEAS_DTYPE)
    out = np.zeros(1, PREINT_DTYPE)
    bg, ba = np.asarray(gyro_bias, np.float32), np.asarray(acc_bias, np.float32)
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
    lib().orc_imu_integrate(C.byref(calib.view()), s.ctypes.data_as(C.POINTER(_lib.ImuMeasurement)), s.shape[0],
                            fp(bg), fp(ba), out.ctypes.data_as(C.POINTER(_lib.ImuPreintegration)))
    return out[0]


def integrate_f64(calib, samples, gyro_bias=(0, 0, 0), acc_bias=(0, 0, 0)):
    """dict of the FP64 integration: dR, dV, dP, JRg, JVg, JVa, JPg, JPa, cov (9x9), cov_grw, cov_arw, dT."""
    from cuvslam_amd.imu import MEAS_DTYPE
    s = np.ascontiguousarray(samples, MEAS_DTYPE)
    out = np.zeros(160)
    lib().orc_imu_integrate_f64(C.byref(calib.view()), s.ctypes.data_as(C.POINTER(_lib.ImuMeasurement)),
                                s.shape[0], _d(np.asarray(gyro_bias, np.float64).copy()),
                                _d(np.asarray(acc_bias, np.float64).copy()), _d(out))
    r, o = {}, 0
    for name, shape in (("dR", (3, 3)), ("dV", (3,)), ("dP", (3,)), ("JRg", (3, 3)), ("JVg", (3, 3)),
                        ("JVa", (3, 3)), ("JPg", (3, 3)), ("JPa", (3, 3)), ("cov", (9, 9)), ("cov_grw", (3, 3)),
                        ("cov_arw", (3, 3)), ("dT", ())):
        n = int(np.prod(shape)) if shape else 1
        r[name] = out[o:o + n].reshape(shape) if shape else float(out[o])
        o += n
    return r


def information(pre):
    from cuvslam_amd.imu import PREINT_DTYPE
    rec = np.ascontiguousarray(np.asarray(pre, PREINT_DTYPE).reshape(1))
    i9, ig, ia = np.zeros(81), np.zeros(9), np.zeros(9)
    lib().orc_imu_information(rec.ctypes.data_as(C.POINTER(_lib.ImuPreintegration)), _d(i9), _d(ig), _d(ia))
    return i9.reshape(9, 9), ig.reshape(3, 3), ia.reshape(3, 3)


def solve(problem, calib, trace=64):
    """The FP64 solve: dict with points (P,3), w_from_imu (K,4,4), velocity, gyro_bias, acc_bias (K,3),
    preintegration (K records), status, iterations, accepted, initial_cost, last_cost, lam and the LM trace."""
    from cuvslam_amd.imu import PREINT_DTYPE
    v = problem.view()
    K, P = problem.w_from_imu.shape[0], problem.points.shape[0]
    pts, poses, vb = np.zeros((P, 3)), np.zeros((K, 12)), np.zeros((K, 9))
    pre = np.zeros(K, PREINT_DTYPE)
    tl, tc, tr_, tf = np.zeros(trace), np.zeros(trace), np.zeros(trace), np.zeros(trace, np.int32)
    r = OrcImuResult()
    r.points, r.poses, r.vel_bias = _d(pts), _d(poses), _d(vb)
    r.preintegration = pre.ctypes.data_as(C.POINTER(_lib.ImuPreintegration))
    r.trace_capacity = trace
    r.trace_lambda, r.trace_cost, r.trace_rho = _d(tl), _d(tc), _d(tr_)
    r.trace_flags = tf.ctypes.data_as(C.POINTER(C.c_int32))
    lib().orc_imu_solve(C.byref(calib.view()), C.byref(v), C.byref(r))
    T = np.zeros((K, 4, 4))
    T[:, :3, :3] = poses[:, :9].reshape(K, 3, 3)
    T[:, :3, 3] = poses[:, 9:]
    T[:, 3, 3] = 1
    n = min(r.iterations, trace)
    return {"points": pts, "w_from_imu": T, "velocity": vb[:, :3], "gyro_bias": vb[:, 3:6], "acc_bias": vb[:, 6:],
            "preintegration": pre, "status": r.status, "iterations": r.iterations, "accepted": r.accepted,
            "initial_cost": r.initial_cost, "last_cost": r.last_cost, "lam": r.lambda_,
            "trace_lambda": tl[:n], "trace_cost": tc[:n], "trace_rho": tr_[:n], "trace_flags": tf[:n]}


def _state_args(points=None, w_from_imu=None, velocity=None, gyro_bias=None, acc_bias=None, K=0):
    pts = None if points is None else np.ascontiguousarray(points, np.float64)
    poses = None
    if w_from_imu is not None:
        T = np.asarray(w_from_imu, np.float64).reshape(-1, 4, 4)
        poses = np.ascontiguousarray(np.concatenate([T[:, :3, :3].reshape(-1, 9), T[:, :3, 3]], axis=1))
    vb = None
    if velocity is not None:
        vb = np.ascontiguousarray(np.concatenate([np.asarray(velocity, np.float64).reshape(-1, 3),
                                                  np.asarray(gyro_bias, np.float64).reshape(-1, 3),
                                                  np.asarray(acc_bias, np.float64).reshape(-1, 3)], axis=1))
    return pts, poses, vb


def cost(problem, calib, points=None, w_from_imu=None, velocity=None, gyro_bias=None, acc_bias=None):
    """The FP64 cost of a state (the problem's where an argument is None; velocity and biases go together)."""
    pts, poses, vb = _state_args(points, w_from_imu, velocity, gyro_bias, acc_bias)
    v = problem.view()
    n = lambda a: None if a is None else _d(a)  # noqa: E731
    return lib().orc_imu_cost(C.byref(calib.view()), C.byref(v), n(pts), n(poses), n(vb))


def model(problem, calib):
    """(edge residuals (K-1, 15): r (9), r_gyro_rw (3), r_acc_rw (3); edge Jacobians (K-1, 8, 9, 3); observation
    residuals (N, 2), weights (N,), Jacobians (N, 3, 2, 3): JR, Jt, Jp)."""
    K, N = problem.w_from_imu.shape[0], problem.num_observations
    er, ej = np.zeros((max(K - 1, 0), 15)), np.zeros((max(K - 1, 0), 8, 9, 3))
    orr, ow, oj = np.zeros((N, 2)), np.zeros(N), np.zeros((N, 3, 2, 3))
    v = problem.view()
    lib().orc_imu_model(C.byref(calib.view()), C.byref(v), _d(er), _d(ej), _d(orr), _d(ow), _d(oj))
    return er, ej, orr, ow, oj


def system(problem, calib, lam):
    """dict of the full system (U, ea, V (P,3,3), eb, W (3P, 6n)) at the problem's state, the reduced system
    (A, rhs) at damping lam, the step (dc, dp) and whether the factorization succeeded."""
    K, P, F = problem.w_from_imu.shape[0], problem.points.shape[0], problem.num_fixed_key_frames
    n = K - F
    U, ea = np.zeros((15 * n, 15 * n)), np.zeros(15 * n)
    V, eb, W = np.zeros((P, 3, 3)), np.zeros(3 * P), np.zeros((3 * P, 6 * n))
    A, rhs, dc, dp = np.zeros((15 * n, 15 * n)), np.zeros(15 * n), np.zeros(15 * n), np.zeros(3 * P)
    v = problem.view()
    ok = lib().orc_imu_system(C.byref(calib.view()), C.byref(v), float(lam), _d(U), _d(ea), _d(V), _d(eb), _d(W),
                              _d(A), _d(rhs), _d(dc), _d(dp))
    return {"U": U, "ea": ea, "V": V, "eb": eb, "W": W, "A": A, "rhs": rhs, "dc": dc, "dp": dp.reshape(P, 3),
            "ok": bool(ok)}


def so3(which, x):
    """0 Exp, 1 Log, 2 J_l, 3 J_l^-1, 4 J_r, 5 J_r^-1, 6 nearest rotation."""
    x = np.ascontiguousarray(x, np.float64).reshape(-1)
    out = np.zeros(9)
    lib().orc_so3(which, _d(x), _d(out))
    return out[:3].copy() if which == 1 else out.reshape(3, 3)
