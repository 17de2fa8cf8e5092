"""Python mirror of cuVSLAM's inertial SBA API over the MI355X C ABI (include/cuvslam_sba_imu.h).

``ImuBAProblem`` has the fields of libs/imu/imu_sba_problem.h:48-91 (as numpy arrays; the per-key-frame
``IMUPreintegration`` objects are records of ``PREINT_DTYPE`` plus one shared array of IMU samples) and
``IMUBundlerGpuFixedVel.solve`` the contract of libs/imu/imu_sba_gpu.h:42-47: the problem is updated in place and
the return value is True when the final cost is below the initial one.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _lib
from .sba import SbaError

PREINT_DTYPE = np.dtype(_lib.ImuPreintegration)
MEAS_DTYPE = np.dtype(_lib.ImuMeasurement)


@dataclass
class ImuCalibration:
    """imu::ImuCalibration (common/imu_calibration.h:29-60); the defaults are its EuRoC values."""
    rig_from_imu: np.ndarray = field(default_factory=lambda: np.eye(4, dtype=np.float32))
    gyroscope_noise_density: float = 0.00016968
    gyroscope_random_walk: float = 0.000019393
    accelerometer_noise_density: float = 0.002
    accelerometer_random_walk: float = 0.003
    frequency: float = 200.0

    def view(self) -> _lib.ImuCalibration:
        c = _lib.ImuCalibration()
        c.rig_from_imu[:] = np.asarray(self.rig_from_imu, np.float32).T.reshape(-1).tolist()
        c.gyroscope_noise_density = self.gyroscope_noise_density
        c.gyroscope_random_walk = self.gyroscope_random_walk
        c.accelerometer_noise_density = self.accelerometer_noise_density
        c.accelerometer_random_walk = self.accelerometer_random_walk
        c.frequency = self.frequency
        return c


@dataclass
class ImuBAProblem:
    """In/out problem of one inertial window. Matrices are indexed [row, col]; the first ``num_fixed_key_frames``
    key frames do not move. ``preintegration[i]`` integrates ``measurements[first : first + num]`` from key frame i
    to key frame i + 1 (its record fields are column-major, as the C ABI)."""
    points: np.ndarray              # (P, 3) float32
    observation_xys: np.ndarray     # (N, 2) float32
    observation_infos: np.ndarray   # (N, 2, 2) float32
    point_ids: np.ndarray           # (N,) int32
    pose_ids: np.ndarray            # (N,) int32
    camera_ids: np.ndarray          # (N,) int8
    camera_from_rig: np.ndarray     # (C, 4, 4) float32
    w_from_imu: np.ndarray          # (K, 4, 4) float32
    velocity: np.ndarray            # (K, 3) float32
    gyro_bias: np.ndarray           # (K, 3) float32
    acc_bias: np.ndarray            # (K, 3) float32
    preintegration: np.ndarray      # (K,) PREINT_DTYPE
    measurements: np.ndarray        # (M,) MEAS_DTYPE
    gravity: np.ndarray = field(default_factory=lambda: np.zeros(3, np.float32))
    num_fixed_key_frames: int = 0
    max_iterations: int = 7
    robustifier_scale: float = 1.0
    robustifier_scale_pose: float = 10.0
    prior_gyro: float = 0.0
    prior_acc: float = 0.0
    imu_penalty: float = 1.0
    acc_rw_penalty: float = -1.0
    boundary_imu_penalty: float = 1e-2
    reintegration_thresh: float = 1e-4
    # solver statistics (ImuBAProblem::iterations, initial_cost) and the MI355X solver's extras
    iterations: int = 0
    initial_cost: float = 0.0
    initial_cost_f64: float = 0.0
    last_cost_f64: float = 0.0
    accepted: int = 0
    status: int = 0
    final_lambda: float = 0.0
    _keep: list = field(default_factory=list, repr=False)

    def copy(self) -> "ImuBAProblem":
        q = ImuBAProblem(*[np.copy(getattr(self, f)) for f in (
            "points", "observation_xys", "observation_infos", "point_ids", "pose_ids", "camera_ids",
            "camera_from_rig", "w_from_imu", "velocity", "gyro_bias", "acc_bias", "preintegration", "measurements",
            "gravity")])
        for f in ("num_fixed_key_frames", "max_iterations", "robustifier_scale", "robustifier_scale_pose",
                  "prior_gyro", "prior_acc", "imu_penalty", "acc_rw_penalty", "boundary_imu_penalty",
                  "reintegration_thresh"):
            setattr(q, f, getattr(self, f))
        return q

    @property
    def num_observations(self) -> int:
        return int(self.point_ids.shape[0])

    def view(self) -> _lib.ImuProblemView:
        """A C view whose buffers stay alive as long as this problem (column-major copies of the matrices)."""
        f32 = lambda a, shape: np.ascontiguousarray(np.asarray(a, np.float32).reshape(shape))  # noqa: E731
        pts = f32(self.points, (-1, 3))
        xy = f32(self.observation_xys, (-1, 2))
        info = np.ascontiguousarray(f32(self.observation_infos, (-1, 2, 2)).transpose(0, 2, 1))
        pid = np.ascontiguousarray(self.point_ids, np.int32)
        qid = np.ascontiguousarray(self.pose_ids, np.int32)
        cid = np.ascontiguousarray(self.camera_ids, np.int8)
        cams = np.ascontiguousarray(f32(self.camera_from_rig, (-1, 4, 4)).transpose(0, 2, 1))
        poses = np.ascontiguousarray(f32(self.w_from_imu, (-1, 4, 4)).transpose(0, 2, 1))
        vel, bg, ba = f32(self.velocity, (-1, 3)), f32(self.gyro_bias, (-1, 3)), f32(self.acc_bias, (-1, 3))
        pre = np.ascontiguousarray(self.preintegration, PREINT_DTYPE)
        meas = np.ascontiguousarray(self.measurements, MEAS_DTYPE)
        self._keep = [pts, xy, info, pid, qid, cid, cams, poses, vel, bg, ba, pre, meas]
        v = _lib.ImuProblemView()
        v.num_points, v.num_observations, v.num_poses = pts.shape[0], pid.shape[0], poses.shape[0]
        v.num_fixed_key_frames, v.num_cameras, v.num_measurements = int(self.num_fixed_key_frames), cams.shape[0], \
            meas.shape[0]
        p = lambda a, t=C.c_float: a.ctypes.data_as(C.POINTER(t))  # noqa: E731
        v.points, v.observation_xys, v.observation_infos = p(pts), p(xy), p(info)
        v.point_ids, v.pose_ids, v.camera_ids = p(pid, C.c_int32), p(qid, C.c_int32), p(cid, C.c_int8)
        v.camera_from_rig, v.w_from_imu, v.velocity, v.gyro_bias, v.acc_bias = p(cams), p(poses), p(vel), p(bg), p(ba)
        v.preintegration = p(pre, _lib.ImuPreintegration)
        v.measurements = p(meas, _lib.ImuMeasurement)
        v.gravity[:] = [float(x) for x in self.gravity]
        v.max_iterations = int(self.max_iterations)
        for name in ("robustifier_scale", "robustifier_scale_pose", "prior_gyro", "prior_acc", "imu_penalty",
                     "acc_rw_penalty", "boundary_imu_penalty", "reintegration_thresh"):
            setattr(v, name, float(getattr(self, name)))
        return v


def integrate(calib: ImuCalibration, samples: np.ndarray, gyro_bias=(0, 0, 0), acc_bias=(0, 0, 0)) -> np.ndarray:
    """IMUPreintegration(gyro_bias, acc_bias) + IntegrateNewMeasurement of every sample (one PREINT_DTYPE record)."""
    s = np.ascontiguousarray(samples, MEAS_DTYPE)
    out = np.zeros(1, PREINT_DTYPE)
    bg = np.asarray(gyro_bias, np.float32)
    ba = np.asarray(acc_bias, np.float32)
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
    rc = _lib.lib().sba_imu_integrate(C.byref(calib.view()), s.ctypes.data_as(C.POINTER(_lib.ImuMeasurement)),
                                      s.shape[0], fp(bg), fp(ba), out.ctypes.data_as(C.POINTER(_lib.ImuPreintegration)))
    if rc != _lib.SBA_OK:
        raise SbaError(rc, "sba_imu_integrate failed")
    return out[0]


def information(pre) -> tuple:
    """(InfoMatrix 9x9, InfoGyroRWMatrix 3x3, InfoAccRWMatrix 3x3) of one preintegration record, [row, col]."""
    rec = np.ascontiguousarray(np.asarray(pre, PREINT_DTYPE).reshape(1))
    i9, ig, ia = np.zeros(81, np.float32), np.zeros(9, np.float32), np.zeros(9, np.float32)
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
    rc = _lib.lib().sba_imu_information(rec.ctypes.data_as(C.POINTER(_lib.ImuPreintegration)), fp(i9), fp(ig), fp(ia))
    if rc != _lib.SBA_OK:
        raise SbaError(rc, "sba_imu_information failed")
    return i9.reshape(9, 9).T.copy(), ig.reshape(3, 3).T.copy(), ia.reshape(3, 3).T.copy()


class IMUBundlerGpuFixedVel:
    """cuvslam::sba_imu::IMUBundlerGpuFixedVel on MI355X (imu_sba_gpu.h:42-47)."""

    def __init__(self, calib: Optional[ImuCalibration] = None, device: int = 0):
        self.calib = calib if calib is not None else ImuCalibration()
        L = _lib.lib()
        opt = _lib.Options()
        L.sba_default_options(C.byref(opt))
        opt.device = device
        h = C.c_void_p()
        rc = L.sba_create(C.byref(opt), C.byref(h))
        if rc != _lib.SBA_OK:
            raise SbaError(rc, "sba_create failed (is a GPU visible?)")
        self._L, self._h = L, h

    def close(self):
        if getattr(self, "_h", None):
            self._L.sba_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def last_error(self) -> str:
        return self._L.sba_last_error(self._h).decode()

    def solve(self, problem: ImuBAProblem) -> bool:
        v = problem.view()
        K, P = problem.w_from_imu.shape[0], problem.points.shape[0]
        pts = np.zeros((P, 3), np.float32)
        poses = np.zeros((K, 4, 4), np.float32)
        vel, bg, ba = (np.zeros((K, 3), np.float32) for _ in range(3))
        pre = np.zeros(K, PREINT_DTYPE)
        r = _lib.ImuResult()
        fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
        r.points, r.w_from_imu, r.velocity, r.gyro_bias, r.acc_bias = fp(pts), fp(poses), fp(vel), fp(bg), fp(ba)
        r.preintegration = pre.ctypes.data_as(C.POINTER(_lib.ImuPreintegration))
        calib = self.calib.view()
        rc = self._L.sba_imu_solve(self._h, C.byref(calib), C.byref(v), C.byref(r))
        if rc in (_lib.SBA_E_DEVICE, _lib.SBA_E_CAPACITY, _lib.SBA_E_INVALID):
            raise SbaError(rc, self.last_error())
        problem.status = int(r.status)
        problem.iterations = int(r.iterations)
        problem.accepted = int(r.accepted)
        problem.initial_cost = float(r.initial_cost)
        problem.initial_cost_f64 = float(r.initial_cost_f64)
        problem.last_cost_f64 = float(r.last_cost_f64)
        problem.final_lambda = float(r.lambda_)
        if rc in (_lib.SBA_OK, _lib.SBA_E_NO_PROGRESS):
            problem.points = pts
            problem.w_from_imu = poses.transpose(0, 2, 1).copy()
            problem.velocity, problem.gyro_bias, problem.acc_bias = vel, bg, ba
            problem.preintegration = pre
        return rc == _lib.SBA_OK
