"""Synthetic inertial SBA windows: the reference's own test scene (libs/cuda_modules/test/imu_sba_test.cpp:40-208,
MakeScenario) restated with numpy's generator.

A rig with an IMU and two cameras (0.12 m baseline) moves with a constant angular rate and a smooth acceleration;
40 IMU samples at 200 Hz lie between consecutive key frames. The true states are integrated with the same discrete
scheme as IMUPreintegration::update_state, so the preintegrated measurements agree with them exactly; points are
seen by `track_length` consecutive key frames. The start perturbs the free key frames (0.01 rad, 0.03 m) and the
points (0.02 m), and every bias starts at zero although the true gyroscope bias is (0.012, -0.02, 0.016). The
solver settings are those of run_imu_sba() (libs/pipelines/service_sba.h). `exact` is the reference's exact
variant: no rotation, no bias, no noise, and the free key frames moved by (0.01, -0.01, 0.01) m.
"""
from __future__ import annotations

import numpy as np

from .imu import MEAS_DTYPE, PREINT_DTYPE, ImuBAProblem, ImuCalibration, integrate as library_integrate

SAMPLES_PER_KEYFRAME = 40
FOCAL = 400.0
PARITY_SHAPE = (7, 2, 200, 7)     # key frames, fixed, points, track length: every point seen by every key frame
INIT_SHAPE = (20, 1, 5900, 2)     # the gravity-initialization refinement: 20 key frames, short tracks


def exp_so3(w):
    w = np.asarray(w, np.float64)
    th = np.linalg.norm(w)
    K = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    if th < 1e-12:
        return np.eye(3) + K
    return np.eye(3) + np.sin(th) / th * K + (1 - np.cos(th)) / th ** 2 * K @ K


def calibration():
    T = np.eye(4)
    T[:3, :3] = exp_so3([0.02, -0.03, 0.01])
    T[:3, 3] = [0.05, -0.01, 0.02]
    return ImuCalibration(T.astype(np.float32), 0.00016968, 0.000019393, 0.002, 0.003, 200.0)


def make_scenario(seed=0, shape=PARITY_SHAPE, exact=False, integrate=None):
    """(ImuBAProblem, ImuCalibration, truth dict). `integrate(calib, samples, bg, ba)` returns one preintegration
    record (default: the library's sba_imu_integrate)."""
    if integrate is None:
        integrate = library_integrate
    K, F, P, track = shape
    rng = np.random.default_rng(seed)
    calib = calibration()
    freq = 200.0
    dt = 1.0 / freq
    dt_ns = int(1e9 / freq)
    g = np.array([0.0, 9.81, 0.0])
    bg_true = np.zeros(3) if exact else np.array([0.012, -0.02, 0.016])
    omega = np.zeros(3) if exact else np.array([0.15, 0.3, -0.25])
    cams = np.tile(np.eye(4, dtype=np.float32), (2, 1, 1))
    cams[1, 0, 3] = -0.12

    R, pos, vel = np.eye(3), np.zeros(3), np.array([0.6, 0.0, 0.1])
    t_ns, t = 0, 0.0
    truth_T, truth_v = [], []
    meas = []
    ranges = []
    for k in range(K):
        T = np.eye(4)
        T[:3, :3], T[:3, 3] = R, pos
        truth_T.append(T)
        truth_v.append(vel.copy())
        first = len(meas)
        if k + 1 < K:
            for _ in range(SAMPLES_PER_KEYFRAME):
                t_ns += dt_ns
                t += dt
                acc_w = np.array([0.4 * np.sin(2 * t), 0.3 * np.cos(1.5 * t), -0.2 * np.sin(t)])
                a_s = R.T @ (acc_w - g)
                meas.append((t_ns, a_s.astype(np.float32), (omega + bg_true).astype(np.float32)))
                pos = pos + vel * dt + 0.5 * g * dt * dt + 0.5 * R @ a_s * dt * dt
                vel = vel + (g + R @ a_s) * dt
                R = R @ exp_so3(omega * dt)
        ranges.append((first, len(meas) - first))
    m = np.zeros(len(meas), MEAS_DTYPE)
    for i, (tn, a, w) in enumerate(meas):
        m[i]["time_ns"], m[i]["linear_acceleration"], m[i]["angular_velocity"] = tn, a, w
    pre = np.zeros(K, PREINT_DTYPE)
    for k, (first, n) in enumerate(ranges):
        rec = integrate(calib, m[first:first + n])
        rec["first_measurement"], rec["num_measurements"] = first, n
        pre[k] = rec

    pts = np.stack([rng.uniform(-4, 4, P) + 1.0, rng.uniform(-4, 4, P), rng.uniform(4, 12, P)], axis=1)
    first_pose = rng.integers(0, K - track + 1, P) if track < K else np.zeros(P, int)
    rig_from_imu = np.asarray(calib.rig_from_imu, np.float64)
    xys, kid, pid, cid = [], [], [], []
    for k in range(K):
        imu_from_w = np.linalg.inv(truth_T[k])
        seen = np.nonzero((first_pose <= k) & (k < first_pose + track))[0]
        for i in seen:
            for c in range(2):
                pc = cams[c].astype(np.float64) @ rig_from_imu @ imu_from_w @ np.append(pts[i], 1.0)
                if pc[2] < 1.0:
                    continue
                uv = pc[:2] / pc[2]
                if np.abs(uv).max() > 1.0:
                    continue
                if not exact:
                    uv = uv + rng.normal(0, 0.3 / FOCAL, 2)
                xys.append(uv)
                kid.append(k)
                pid.append(i)
                cid.append(c)
    N = len(xys)
    infos = np.tile(np.eye(2, dtype=np.float32) * FOCAL * FOCAL, (N, 1, 1))
    poses = np.array(truth_T)
    for k in range(F, K):
        if exact:
            poses[k, :3, 3] += [0.01, -0.01, 0.01]
        else:
            poses[k, :3, :3] = poses[k, :3, :3] @ exp_so3(rng.normal(0, 0.01, 3))
            poses[k, :3, 3] += rng.normal(0, 0.03, 3)
    start_pts = pts if exact else pts + rng.normal(0, 0.02, pts.shape)
    prob = ImuBAProblem(
        points=start_pts.astype(np.float32), observation_xys=np.array(xys, np.float32).reshape(-1, 2),
        observation_infos=infos, point_ids=np.array(pid, np.int32), pose_ids=np.array(kid, np.int32),
        camera_ids=np.array(cid, np.int8), camera_from_rig=cams, w_from_imu=poses.astype(np.float32),
        velocity=np.array(truth_v, np.float32), gyro_bias=np.zeros((K, 3), np.float32),
        acc_bias=np.zeros((K, 3), np.float32), preintegration=pre, measurements=m, gravity=g.astype(np.float32),
        num_fixed_key_frames=F, max_iterations=10, robustifier_scale=1.5, robustifier_scale_pose=-1.0,
        prior_gyro=0.0, prior_acc=0.0, imu_penalty=1e-2, boundary_imu_penalty=1e-2, acc_rw_penalty=0.1,
        reintegration_thresh=1e-3)
    truth = {"w_from_imu": np.array(truth_T), "velocity": np.array(truth_v), "points": pts, "gyro_bias": bg_true}
    return prob, calib, truth


def use_init_settings(problem):
    """RunImuSbaInit() (libs/pipelines/track_online_inertial.cpp): a strong accelerometer-bias prior."""
    problem.prior_acc = 10.0
    return problem
