"""The C++ drop-in class of the inertial path (include/sba/imu_sba.h) run by a plain g++ program of the reference's
API, tests/cpp/dropin_imu.cpp, as a child process: it rebuilds the window's preintegrations sample by sample
through IMUPreintegration::IntegrateNewMeasurement, solves with IMUBundlerGpuFixedVel and is checked against the FP64
oracle's solve of the same window."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import imu_cases as IC  # noqa: E402
from oracle import imu_oracle as IO  # noqa: E402

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "dropin_imu")


def write_problem(path, p, calib):
    K = p.w_from_imu.shape[0]
    with open(path, "wb") as f:
        f.write(np.array([p.points.shape[0], p.num_observations, K, p.num_fixed_key_frames,
                          p.camera_from_rig.shape[0]], np.int32).tobytes())
        f.write(np.asarray(p.points, np.float32).tobytes())
        f.write(np.asarray(p.observation_xys, np.float32).tobytes())
        f.write(np.asarray(p.observation_infos, np.float32).transpose(0, 2, 1).tobytes())
        for a in (p.point_ids, p.pose_ids, p.camera_ids):
            f.write(np.asarray(a, np.int32).tobytes())
        f.write(np.asarray(p.camera_from_rig, np.float32).transpose(0, 2, 1).tobytes())
        f.write(np.asarray(p.w_from_imu, np.float32).transpose(0, 2, 1).tobytes())
        for a in (p.velocity, p.gyro_bias, p.acc_bias):
            f.write(np.asarray(a, np.float32).tobytes())
        f.write(np.asarray(p.preintegration["num_measurements"], np.int32).tobytes())
        for rec in p.preintegration:
            for s in p.measurements[rec["first_measurement"]:rec["first_measurement"] + rec["num_measurements"]]:
                f.write(np.int64(s["time_ns"]).tobytes())
                f.write(np.asarray(s["linear_acceleration"], np.float32).tobytes())
                f.write(np.asarray(s["angular_velocity"], np.float32).tobytes())
        f.write(np.asarray(p.gravity, np.float32).tobytes())
        f.write(np.asarray(calib.rig_from_imu, np.float32).T.tobytes())
        f.write(np.array([calib.gyroscope_noise_density, calib.gyroscope_random_walk,
                          calib.accelerometer_noise_density, calib.accelerometer_random_walk, calib.frequency],
                         np.float32).tobytes())
        f.write(np.int32(p.max_iterations).tobytes())
        f.write(np.array([p.robustifier_scale, p.robustifier_scale_pose, p.prior_gyro, p.prior_acc, p.imu_penalty,
                          p.acc_rw_penalty, p.boundary_imu_penalty, p.reintegration_thresh], np.float32).tobytes())


@pytest.mark.parametrize("seed,shape", [(1, IC.PARITY_SHAPE), (2, IC.INIT_SHAPE)])
def test_dropin_class_matches_oracle(tmp_path, seed, shape):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (__graft_entry__.build() builds it)")
    p, calib, _ = IC.make_scenario(seed, shape=shape)
    path = os.path.join(tmp_path, "window.bin")
    write_problem(path, p, calib)
    out = subprocess.run([EXE, path], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    got = json.loads(out.stdout.strip().splitlines()[-1])
    ref = IO.solve(p, calib)
    assert got["ok"] is True and got["iterations"] == ref["iterations"]
    assert abs(got["initial_cost"] - ref["initial_cost"]) <= 1e-6 * ref["initial_cost"]
    np.testing.assert_allclose(got["p0"], ref["points"][0], rtol=0, atol=1e-5)
    np.testing.assert_allclose(np.array(got["t"]), ref["w_from_imu"][:, :3, 3], rtol=0, atol=1e-5)
    np.testing.assert_allclose(np.array(got["bg"]), ref["gyro_bias"], rtol=0, atol=1e-7)
    # the first free key frame's preintegration, as the reference leaves it: re-integrated at the final biases when
    # they moved past the threshold, with its samples kept
    F = p.num_fixed_key_frames
    np.testing.assert_allclose(got["pre_bg"], ref["preintegration"][F]["gyro_bias"], rtol=0, atol=1e-7)
    assert got["pre_samples"] == p.preintegration[F]["num_measurements"]
