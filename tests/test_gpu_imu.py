"""GPU parity of the inertial SBA (sba_imu_solve, the path of IMUBundlerGpuFixedVel::solve,
libs/imu/imu_sba_gpu_fixed_vel.cpp:75-363) against the FP64 oracle on the reference's own test scenes
(libs/cuda_modules/test/imu_sba_test.cpp): both compute in FP64, so every LM decision is the oracle's and the final
costs agree to 1e-9 relative; the returned FP32 state agrees to its rounding."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import imu_cases as IC  # noqa: E402
from cuvslam_amd import _lib, imu  # noqa: E402
from oracle import imu_oracle as IO  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bundler():
    return imu.IMUBundlerGpuFixedVel(IC.calibration())


def check_parity(q, ref, rel=1e-9):
    assert q.status == ref["status"]
    assert (q.iterations, q.accepted) == (ref["iterations"], ref["accepted"])
    assert q.final_lambda == np.float32(ref["lam"])
    assert abs(q.initial_cost_f64 - ref["initial_cost"]) <= 1e-12 * ref["initial_cost"]
    assert abs(q.last_cost_f64 - ref["last_cost"]) <= rel * ref["last_cost"], (q.last_cost_f64, ref["last_cost"])
    np.testing.assert_allclose(q.points, ref["points"], rtol=0, atol=1e-5)
    np.testing.assert_allclose(q.w_from_imu[:, :3, :3], ref["w_from_imu"][:, :3, :3], rtol=0, atol=1e-6)
    np.testing.assert_allclose(q.w_from_imu[:, :3, 3], ref["w_from_imu"][:, :3, 3], rtol=0, atol=1e-5)
    np.testing.assert_allclose(q.velocity, ref["velocity"], rtol=0, atol=1e-5)
    np.testing.assert_allclose(q.gyro_bias, ref["gyro_bias"], rtol=0, atol=1e-7)
    np.testing.assert_allclose(q.acc_bias, ref["acc_bias"], rtol=0, atol=1e-6)
    for f in ("gyro_bias", "acc_bias", "gyro_bias_updated", "acc_bias_updated", "dR", "dV", "dP"):
        np.testing.assert_allclose(q.preintegration[f], ref["preintegration"][f], rtol=0, atol=2e-6, err_msg=f)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_parity_scene_matches_oracle(bundler, seed):
    # MatchesCpuBundler's scene (imu_sba_test.cpp:232-246): 7 key frames (2 fixed), 200 points seen by all
    p, calib, _ = IC.make_scenario(seed)
    ref = IO.solve(p, calib)
    q = p.copy()
    assert bundler.solve(q)
    check_parity(q, ref)


@pytest.mark.parametrize("seed", [1, 2])
def test_init_scene_matches_oracle(bundler, seed):
    # MatchesCpuBundlerWithInitSettings (imu_sba_test.cpp:258-282): 20 key frames, 5900 points on two-frame tracks
    p, calib, _ = IC.make_scenario(seed, shape=IC.INIT_SHAPE)
    IC.use_init_settings(p)
    ref = IO.solve(p, calib)
    q = p.copy()
    assert bundler.solve(q)
    check_parity(q, ref)
    assert q.last_cost_f64 < 1e-2 * q.initial_cost_f64


def test_exact_scene_early_exit(bundler):
    # HandlesNearIdentityRotationResiduals (imu_sba_test.cpp:284-297)
    p, calib, truth = IC.make_scenario(1, exact=True)
    ref = IO.solve(p, calib)
    q = p.copy()
    assert bundler.solve(q)
    assert q.iterations == ref["iterations"] < p.max_iterations
    assert q.accepted == ref["accepted"]
    np.testing.assert_allclose(q.w_from_imu[:, :3, 3], truth["w_from_imu"][:, :3, 3], atol=1e-4)


def test_empty_problem_fails(bundler):
    p, calib, _ = IC.make_scenario(1)
    q = p.copy()
    keep = q.pose_ids == 0
    q.w_from_imu, q.velocity, q.gyro_bias, q.acc_bias = (a[:1] for a in (q.w_from_imu, q.velocity, q.gyro_bias,
                                                                          q.acc_bias))
    q.preintegration = q.preintegration[:1]
    q.num_fixed_key_frames = 0
    for f in ("observation_xys", "observation_infos", "point_ids", "pose_ids", "camera_ids"):
        setattr(q, f, getattr(q, f)[keep])
    assert not bundler.solve(q)
    assert q.status == _lib.SBA_E_EMPTY


def test_repeated_solves_are_bitwise_identical(bundler):
    p, calib, _ = IC.make_scenario(4, shape=IC.INIT_SHAPE)
    a, b = p.copy(), p.copy()
    assert bundler.solve(a) and bundler.solve(b)
    assert a.last_cost_f64 == b.last_cost_f64
    np.testing.assert_array_equal(a.points, b.points)
    np.testing.assert_array_equal(a.w_from_imu, b.w_from_imu)


def test_more_free_key_frames_than_supported_is_a_capacity_error(bundler):
    p, calib, _ = IC.make_scenario(1, shape=(70, 1, 300, 3))
    with pytest.raises(Exception) as e:
        bundler.solve(p)
    assert e.value.code == _lib.SBA_E_CAPACITY
