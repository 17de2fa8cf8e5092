"""Inertial SBA windows for the tests: the package's restatement of the reference's test scene
(cuvslam_amd/imu_scene.py, libs/cuda_modules/test/imu_sba_test.cpp:40-208) with the preintegrations computed by the
oracle, so the inputs of a parity test do not depend on the code under test."""
from cuvslam_amd import imu_scene
from cuvslam_amd.imu_scene import INIT_SHAPE, PARITY_SHAPE, calibration, exp_so3, use_init_settings  # noqa: F401


def make_scenario(seed=0, shape=PARITY_SHAPE, exact=False, integrate=None):
    if integrate is None:
        from oracle import imu_oracle
        integrate = imu_oracle.integrate
    return imu_scene.make_scenario(seed, shape=shape, exact=exact, integrate=integrate)
