"""Host-side parts of the inertial path without a GPU: the product's IMU preintegration (sba_imu_integrate, the FP64
integration of imu_preintegration.cpp stored in FP32) and information matrices (sba_imu_information) against the
oracle's, and the argument checks of the C ABI."""
import ctypes as C
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import imu_cases as IC  # noqa: E402
from cuvslam_amd import _lib, imu  # noqa: E402
from oracle import imu_oracle as IO  # noqa: E402

FIELDS = ("dR", "dV", "dP", "JRg", "JVg", "JVa", "JPg", "JPa", "cov", "acc_rw_cov", "gyro_rw_cov", "dT_s",
          "gyro_bias", "acc_bias", "gyro_bias_updated", "acc_bias_updated")


def test_integration_matches_the_oracle_bit_for_bit():
    p, calib, _ = IC.make_scenario(1)
    for k in range(3):
        rec = p.preintegration[k]
        m = p.measurements[rec["first_measurement"]:rec["first_measurement"] + rec["num_measurements"]]
        for bg, ba in (((0, 0, 0), (0, 0, 0)), ((0.01, -0.02, 0.005), (0.1, 0.0, -0.05))):
            a, b = imu.integrate(calib, m, bg, ba), IO.integrate(calib, m, bg, ba)
            for f in FIELDS:
                np.testing.assert_array_equal(a[f], b[f], err_msg=f)


def test_information_matches_the_oracle():
    p, calib, _ = IC.make_scenario(2)
    for k in range(p.preintegration.shape[0]):
        mine, ref = imu.information(p.preintegration[k]), IO.information(p.preintegration[k])
        for x, y in zip(mine, ref):
            np.testing.assert_allclose(x, y, rtol=1e-6, atol=1e-6 * max(np.abs(y).max(), 1e-30))


def test_integration_rejects_bad_arguments():
    calib = IC.calibration()
    out = np.zeros(1, imu.PREINT_DTYPE)
    z = np.zeros(3, np.float32)
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
    cal = calib.view()
    L = _lib.lib()
    po = out.ctypes.data_as(C.POINTER(_lib.ImuPreintegration))
    assert L.sba_imu_integrate(C.byref(cal), None, 3, fp(z), fp(z), po) == _lib.SBA_E_INVALID
    assert L.sba_imu_integrate(C.byref(cal), None, -1, fp(z), fp(z), po) == _lib.SBA_E_INVALID
    cal.frequency = 0.0
    assert L.sba_imu_integrate(C.byref(cal), None, 0, fp(z), fp(z), po) == _lib.SBA_E_INVALID
