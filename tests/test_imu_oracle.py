"""Pins of the inertial-SBA oracle (oracle/imu_oracle.cpp) against what the mathematics fixes and the reference's own
test expectations (libs/cuda_modules/test/imu_sba_test.cpp):
  * SO(3): Exp against scipy's matrix exponential, Log as its inverse (near the identity and near pi), the left and
    right Jacobians by finite differences, the nearest rotation against numpy's SVD;
  * preintegration: closed forms of constant rates, the bias Jacobians by finite differences, the covariance by a
    Monte Carlo of noisy integrations, the random-walk covariances, the information matrices as inverses;
  * the reference scene's construction: its true states satisfy the inertial residuals exactly;
  * the model: every inertial and reprojection Jacobian by central differences of the residuals;
  * the full system: its right-hand side is -1/2 the cost's gradient and its matrix 1/2 the cost's curvature at a
    zero-residual state (second differences along random directions);
  * the step: the Schur path against a dense solve of the damped full system;
  * the LM loop: the damping schedule of its trace, the reference's convergence expectations, the empty problem.
"""
import os
import sys

import numpy as np
import pytest
from scipy.linalg import expm

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import imu_cases as IC  # noqa: E402
from cuvslam_amd import _lib  # noqa: E402
from cuvslam_amd.imu import MEAS_DTYPE  # noqa: E402
from oracle import imu_oracle as IO  # noqa: E402


def skew(w):
    return np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])


def log_np(R):
    return IO.so3(1, R)


# ---- SO(3) --------------------------------------------------------------------------------------------------------

@pytest.mark.parametrize("w", [[0.3, -0.2, 0.5], [1e-7, 2e-7, -1e-7], [0.0, 3.1, 0.0], [2.0, 1.0, -1.5]])
def test_exp_log(w):
    w = np.array(w)
    R = IO.so3(0, w)
    np.testing.assert_allclose(R, expm(skew(w)), atol=1e-13)
    np.testing.assert_allclose(log_np(R), w, atol=1e-9 if np.linalg.norm(w) > 3 else 1e-13)


def test_log_near_pi_and_identity():
    axis = np.array([0.3, -0.5, 0.8]) / np.linalg.norm([0.3, -0.5, 0.8])
    for th in (1e-9, 1e-5, 0.5, 3.0, np.pi - 1e-6):
        np.testing.assert_allclose(log_np(expm(skew(th * axis))), th * axis, atol=1e-8 * max(1, th))


@pytest.mark.parametrize("x", [[0.2, -0.4, 0.1], [1e-9, 0, 2e-9], [1.2, 0.3, -0.7]])
def test_jacobians_by_finite_differences(x):
    x = np.array(x)
    Jl, Jli, Jr, Jri = (IO.so3(k, x) for k in (2, 3, 4, 5))
    np.testing.assert_allclose(Jl @ Jli, np.eye(3), atol=1e-12)
    np.testing.assert_allclose(Jr @ Jri, np.eye(3), atol=1e-12)
    h = 1e-6
    for k in range(3):
        d = np.zeros(3)
        d[k] = h
        # Exp(x + d) = Exp(Jl d) Exp(x) = Exp(x) Exp(Jr d)
        fl = (log_np(expm(skew(x + d)) @ expm(skew(x)).T) - log_np(expm(skew(x - d)) @ expm(skew(x)).T)) / (2 * h)
        fr = (log_np(expm(skew(x)).T @ expm(skew(x + d))) - log_np(expm(skew(x)).T @ expm(skew(x - d)))) / (2 * h)
        np.testing.assert_allclose(fl, Jl[:, k], atol=1e-8)
        np.testing.assert_allclose(fr, Jr[:, k], atol=1e-8)


def test_nearest_rotation_is_the_svd_polar_factor():
    rng = np.random.default_rng(0)
    for _ in range(5):
        A = expm(skew(rng.normal(size=3))) + 0.05 * rng.normal(size=(3, 3))
        U, _, Vt = np.linalg.svd(A)
        np.testing.assert_allclose(IO.so3(6, A), U @ Vt, atol=1e-12)


# ---- preintegration -----------------------------------------------------------------------------------------------

def samples(n, acc, gyro, dt_ns=5_000_000):
    m = np.zeros(n, MEAS_DTYPE)
    m["time_ns"] = np.arange(1, n + 1) * dt_ns
    m["linear_acceleration"] = np.asarray(acc, np.float32)
    m["angular_velocity"] = np.asarray(gyro, np.float32)
    return m


def test_preintegration_closed_forms():
    calib = IC.calibration()
    n, dt = 40, 1.0 / 200
    w = np.array([0.15, 0.3, -0.25], np.float32).astype(np.float64)
    r = IO.integrate_f64(calib, samples(n, [0, 0, 0], w))
    np.testing.assert_allclose(r["dR"], expm(skew(w * n * dt)), atol=1e-12)
    np.testing.assert_allclose(r["dV"], 0, atol=1e-15)
    np.testing.assert_allclose(r["dP"], 0, atol=1e-15)
    assert abs(r["dT"] - n * dt) < 1e-12
    a = np.array([0.4, -9.7, 0.3], np.float32).astype(np.float64)
    r = IO.integrate_f64(calib, samples(n, a, [0, 0, 0]))
    T = n * dt
    np.testing.assert_allclose(r["dR"], np.eye(3), atol=1e-15)
    np.testing.assert_allclose(r["dV"], a * T, rtol=1e-12)
    np.testing.assert_allclose(r["dP"], 0.5 * a * T * T, rtol=1e-12)
    # the random walks accumulate one sample's covariance per sample (imu_calibration.cpp:44-45)
    np.testing.assert_allclose(r["cov_grw"], np.eye(3) * 0.000019393 ** 2 / 200 * n, rtol=1e-6)
    np.testing.assert_allclose(r["cov_arw"], np.eye(3) * 0.003 ** 2 / 200 * n, rtol=1e-6)


def test_preintegration_bias_jacobians_by_finite_differences():
    calib = IC.calibration()
    rng = np.random.default_rng(1)
    m = samples(40, [0, 0, 0], [0, 0, 0])
    t = np.arange(40) / 200
    m["linear_acceleration"] = np.stack([0.4 * np.sin(2 * t), -9.81 + 0.3 * np.cos(t), 0.2 * t], 1).astype(np.float32)
    m["angular_velocity"] = (np.array([0.15, 0.3, -0.25]) + 0.05 * rng.normal(size=(40, 3))).astype(np.float32)
    bg0, ba0 = np.array([0.01, -0.02, 0.005]), np.array([0.1, 0.0, -0.05])
    r0 = IO.integrate_f64(calib, m, bg0, ba0)
    h = 1e-5
    for k in range(3):
        d = np.zeros(3)
        d[k] = h
        rp, rm = IO.integrate_f64(calib, m, bg0 + d, ba0), IO.integrate_f64(calib, m, bg0 - d, ba0)
        # dR(bg + d) = dR(bg) Exp(JRg d)
        np.testing.assert_allclose((log_np(r0["dR"].T @ rp["dR"]) - log_np(r0["dR"].T @ rm["dR"])) / (2 * h),
                                   r0["JRg"][:, k], atol=1e-7)
        np.testing.assert_allclose((rp["dV"] - rm["dV"]) / (2 * h), r0["JVg"][:, k], atol=1e-7)
        np.testing.assert_allclose((rp["dP"] - rm["dP"]) / (2 * h), r0["JPg"][:, k], atol=1e-7)
        rp, rm = IO.integrate_f64(calib, m, bg0, ba0 + d), IO.integrate_f64(calib, m, bg0, ba0 - d)
        np.testing.assert_allclose((rp["dV"] - rm["dV"]) / (2 * h), r0["JVa"][:, k], atol=1e-8)
        np.testing.assert_allclose((rp["dP"] - rm["dP"]) / (2 * h), r0["JPa"][:, k], atol=1e-8)


def test_preintegration_covariance_by_monte_carlo():
    # noisy integrations of the same motion, errors measured the way the propagation defines them: rotation on the
    # right (dR_noisy = dR Exp(e_R)), velocity and position by difference
    calib = IC.calibration()
    n, dt, trials = 40, 1.0 / 200, 20000
    t = np.arange(n) * dt
    acc = np.stack([0.4 * np.sin(2 * t), -9.81 + 0.3 * np.cos(t), 0.2 * t], 1).astype(np.float32).astype(np.float64)
    gyr = np.tile(np.array([0.15, 0.3, -0.25], np.float32).astype(np.float64), (n, 1))
    ref = IO.integrate_f64(calib, samples(n, acc, gyr))
    rng = np.random.default_rng(2)
    sg, sa = 0.00016968 * np.sqrt(200), 0.002 * np.sqrt(200)
    R = np.tile(np.eye(3), (trials, 1, 1))
    V, P = np.zeros((trials, 3)), np.zeros((trials, 3))
    for k in range(n):
        a = acc[k] + sa * rng.normal(size=(trials, 3))
        w = gyr[k] + sg * rng.normal(size=(trials, 3))
        Ra = np.einsum("tij,tj->ti", R, a)
        P = P + V * dt + 0.5 * Ra * dt * dt
        V = V + Ra * dt
        th = np.linalg.norm(w * dt, axis=1)[:, None, None]
        K = np.zeros((trials, 3, 3))
        wd = w * dt
        K[:, 0, 1], K[:, 0, 2], K[:, 1, 2] = -wd[:, 2], wd[:, 1], -wd[:, 0]
        K = K - K.transpose(0, 2, 1)
        E = np.eye(3) + np.sin(th) / th * K + (1 - np.cos(th)) / th ** 2 * K @ K
        R = R @ E
    eR = np.array([log_np(ref["dR"].T @ Rk) for Rk in R])
    err = np.concatenate([eR, V - ref["dV"], P - ref["dP"]], axis=1)
    C = np.cov(err.T)
    cov = ref["cov"]
    s = np.sqrt(np.diag(cov))
    # correlations and variances agree within the Monte Carlo error (20000 trials: ~2% on variances)
    np.testing.assert_allclose(np.diag(C) / np.diag(cov), 1, atol=0.05)
    np.testing.assert_allclose(C / np.outer(s, s), cov / np.outer(s, s), atol=0.03)


def test_information_is_the_inverse():
    calib = IC.calibration()
    pre = IO.integrate(calib, samples(40, [0.3, -9.8, 0.1], [0.1, 0.2, -0.3]))
    info, ig, ia = IO.information(pre)
    cov = np.asarray(pre["cov"], np.float64).reshape(9, 9).T
    np.testing.assert_allclose(info @ cov, np.eye(9), atol=1e-8)
    np.testing.assert_allclose(ig @ np.asarray(pre["gyro_rw_cov"], np.float64).reshape(3, 3).T, np.eye(3), atol=1e-9)
    np.testing.assert_allclose(ia @ np.asarray(pre["acc_rw_cov"], np.float64).reshape(3, 3).T, np.eye(3), atol=1e-9)
    empty = np.zeros(1, pre.dtype)[0]
    assert not IO.information(empty)[0].any()


# ---- the scene and the model --------------------------------------------------------------------------------------

def at_state(p, pts=None, T=None, vel=None, bg=None, ba=None):
    q = p.copy()
    if pts is not None:
        q.points = np.asarray(pts, np.float32)
    if T is not None:
        q.w_from_imu = np.asarray(T, np.float32)
    if vel is not None:
        q.velocity = np.asarray(vel, np.float32)
    if bg is not None:
        q.gyro_bias = np.asarray(bg, np.float32)
    if ba is not None:
        q.acc_bias = np.asarray(ba, np.float32)
    return q


def test_true_states_satisfy_the_inertial_residuals():
    # the reference's construction (imu_sba_test.cpp:84-86): truth is integrated with update_state's scheme, so at
    # the true states (and the preintegration at the true biases) the inertial residuals vanish
    p, calib, truth = IC.make_scenario(3)
    K = p.w_from_imu.shape[0]
    bg = np.tile(truth["gyro_bias"], (K, 1))
    q = at_state(p, truth["points"], truth["w_from_imu"], truth["velocity"], bg, np.zeros((K, 3)))
    for k in range(K):
        rec = p.preintegration[k]
        f, n = rec["first_measurement"], rec["num_measurements"]
        new = IO.integrate(calib, p.measurements[f:f + n], bg[k], (0, 0, 0))
        new["first_measurement"], new["num_measurements"] = f, n
        q.preintegration[k] = new
    er, _, _, _, _ = IO.model(q, calib)
    assert np.abs(er[:, :3]).max() < 2e-6      # rad (float32 storage of states and samples)
    assert np.abs(er[:, 3:9]).max() < 1e-5     # m/s, m
    assert np.abs(er[:, 9:]).max() == 0        # equal biases


def perturbed(p, k, var, h):
    """The problem with key frame k's variable `var` (0 rotation (right), 1 translation (t + R d), 2 velocity,
    3 gyro bias, 4 acc bias) moved by the 3-vector h (FP64 state kept in a float32 problem: h must be large enough)."""
    q = p.copy()
    T = q.w_from_imu.astype(np.float64)
    if var == 0:
        T[k, :3, :3] = T[k, :3, :3] @ expm(skew(h))
    elif var == 1:
        T[k, :3, 3] = T[k, :3, 3] + T[k, :3, :3] @ h
    q.w_from_imu = T.astype(np.float32)
    if var == 2:
        q.velocity = q.velocity.astype(np.float64)
        q.velocity[k] += h
    if var == 3:
        q.gyro_bias = q.gyro_bias.astype(np.float64)
        q.gyro_bias[k] += h
    if var == 4:
        q.acc_bias = q.acc_bias.astype(np.float64)
        q.acc_bias[k] += h
    return q


def residual_vector(er):
    # the inertial residual with its rotation part compared on the manifold is linear enough for central differences
    return er[:, :9]


def test_inertial_jacobians_by_finite_differences():
    p, calib, _ = IC.make_scenario(4, shape=(4, 1, 60, 4))
    # a consistent linearization point: biases away from the integration biases, SetNewBias applied
    K = p.w_from_imu.shape[0]
    rng = np.random.default_rng(4)
    p.gyro_bias = (0.01 * rng.normal(size=(K, 3))).astype(np.float32)
    p.acc_bias = (0.05 * rng.normal(size=(K, 3))).astype(np.float32)
    p.preintegration["gyro_bias_updated"] = p.gyro_bias
    p.preintegration["acc_bias_updated"] = p.acc_bias
    er, ej, _, _, _ = IO.model(p, calib)
    h = 1e-3
    for i in range(K - 1):
        for side, k in ((0, i), (1, i + 1)):
            for var in range(5 if side == 0 else 3):
                j = var if side == 0 else 5 + var
                num = np.zeros((9, 3))
                for c in range(3):
                    d = np.zeros(3)
                    d[c] = h
                    qp, qm = perturbed(p, k, var, d), perturbed(p, k, var, -d)
                    if var == 3:  # the Jacobian is taken at the stored updated bias
                        qp.preintegration["gyro_bias_updated"] = qp.gyro_bias
                        qm.preintegration["gyro_bias_updated"] = qm.gyro_bias
                    rp, rm = IO.model(qp, calib)[0][i, :9], IO.model(qm, calib)[0][i, :9]
                    num[:, c] = (rp - rm) / (2 * h)
                # float32 states: the finite differences carry ~1e-4 relative rounding
                np.testing.assert_allclose(num, ej[i, j], atol=2e-4 * max(1.0, np.abs(ej[i, j]).max()),
                                           err_msg=f"edge {i} jacobian {j}")


def test_reprojection_jacobians_by_finite_differences():
    p, calib, _ = IC.make_scenario(5, shape=(3, 1, 20, 3))
    _, _, orr, ow, oj = IO.model(p, calib)
    h = 1e-3
    for o in range(0, p.num_observations, 7):
        k, i = p.pose_ids[o], p.point_ids[o]
        for var, J in ((0, oj[o, 0]), (1, oj[o, 1])):
            num = np.zeros((2, 3))
            for c in range(3):
                d = np.zeros(3)
                d[c] = h
                num[:, c] = (IO.model(perturbed(p, k, var, d), calib)[2][o]
                             - IO.model(perturbed(p, k, var, -d), calib)[2][o]) / (2 * h)
            np.testing.assert_allclose(num, J, atol=1e-4 * max(1, np.abs(J).max()))
        num = np.zeros((2, 3))
        for c in range(3):
            qp, qm = p.copy(), p.copy()
            qp.points = qp.points.astype(np.float64)
            qm.points = qm.points.astype(np.float64)
            qp.points[i, c] += h
            qm.points[i, c] -= h
            num[:, c] = (IO.model(qp, calib)[2][o] - IO.model(qm, calib)[2][o]) / (2 * h)
        np.testing.assert_allclose(num, oj[o, 2], atol=1e-4 * max(1, np.abs(oj[o, 2]).max()))


def free_variables(p):
    F, K, P = p.num_fixed_key_frames, p.w_from_imu.shape[0], p.points.shape[0]
    return [(k, var) for k in range(F, K) for var in range(5)], P


def cost_along(p, calib, dx, pose_vars, P):
    """The oracle's cost after moving the free variables by dx (pose blocks of 15, then 3 per point), in FP64."""
    T = p.w_from_imu.astype(np.float64).copy()
    vel, bg, ba = (a.astype(np.float64).copy() for a in (p.velocity, p.gyro_bias, p.acc_bias))
    for n, (k, var) in enumerate(pose_vars):
        d = dx[3 * n:3 * n + 3]
        if var == 0:
            T[k, :3, :3] = T[k, :3, :3] @ expm(skew(d))
        elif var == 1:
            T[k, :3, 3] += T[k, :3, :3] @ d
        elif var == 2:
            vel[k] += d
        elif var == 3:
            bg[k] += d
        else:
            ba[k] += d
    pts = p.points.astype(np.float64) + dx[3 * len(pose_vars):].reshape(P, 3)
    return IO.cost(p, calib, pts, T, vel, bg, ba)


def test_full_system_is_half_the_gradient_and_curvature_of_the_cost():
    p, calib, truth = IC.make_scenario(6, shape=(4, 1, 30, 4), exact=True)
    # the exact scene at its true states: zero residuals, inliers of the Huber loss, so the cost is quadratic to
    # second order and 1/2 its Hessian is the Gauss-Newton matrix
    K = p.w_from_imu.shape[0]
    q = at_state(p, truth["points"], truth["w_from_imu"], truth["velocity"], np.zeros((K, 3)), np.zeros((K, 3)))
    pose_vars, P = free_variables(q)
    S = IO.system(q, calib, 0.0)
    n15 = S["U"].shape[0]
    H = np.zeros((n15 + 3 * P, n15 + 3 * P))
    H[:n15, :n15] = S["U"]
    for i in range(P):
        H[n15 + 3 * i:n15 + 3 * i + 3, n15 + 3 * i:n15 + 3 * i + 3] = S["V"][i]
    for i in range(P):
        for j in range(S["W"].shape[1] // 6):
            blk = S["W"][3 * i:3 * i + 3, 6 * j:6 * j + 6]
            H[n15 + 3 * i:n15 + 3 * i + 3, 15 * j:15 * j + 6] = blk
            H[15 * j:15 * j + 6, n15 + 3 * i:n15 + 3 * i + 3] = blk.T
    rng = np.random.default_rng(6)
    c0 = cost_along(q, calib, np.zeros(n15 + 3 * P), pose_vars, P)
    for _ in range(4):
        v = rng.normal(size=n15 + 3 * P)
        v *= 1e-4 / np.linalg.norm(v)
        cp, cm = cost_along(q, calib, v, pose_vars, P), cost_along(q, calib, -v, pose_vars, P)
        curv = cp - 2 * c0 + cm
        assert abs(curv - 2 * v @ H @ v) < 2e-3 * abs(curv) + 1e-12, (curv, 2 * v @ H @ v)
    # the right-hand side at a state with residuals: e = -1/2 grad cost (central differences)
    S = IO.system(p, calib, 0.0)
    pose_vars, P = free_variables(p)
    e = np.concatenate([S["ea"], S["eb"]])
    h = 1e-5
    for idx in list(range(0, n15, 2)) + list(range(n15, n15 + 3 * P, 5)):
        d = np.zeros(n15 + 3 * P)
        d[idx] = h
        g = (cost_along(p, calib, d, pose_vars, P) - cost_along(p, calib, -d, pose_vars, P)) / (2 * h)
        assert abs(-0.5 * g - e[idx]) < 1e-4 * max(1.0, abs(e[idx])), (idx, -0.5 * g, e[idx])


def test_step_matches_a_dense_solve_of_the_damped_full_system():
    p, calib, _ = IC.make_scenario(7, shape=(5, 2, 40, 5))
    for lam in (1e2, 1.0, 1e-3):
        S = IO.system(p, calib, lam)
        assert S["ok"]
        n15, P = S["U"].shape[0], S["V"].shape[0]
        H = np.zeros((n15 + 3 * P, n15 + 3 * P))
        H[:n15, :n15] = S["U"]
        for i in range(P):
            H[n15 + 3 * i:n15 + 3 * i + 3, n15 + 3 * i:n15 + 3 * i + 3] = S["V"][i]
            for j in range(S["W"].shape[1] // 6):
                blk = S["W"][3 * i:3 * i + 3, 6 * j:6 * j + 6]
                H[n15 + 3 * i:n15 + 3 * i + 3, 15 * j:15 * j + 6] = blk
                H[15 * j:15 * j + 6, n15 + 3 * i:n15 + 3 * i + 3] = blk.T
        Hd = H + lam * np.diag(np.diag(H))
        e = np.concatenate([S["ea"], S["eb"]])
        x = np.linalg.solve(Hd, e)
        y = np.concatenate([S["dc"], S["dp"].reshape(-1)])
        # backward error of the Schur path's step on the dense system, and the forward error its condition allows
        # (the inertial information reaches 1e8 against 1e5 for the reprojections: cond ~1e9)
        assert np.linalg.norm(Hd @ y - e) <= 1e-10 * np.linalg.norm(e)
        bound = 100 * np.linalg.cond(Hd) * np.finfo(float).eps
        assert np.linalg.norm(y - x) <= bound * np.linalg.norm(x), (np.linalg.norm(y - x) / np.linalg.norm(x), bound)


def test_unobserved_points_step_zero():
    # the reference's pivoted LDLT leaves a zero block's solution at zero (cuda_sba_common.h:392-413)
    p, calib, _ = IC.make_scenario(8, shape=(4, 1, 30, 4))
    p.points = np.concatenate([p.points, [[0.5, 0.5, 6.0]]]).astype(np.float32)
    S = IO.system(p, calib, 1.0)
    assert S["ok"] and np.all(S["dp"][-1] == 0) and np.isfinite(S["dp"]).all()


# ---- the LM loop --------------------------------------------------------------------------------------------------

def check_schedule(r):
    """The damping each iteration used: 1e2 first (imu_sba_gpu_fixed_vel.cpp:134), x0.125 after a gain ratio above
    0.75, kept after one in (0.25, 0.75], x5 after a rejection, a failed factorization or a non-finite trial cost."""
    lam = np.float32(1e2)
    for k, (f, l, rho) in enumerate(zip(r["trace_flags"], r["trace_lambda"], r["trace_rho"])):
        assert l == lam, (k, l, lam)
        if f == 1:
            assert rho > 0.25
            lam = lam * np.float32(0.125) if rho > 0.75 else lam
        elif f == 0:
            assert rho <= 0.25
            lam = lam * np.float32(5)
        elif f in (2, 8):
            lam = lam * np.float32(5)
    if r["trace_flags"][-1] != 4:
        assert r["lam"] == lam


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_reference_parity_scene_converges_with_the_reference_schedule(seed):
    p, calib, _ = IC.make_scenario(seed)
    r = IO.solve(p, calib)
    assert r["status"] == _lib.SBA_OK and r["iterations"] == 10
    assert r["last_cost"] < 1e-2 * r["initial_cost"]
    check_schedule(r)
    # the returned state has the reported cost (up to the re-integrations at the moved biases inside the solve)
    np.testing.assert_allclose(IO.cost(p, calib, r["points"], r["w_from_imu"], r["velocity"], r["gyro_bias"],
                                       r["acc_bias"]), r["last_cost"], rtol=1e-5)


@pytest.mark.parametrize("seed", [1, 2])
def test_reference_init_scene_reaches_one_percent_of_its_cost(seed):
    # MatchesCpuBundlerWithInitSettings (imu_sba_test.cpp:258-282): cpu_cost < 1e-2 * initial_cost
    p, calib, _ = IC.make_scenario(seed, shape=IC.INIT_SHAPE)
    IC.use_init_settings(p)
    r = IO.solve(p, calib)
    assert r["status"] == _lib.SBA_OK
    assert r["last_cost"] < 1e-2 * r["initial_cost"]
    check_schedule(r)


def test_exact_scene_converges_with_the_early_exit():
    # HandlesNearIdentityRotationResiduals (imu_sba_test.cpp:284-297): the exact scene, identity rotation residuals
    p, calib, truth = IC.make_scenario(1, exact=True)
    r = IO.solve(p, calib)
    assert r["status"] == _lib.SBA_OK
    assert r["trace_flags"][-1] == 4 and r["iterations"] < 10
    assert r["last_cost"] < 1e-6 * r["initial_cost"]
    np.testing.assert_allclose(r["w_from_imu"][:, :3, 3], truth["w_from_imu"][:, :3, 3], atol=1e-4)


def test_empty_problem_fails_like_the_reference():
    # EmptyProblemFailsLikeCpu (imu_sba_test.cpp:346-361): one key frame, no inertial edge
    p, calib, _ = IC.make_scenario(1)
    q = p.copy()
    keep = q.pose_ids == 0
    q.w_from_imu, q.velocity, q.gyro_bias, q.acc_bias = (a[:1] for a in (q.w_from_imu, q.velocity, q.gyro_bias,
                                                                          q.acc_bias))
    q.preintegration = q.preintegration[:1]
    q.num_fixed_key_frames = 0
    for f in ("observation_xys", "observation_infos", "point_ids", "pose_ids", "camera_ids"):
        setattr(q, f, getattr(q, f)[keep])
    assert IO.solve(q, calib)["status"] == _lib.SBA_E_EMPTY


def test_returned_preintegrations_are_integrated_at_the_final_biases():
    p, calib, _ = IC.make_scenario(2, shape=IC.INIT_SHAPE)
    p.reintegration_thresh = 1e-4
    r = IO.solve(p, calib)
    moved = 0
    for k in range(p.num_fixed_key_frames, p.w_from_imu.shape[0] - 1):
        rec = r["preintegration"][k]
        f, n = rec["first_measurement"], rec["num_measurements"]
        if np.linalg.norm(r["gyro_bias"][k] - rec["gyro_bias"]) > 1e-4 or \
                np.linalg.norm(r["acc_bias"][k] - rec["acc_bias"]) > 1e-4:
            raise AssertionError(f"key frame {k}: biases moved past the threshold without a re-integration")
        if np.abs(rec["gyro_bias"]).max() > 0 or np.abs(rec["acc_bias"]).max() > 0:
            moved += 1
            want = IO.integrate(calib, p.measurements[f:f + n], rec["gyro_bias"], rec["acc_bias"])
            np.testing.assert_allclose(rec["dR"], want["dR"], atol=1e-7)
            np.testing.assert_allclose(rec["dP"], want["dP"], atol=1e-7)
    assert moved > 0
