"""Pins of the CPU oracle (runs without a GPU).

Each function of oracle/sba_oracle.cpp is checked against something it does not compute itself:
  * golden fixtures from tests/golden/make_fixtures.py (numpy/scipy: finite-difference Jacobians and a dense
    solve of the damped normal equations) -- linearization, full system, Schur complement, point-block
    inverse, Cholesky, back-substitution, damping, trial cost, predicted-reduction terms;
  * the reference's own known answers (libs/cuda_modules/test/sba_test.cpp: the 3x3 Cholesky system, the
    near-plane bands, convergence on exact observations);
  * closed forms and library routines (scipy.linalg.expm, numpy SVD / pinv / cholesky).
"""
import json
import os

import numpy as np
import pytest
from scipy.linalg import expm

import cuvslam_amd as ca
from lm_cases import reference_style_problem
from oracle import sba_oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


def fixture_problems():
    d = json.load(open(os.path.join(HERE, "golden", "lm_step_fixtures.json")))
    out = []
    for c in d["cases"]:
        p = ca.BundleAdjustmentProblem(
            np.array(c["points"], np.float32), np.array(c["xy"], np.float32), np.array(c["info"], np.float32),
            np.array(c["point_ids"], np.int32), np.array(c["pose_ids"], np.int32), np.array(c["camera_ids"], np.int8),
            np.array(c["rig_from_world"], np.float32), np.array(c["camera_from_rig"], np.float32),
            num_fixed_points=c["num_fixed_points"], num_fixed_key_frames=c["num_fixed_key_frames"],
            max_iterations=10, robustifier_scale=c["delta"])
        out.append((p, c))
    return out


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("case", range(3))
def test_step_matches_dense_fixture(case):
    p, c = fixture_problems()[case]
    for st in c["steps"]:
        S, e, dc, dp, sc, rc = O.step(p, st["lam"], loss=c["loss"])
        assert rc == 0
        # limited by the finite-difference Jacobians of the fixture (central differences, h = 1e-6)
        assert rel(S, st["S"]) < 1e-6
        assert rel(e, st["e"]) < 1e-6
        assert rel(dc, st["dc"]) < 1e-5
        assert rel(dp, st["dp"]) < 1e-5
        assert abs(sc[0] - st["trial_cost"]) / st["trial_cost"] < 1e-7
        assert abs(sc[2] - st["hess"]) / st["hess"] < 1e-6
        assert abs(sc[3] - st["scal"]) / st["scal"] < 1e-5
        assert abs(sc[6] - c["initial_cost"]) / c["initial_cost"] < 1e-12


def test_dense_schur_equals_block_sparse():
    # the reference CPU path's dense temp * point_pose_block product (cpu.cpp:264-277) against the block form
    p, c = fixture_problems()[1]
    a = O.step(p, 1e-2, loss=c["loss"], dense_schur=1)
    b = O.step(p, 1e-2, loss=c["loss"], dense_schur=0)
    for x, y in zip(a[:4], b[:4]):
        assert rel(x, y) < 1e-12


def test_pseudo_inverse_solver_matches_cholesky_on_pd_system():
    # cpu.cpp:337-346 (Jacobi pseudo-inverse, threshold 1e-6) and the GPU bundler's Cholesky agree on a
    # well-conditioned reduced system
    p, c = fixture_problems()[0]
    a = O.step(p, 1e-2, loss=c["loss"], rcs_solver=0)
    b = O.step(p, 1e-2, loss=c["loss"], rcs_solver=1)
    assert rel(a[2], b[2]) < 1e-9


def test_linearize_jacobians_vs_central_differences():
    p, _ = fixture_problems()[0]
    r, Jc, Jp, w = O.linearize(p, loss=1)
    pts = p.points.astype(np.float64)
    poses = p.rig_from_world.astype(np.float64)
    cams = p.camera_from_rig.astype(np.float64)

    def hat(tw):
        M = np.zeros((4, 4))
        M[:3, :3] = [[0, -tw[2], tw[1]], [tw[2], 0, -tw[0]], [-tw[1], tw[0], 0]]
        M[:3, 3] = tw[3:]
        return M

    for o in range(0, p.num_observations, 7):
        i, j, c = p.point_ids[o], p.pose_ids[o], p.camera_ids[o]

        def proj(T, X):
            q = cams[c] @ T @ np.append(X, 1.0)
            return q[:2] / q[2]

        h = 1e-6
        num_c = np.zeros((2, 6))
        for k in range(6):
            e = np.zeros(6)
            e[k] = h
            num_c[:, k] = (proj(expm(hat(e)) @ poses[j], pts[i]) - proj(expm(hat(-e)) @ poses[j], pts[i])) / (2 * h)
        num_p = np.zeros((2, 3))
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            num_p[:, k] = (proj(poses[j], pts[i] + e) - proj(poses[j], pts[i] - e)) / (2 * h)
        np.testing.assert_allclose(Jc[o], num_c, rtol=1e-6, atol=1e-8)
        np.testing.assert_allclose(Jp[o], num_p, rtol=1e-6, atol=1e-8)
        np.testing.assert_allclose(r[o], p.observation_xys[o].astype(np.float64) - proj(poses[j], pts[i]), atol=1e-12)


@pytest.mark.parametrize("kind", [0, 1])
def test_robust_loss_closed_forms(kind):
    d = 0.7
    for s in [0.0, 0.1, 0.48, 0.49, 0.5, 2.0, 30.0]:
        rho = O.loss(kind, s, d)
        if kind == 0:
            assert rho == pytest.approx((d + 1) * np.log(1 + s / d), rel=1e-14)
            assert O.loss(kind, s, d, True) == pytest.approx((d + 1) / (d + s), rel=1e-14)
        else:
            want = 0.5 * s if s < d * d else d * np.sqrt(s) - 0.5 * d * d
            assert rho == pytest.approx(want, rel=1e-14, abs=1e-300)
        if s > 0.05 and abs(s - d * d) > 1e-3:  # rho' is the derivative of rho (away from the Huber knee)
            h = 1e-6
            fd = (O.loss(kind, s + h, d) - O.loss(kind, s - h, d)) / (2 * h)
            assert O.loss(kind, s, d, True) == pytest.approx(fd, rel=1e-7)
    # Huber is continuous at the knee s = d^2 (robust_cost_function.cpp:62-75)
    assert O.loss(1, d * d * (1 - 1e-12), d) == pytest.approx(O.loss(1, d * d, d), rel=1e-9)


def test_exp_matches_matrix_exponential():
    rng = np.random.default_rng(0)
    for _ in range(20):
        tw = np.concatenate([rng.normal(0, 0.8, 3), rng.normal(0, 2, 3)])
        M = np.zeros((4, 4))
        M[:3, :3] = [[0, -tw[2], tw[1]], [tw[2], 0, -tw[0]], [-tw[1], tw[0], 0]]
        M[:3, 3] = tw[3:]
        np.testing.assert_allclose(O.exp(tw), expm(M), atol=1e-12)
    np.testing.assert_allclose(O.exp(np.zeros(6)), np.eye(4), atol=1e-15)


def test_nearest_rotation_is_svd_polar_factor():
    rng = np.random.default_rng(1)
    for _ in range(20):
        R0 = expm(np.array([[0, -0.3, 0.2], [0.3, 0, -0.1], [-0.2, 0.1, 0]]) * rng.normal())
        M = R0 + rng.normal(0, 1e-3, (3, 3))
        U, _, Vt = np.linalg.svd(M)
        want = U @ Vt
        if np.linalg.det(want) < 0:
            want = want - 2 * np.outer(U[:, 2], Vt[2])
        R = O.nearest_rotation(M)
        np.testing.assert_allclose(R, want, atol=1e-12)
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-12)
        assert np.linalg.det(R) == pytest.approx(1.0, abs=1e-12)


def test_cholesky_known_answer_and_failure():
    # libs/cuda_modules/test/sba_test.cpp:700-719
    A = np.array([[4, 1, 0], [1, 3, 1], [0, 1, 2]], float)
    x, rc = O.chol_solve(A, np.array([6, 10, 8], float))
    assert rc == 0
    np.testing.assert_allclose(x, [1, 2, 3], atol=1e-14)
    _, rc = O.chol_solve(np.array([[1, 2, 0], [2, 1, 0], [0, 0, 1]], float), np.array([6, 10, 8], float))
    assert rc == 1
    rng = np.random.default_rng(2)
    M = rng.standard_normal((30, 30))
    A = M @ M.T + 30 * np.eye(30)
    b = rng.standard_normal(30)
    x, rc = O.chol_solve(A, b)
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-12)


def test_point_pseudo_inverse_rank_rule():
    rng = np.random.default_rng(3)
    Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    full = Q @ np.diag([5.0, 2.0, 0.5]) @ Q.T
    inv, rank = O.point_pinv(full)
    assert rank == 3
    np.testing.assert_allclose(inv, np.linalg.inv(full), rtol=1e-12)
    # rank 2: the smallest singular value is below 1e-6 of the largest -> its direction is dropped
    two = Q @ np.diag([5.0, 2.0, 1e-9]) @ Q.T
    inv, rank = O.point_pinv(two)
    assert rank == 2
    np.testing.assert_allclose(inv, np.linalg.pinv(two, rcond=1e-6), rtol=1e-9, atol=1e-12)
    # rank 1 -> zero (schur_complement_bundler_cpu.cpp:242-254)
    one = Q @ np.diag([5.0, 1e-9, 1e-10]) @ Q.T
    inv, rank = O.point_pinv(one)
    assert rank == 1
    assert np.all(inv == 0)


def test_cost_zero_at_exact_observations_and_infinite_inside_near_plane():
    p, _, _ = ca.scene.generate(seed=4, num_points=100, num_keyframes=4, pixel_noise=0.0, outlier_rate=0.0,
                                pose_sigma_m=0.0, pose_sigma_deg=0.0, point_sigma_m=0.0)
    c, sk = O.cost(p, loss=1)
    assert sk == 0 and c < 1e-6 * p.num_observations  # FP32 storage of the exact projections
    # sba_test.cpp:1402-1462: every point between 0.02 and 0.08 m in front of the cameras -> all skipped, +inf
    q = p.copy()
    q.rig_from_world[:] = np.eye(4, dtype=np.float32)
    q.points[:, 2] = np.linspace(0.02, 0.08, q.points.shape[0], dtype=np.float32)
    q.points[:, :2] *= 0.001
    c, sk = O.cost(q, loss=1)
    assert sk == q.num_observations and np.isinf(c)
    r = O.solve(q, "f64", loss=1)
    assert r["status"] == 2
    # 0.3-0.8 m is in front of the near plane: kept (sba_test.cpp:1361-1396)
    q.points[:, 2] = np.linspace(0.3, 0.8, q.points.shape[0], dtype=np.float32)
    c, sk = O.cost(q, loss=1)
    assert sk == 0 and np.isfinite(c)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_converges_on_exact_observations(precision):
    # sba_test.cpp:1322-1355: exact observations, free key frames shifted by (0.01, -0.01, 0.01), 50 iterations
    p = reference_style_problem(1)
    p.rig_from_world[:-1, :3, 3] += np.array([0.01, -0.01, 0.01], np.float32)
    kw = dict(loss=0) if precision == "f64" else dict(loss=0, rcs_solver=1, dense_schur=1)
    r = O.solve(p, precision, **kw)
    assert r["status"] == 0
    assert np.isfinite(r["initial_cost"]) and np.isfinite(r["last_cost"])
    assert r["last_cost"] < r["initial_cost"]
    assert r["iterations"] < 50


def test_initial_cost_below_epsilon_returns_immediately():
    p, _, _ = ca.scene.generate(seed=5, num_points=60, num_keyframes=3, pixel_noise=0.0, outlier_rate=0.0,
                                pose_sigma_m=0.0, pose_sigma_deg=0.0, point_sigma_m=0.0)
    q = p.copy()
    q.observation_infos *= 1e-12  # exact projections; information scaled so the cost is below FLT_EPSILON
    r = O.solve(q, "f64", loss=1)
    assert r["status"] == 0 and r["iterations"] == 0


def test_fp32_reference_path_tracks_fp64_oracle():
    p, _, _ = ca.scene.generate("c1", seed=1, num_fixed_keyframes=2, num_points=600, max_iterations=10)
    a = O.solve(p, "f64", loss=1)
    b = O.solve(p, "f32", loss=1, rcs_solver=1, dense_schur=1)
    q = p.copy()
    q.points = b["points"].astype(np.float32)
    q.rig_from_world = b["poses"].astype(np.float32)
    c, _ = O.cost(q, loss=1)
    assert abs(c - a["last_cost"]) / a["last_cost"] < 1e-6


def _retract(poses, pts, x, ncf, npf):
    poses = poses.copy()
    pts = pts.copy()
    for j in range(ncf):
        poses[j] = O.exp(x[6 * j:6 * j + 6]) @ poses[j]
    pts[:npf] += x[6 * ncf:].reshape(-1, 3)
    return poses, pts


@pytest.mark.parametrize("loss", [ca.LOSS_HUBER, ca.LOSS_STUDENT_T])
def test_cost_gradient_vs_central_differences(loss):
    # grad F in the solver's tangent coordinates against central differences of the FP64 cost with the
    # update applied the way the LM step applies it (left twist of each free key frame, point offset)
    p, _, _ = ca.scene.generate(seed=3, num_points=40, num_keyframes=4, num_fixed_keyframes=2, pixel_noise=2.0)
    ncf = p.rig_from_world.shape[0] - p.num_fixed_key_frames
    npf = p.points.shape[0] - p.num_fixed_points
    g = O.grad(p, p.points.astype(np.float64), p.rig_from_world.astype(np.float64), loss=loss)
    fd = np.zeros_like(g)
    h = 1e-6
    for k in range(g.size):
        e = np.zeros(g.size)
        e[k] = h
        cp, _ = O.cost(p, dc=e[:6 * ncf], dp=e[6 * ncf:], loss=loss)
        cm, _ = O.cost(p, dc=-e[:6 * ncf], dp=-e[6 * ncf:], loss=loss)
        fd[k] = (cp - cm) / (2 * h)
    assert rel(g, fd) < 1e-6


def test_hessian_product_vs_central_differences_of_gradient():
    # the full Hessian (rho'' and the second-order projection / twist terms) against central differences of
    # grad F along a random tangent direction v, from a state off the optimum so that every term contributes.
    # grad F at Exp(h v) T is taken in coordinates centred there: by Baker-Campbell-Hausdorff,
    # log(Exp(e) Exp(h v)) = h v + (I - (h/2) ad_v) e, so the difference quotient of the key-frame rows is
    # H v - (1/2) ad_v^T g with ad_(w,u) = [[w]x, 0; [u]x, [w]x]; the point rows are plain H v.
    loss = ca.LOSS_STUDENT_T
    p, _, _ = ca.scene.generate(seed=3, num_points=40, num_keyframes=4, num_fixed_keyframes=2, pixel_noise=2.0)
    ncf = p.rig_from_world.shape[0] - p.num_fixed_key_frames
    npf = p.points.shape[0] - p.num_fixed_points
    P0, T0 = p.points.astype(np.float64), p.rig_from_world.astype(np.float64)
    v = np.random.default_rng(0).standard_normal(6 * ncf + 3 * npf)
    Hv = O.hvp(p, v, P0, T0, loss=loss)
    g = O.grad(p, P0, T0, loss=loss)
    h = 1e-6
    Tp, Pp = _retract(T0, P0, h * v, ncf, npf)
    Tm, Pm = _retract(T0, P0, -h * v, ncf, npf)
    fd = (O.grad(p, Pp, Tp, loss=loss) - O.grad(p, Pm, Tm, loss=loss)) / (2 * h)

    def sk(a):
        return np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])

    expect = Hv.copy()
    for j in range(ncf):
        w, u = v[6 * j:6 * j + 3], v[6 * j + 3:6 * j + 6]
        ad = np.zeros((6, 6))
        ad[:3, :3], ad[3:, :3], ad[3:, 3:] = sk(w), sk(u), sk(w)
        expect[6 * j:6 * j + 6] -= 0.5 * ad.T @ g[6 * j:6 * j + 6]
    assert rel(expect[6 * ncf:], fd[6 * ncf:]) < 1e-6
    assert rel(expect[:6 * ncf], fd[:6 * ncf]) < 1e-6
    # without the second-order terms the products would be off by far more than that
    assert rel(Hv[:6 * ncf], fd[:6 * ncf]) > 1e-3


def test_implicit_gradient_vs_finite_differences_of_converged_solves():
    # dL/d pixels and dL/d intrinsics from the oracle's adjoint at the FP64 stationary point, against central
    # differences of fully converged FP64 solves (tests/adjoint_fd.py)
    from adjoint_fd import K0, case, chain_to_pixels, obs
    c = case()
    pts, poses = c.optimum
    xy, info = obs(c.uv, K0)
    g0 = O.grad(c.problem, c.init_pts, c.init_poses, xy=xy, info=info, loss=c.loss)
    assert np.linalg.norm(O.grad(c.problem, pts, poses, xy=xy, info=info, loss=c.loss)) < 1e-9 * np.linalg.norm(g0)
    a, dxy, dinfo, it = O.implicit_grad(c.problem, c.g_of(poses), pts, poses, xy=xy, info=info, loss=c.loss)
    assert it > 0
    duv, dK = chain_to_pixels(dxy, dinfo, c.uv, K0)
    fd_uv, fd_K = c.fd
    assert rel(c.pick(duv), fd_uv) < 1e-4
    for k in range(4):
        assert abs(dK[k] - fd_K[k]) <= 1e-4 * abs(fd_K[k]), (k, dK[k], fd_K[k])


# ---- landmark winsorizing (pipelines/winsorizer.cpp:23-73) ----

def _reproj_sq(problem, kf):
    """Squared reprojection errors of key frame kf's observations from plain matrix products: (error, point)."""
    sel = np.nonzero(problem.pose_ids == kf)[0]
    T = problem.rig_from_world[kf].astype(np.float64)
    Cs = problem.camera_from_rig.astype(np.float64)
    X = np.c_[problem.points[problem.point_ids[sel]].astype(np.float64), np.ones(sel.size)]
    Pc = np.einsum("nij,nj->ni", Cs[problem.camera_ids[sel]], (T @ X.T).T)
    keep = Pc[:, 2] > 0.1
    proj = Pc[keep, :2] / Pc[keep, 2:3]
    e = np.sum((proj - problem.observation_xys[sel[keep]].astype(np.float64)) ** 2, axis=1)
    return e, problem.point_ids[sel[keep]], sel[keep]


def test_winsorize_matches_a_sort_of_the_errors():
    # the 5% largest of n errors, int(5 n / 100) of them, reset their points when above 0.0025^2
    p, _, _ = ca.scene.generate(seed=3, num_points=1500, num_keyframes=8, num_fixed_keyframes=2)
    for kf in range(3):
        e, pt, ob = _reproj_sq(p, kf)
        order = np.lexsort((ob, pt, e))  # ascending by error, then point, then observation
        cut = int(np.float32(5.0) * np.float32(e.size) / np.float32(100.0))
        top = order[e.size - cut:]
        want = np.zeros(p.points.shape[0], bool)
        want[pt[top[e[top] > np.float64(np.float32(0.0025) * np.float32(0.0025))]]] = True
        mask, n = O.winsorize(p, kf)
        assert e.size >= 50 and cut >= 1
        np.testing.assert_array_equal(mask, want)
        assert n == want.sum()


def test_winsorize_keeps_exact_observations_and_catches_planted_outliers():
    p, pts, poses = ca.scene.generate(seed=4, num_points=1200, num_keyframes=6, num_fixed_keyframes=2,
                                      outlier_rate=0.0, pixel_noise=0.0)
    # at the true state every error is ~0, below the minimum size: nothing is reset
    mask, n = O.winsorize(p, 0, pts, poses)
    assert n == 0 and not mask.any()
    # plant exactly int(5 n / 100) gross errors among key frame 0's observations: exactly their points go
    sel = np.nonzero(p.pose_ids == 0)[0]
    e, _, _ = _reproj_sq(p, 0)
    cut = int(np.float32(5.0) * np.float32(e.size) / np.float32(100.0))
    first = {}
    for o in sel:
        first.setdefault(int(p.point_ids[o]), o)
    planted = np.array(list(first.values())[:cut])
    q = p.copy()
    q.observation_xys[planted] += np.float32(0.05)
    mask, n = O.winsorize(q, 0, pts, poses)
    assert n == cut
    np.testing.assert_array_equal(np.nonzero(mask)[0], np.sort(q.point_ids[planted]))


def test_winsorize_needs_fifty_observations():
    p, _, _ = ca.scene.generate(seed=5, num_points=40, num_keyframes=4, num_fixed_keyframes=2)
    e, _, _ = _reproj_sq(p, 0)
    assert e.size < 50
    mask, n = O.winsorize(p, 0)
    assert n == 0 and not mask.any()


def test_oracle_is_clean_under_address_and_undefined_behaviour_sanitizers():
    # every oracle entry point on generated windows, built with -fsanitize=address,undefined (oracle/asan_driver.cpp)
    import subprocess
    here = os.path.join(os.path.dirname(HERE), "oracle")
    subprocess.run(["make", "-s", "-C", here, "asan_driver"], check=True)
    out = subprocess.run([os.path.join(here, "asan_driver")], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0 and out.stdout.strip() == "OK", out.stdout + out.stderr
