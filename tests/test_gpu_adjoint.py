"""GPU implicit gradients: the device adjoint against the FP64 oracle at the same solved state, the device FP64
Newton refinement against the oracle's stationary point, and the end-to-end pixel / intrinsics gradients
against FP64 central finite differences of fully converged solves (tests/adjoint_fd.py)."""
import numpy as np
import pytest

import cuvslam_amd as ca
from adjoint_fd import FX, K0, case, chain_to_pixels, pose_loss_grad
from oracle import sba_oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b)


def solved(p, loss):
    # run to convergence: the reference's step-size exit (sqrt(FLT_EPSILON)) stops short of a stationary point
    solver = ca.SchurComplementBundlerGpu(loss=loss, fixed_iterations=True)
    b = solver.batch([p])
    b.run()
    q = p.copy()
    b.read([q])
    return q, solver, b


@pytest.mark.parametrize("loss", [ca.LOSS_HUBER, ca.LOSS_STUDENT_T])
def test_adjoint_matches_oracle_at_same_state(loss):
    p, _, gt = ca.scene.generate(seed=4, num_points=80, num_keyframes=5, num_fixed_keyframes=2, max_iterations=50)
    q, solver, b = solved(p, loss)
    ncf = p.rig_from_world.shape[0] - p.num_fixed_key_frames
    # the device's FP64 copy of the FP32 solution (free rotations projected onto the rotation group)
    poses, pts, _ = b.refine(newton_steps=0, gauss_newton_steps=0)
    assert np.max(np.abs(poses - q.rig_from_world)) < 1e-6 and np.array_equal(pts, q.points.astype(np.float64))
    g = np.concatenate([pose_loss_grad(poses, gt, ncf), np.zeros(3 * p.points.shape[0])])
    dxy, dinfo, it, a = b.adjoint(g, max_cg=300, tol=1e-12)
    assert 0 < it < 300
    ra, rdxy, rdinfo, rit = O.implicit_grad(p, g, pts, poses, loss=loss, tol=1e-12)
    assert rit > 0
    assert rel(a, ra) < 1e-6
    assert rel(dxy, rdxy) < 1e-6
    assert rel(dinfo, rdinfo) < 1e-6


def test_refine_reaches_the_fp64_stationary_point():
    c = case()
    q, solver, b = solved(c.problem, c.loss)
    g_init = np.linalg.norm(O.grad(c.problem, c.init_pts, c.init_poses, loss=c.loss))
    g_fp32 = np.linalg.norm(O.grad(c.problem, q.points.astype(np.float64), q.rig_from_world.astype(np.float64),
                                   loss=c.loss))
    poses, pts, gn = b.refine(newton_steps=3, max_cg=300, tol=1e-12)
    # the FP32 solve stops at its rounding floor; FP64 Newton steps reach the FP64 one
    assert gn < 1e-6 * g_fp32 and gn < 1e-12 * g_init, (gn, g_fp32, g_init)
    g_ref = np.linalg.norm(O.grad(c.problem, pts, poses, loss=c.loss))
    assert g_ref < 1e-12 * g_init
    # the oracle's stationary point, up to the null space of the unobserved points
    opt_pts, opt_poses = c.optimum
    seen = np.bincount(c.problem.point_ids, minlength=pts.shape[0]) > 0
    assert np.max(np.abs(poses - opt_poses)) < 1e-10
    assert np.max(np.abs(pts - opt_pts)[seen]) < 1e-8
    # the batch's own FP32 results are untouched
    q2 = c.problem.copy()
    b.read([q2])
    np.testing.assert_array_equal(q2.rig_from_world, q.rig_from_world)


def test_pixel_and_intrinsics_gradients_vs_finite_differences():
    # 1e-4 relative, as the oracle's own adjoint is held to in the CPU tests
    c = case()
    q, solver, b = solved(c.problem, c.loss)
    poses, pts, gn = b.refine(newton_steps=3, max_cg=300, tol=1e-12)
    dxy, dinfo, it, _ = b.adjoint(c.g_of(poses), max_cg=300, tol=1e-12)
    duv, dK = chain_to_pixels(dxy, dinfo, c.uv, K0)
    fd_uv, fd_K = c.fd
    assert rel(c.pick(duv), fd_uv) < 1e-4
    for k in range(4):
        assert abs(dK[k] - fd_K[k]) <= 1e-4 * abs(fd_K[k]), (k, dK[k], fd_K[k])
    # and the oracle's adjoint at the device's refined state
    _, rdxy, rdinfo, _ = O.implicit_grad(c.problem, c.g_of(poses), pts, poses, loss=c.loss)
    assert rel(dxy, rdxy) < 1e-6 and rel(dinfo, rdinfo) < 1e-6


def test_torch_autograd_function_matches_adjoint():
    import torch
    from cuvslam_amd import autograd as AG
    c = case()
    pixels = torch.tensor(c.uv, requires_grad=True)
    K = torch.tensor(K0, requires_grad=True)
    solver = ca.SchurComplementBundlerGpu(loss=c.loss, fixed_iterations=True)
    poses, points = AG.solve(pixels, K, c.problem, solver, max_cg=300, tol=1e-12, newton_steps=3)
    G = torch.from_numpy(c.gt.astype(np.float64))
    Lt = ((poses[:c.ncf, :3, :] - G[:c.ncf, :3, :]) ** 2).sum()
    Lt.backward()
    assert torch.isfinite(pixels.grad).all() and torch.isfinite(K.grad).all()
    fd_uv, fd_K = c.fd
    assert rel(c.pick(pixels.grad.numpy()), fd_uv) < 1e-4
    np.testing.assert_allclose(K.grad.numpy(), fd_K, rtol=1e-4)
    # the same numbers through the C ABI directly
    q, solver2, b = solved(c.problem, c.loss)
    T, P, _ = b.refine(newton_steps=3, max_cg=300, tol=1e-12)
    np.testing.assert_allclose(poses.detach().numpy(), T, rtol=0, atol=1e-12)
    dxy, dinfo, it, _ = b.adjoint(c.g_of(T), max_cg=300, tol=1e-12)
    np.testing.assert_allclose(pixels.grad.numpy()[:, 0], dxy[:, 0] / FX, rtol=1e-9, atol=1e-15)


def test_c4_window_gradients_vs_finite_difference_fixture():
    # 20 key frames, 10,000 points: device solve, FP64 refinement and adjoint against the committed FP64 central
    # differences of converged oracle solves (tests/golden/make_c4_fd.py)
    import json
    import os
    from adjoint_fd import C4_SCENE, Case
    with open(os.path.join(os.path.dirname(__file__), "golden", "c4_fd.json")) as f:
        fx = json.load(f)
    c = Case(**C4_SCENE)
    assert c.problem.num_observations == fx["num_observations"]
    q, solver, b = solved(c.problem, c.loss)
    # the FP32 solve of this window stops at |grad F| ~ 2e-4 of its start, where the full Hessian is not yet
    # positive definite: Gauss-Newton steps first, then Newton steps
    poses, pts, gn = b.refine(gauss_newton_steps=8, newton_steps=3, max_cg=300, tol=1e-12)
    g0 = np.linalg.norm(O.grad(c.problem, c.init_pts, c.init_poses, loss=c.loss))
    assert gn < 1e-12 * g0, (gn, g0)
    assert abs(c.L(poses) - fx["loss_at_optimum"]) <= 1e-8 * fx["loss_at_optimum"], (c.L(poses), fx["loss_at_optimum"])
    dxy, dinfo, it, _ = b.adjoint(c.g_of(poses), max_cg=300, tol=1e-12)
    assert it < 300
    duv, dK = chain_to_pixels(dxy, dinfo, c.uv, K0)
    picked = np.array([duv[o, cc] for o, cc in fx["sampled"]])
    assert rel(picked, fx["fd_pixels"]) < 1e-4
    np.testing.assert_allclose(dK, fx["fd_intrinsics"], rtol=1e-4)
