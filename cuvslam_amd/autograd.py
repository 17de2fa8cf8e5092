"""Differentiable SBA: gradients of a loss of the converged solution w.r.t. the observed pixels and the
pinhole intrinsics (fx, fy, cx, cy).

Forward: the observations are normalized (x = (u - c_x) / f_x, y = (v - c_y) / f_y) and their information mapped
to normalized coordinates (L_xy = diag(f) L_uv diag(f), as libs/camera/observation.cpp:57-67 does), then the
window is solved on the GPU and its state polished by FP64 Newton steps on the device (``sba_batch_refine``: the
FP32 LM solve stops at its rounding floor, where the residual gradient would bias the implicit gradient). Backward: the solver's implicit adjoint (``sba_batch_adjoint``: full FP64 Hessian of
the robust cost at the solution, conjugate gradients preconditioned by the factored Gauss-Newton system) gives
dL/dxy and dL/dL_xy per observation, mapped here to pixels and intrinsics by the chain rule above.
"""
from __future__ import annotations

import numpy as np
import torch

from .sba import BundleAdjustmentProblem, SchurComplementBundlerGpu


def observations_from_pixels(uv: np.ndarray, K: np.ndarray, info_uv: np.ndarray):
    fx, fy, cx, cy = K
    xy = np.stack([(uv[:, 0] - cx) / fx, (uv[:, 1] - cy) / fy], 1)
    D = np.array([fx, fy])
    info = info_uv * D[None, :, None] * D[None, None, :]
    return xy, info


def pose_tangent_grad(poses: np.ndarray, dposes: np.ndarray) -> np.ndarray:
    """dL/d(left twist) of each pose from the Euclidean gradient dL/dT of its 4x4 matrix:
    d(Exp(d) T)/d omega_k = [e_k]x T (rows 0-2), d/dv_k = e_k in the translation column."""
    out = np.zeros((poses.shape[0], 6))
    for j in range(poses.shape[0]):
        M = dposes[j, :3, :]
        A = M @ poses[j, :3, :].T
        out[j, :3] = [A[2, 1] - A[1, 2], A[0, 2] - A[2, 0], A[1, 0] - A[0, 1]]
        out[j, 3:] = M[:, 3]
    return out


class SbaSolveFunction(torch.autograd.Function):
    """(pixels (N_o, 2), intrinsics (4,)) -> (rig_from_world (N_c, 4, 4), points (N_p, 3)) of the solved window."""

    @staticmethod
    def forward(ctx, pixels, intrinsics, template: BundleAdjustmentProblem, solver: SchurComplementBundlerGpu,
                info_uv: np.ndarray, max_cg: int, tol: float, newton_steps: int):
        uv = pixels.detach().cpu().double().numpy()
        K = intrinsics.detach().cpu().double().numpy()
        xy, info = observations_from_pixels(uv, K, info_uv)
        prob = template.copy()
        prob.observation_xys = xy.astype(np.float32)
        prob.observation_infos = info.astype(np.float32)
        batch = solver.batch([prob])
        batch.run()
        poses, points, _ = batch.refine(newton_steps, max_cg, tol)
        ctx.batch, ctx.prob, ctx.K, ctx.xy, ctx.info, ctx.poses = batch, prob, K, xy, info, poses
        ctx.max_cg, ctx.tol = max_cg, tol
        dev = pixels.device
        return torch.from_numpy(poses).to(dev), torch.from_numpy(points).to(dev)

    @staticmethod
    def backward(ctx, dposes, dpoints):
        prob, K, xy, info = ctx.prob, ctx.K, ctx.xy, ctx.info
        ncf = prob.rig_from_world.shape[0] - prob.num_fixed_key_frames
        npf = prob.points.shape[0] - prob.num_fixed_points
        gp = np.zeros((ncf, 6)) if dposes is None else pose_tangent_grad(
            ctx.poses[:ncf], dposes.detach().cpu().double().numpy()[:ncf])
        gq = np.zeros((npf, 3)) if dpoints is None else dpoints.detach().cpu().double().numpy()[:npf]
        g = np.concatenate([gp.ravel(), gq.ravel()])
        dxy, dinfo, it, _ = ctx.batch.adjoint(g, ctx.max_cg, ctx.tol)
        ctx.batch.close()
        fx, fy, cx, cy = K
        duv = np.stack([dxy[:, 0] / fx, dxy[:, 1] / fy], 1)
        dK = np.array([
            np.sum(dxy[:, 0] * (-xy[:, 0] / fx) + dinfo[:, 0] * 2 * info[:, 0, 0] / fx + dinfo[:, 1] * info[:, 0, 1] / fx),
            np.sum(dxy[:, 1] * (-xy[:, 1] / fy) + dinfo[:, 2] * 2 * info[:, 1, 1] / fy + dinfo[:, 1] * info[:, 0, 1] / fy),
            np.sum(-dxy[:, 0] / fx),
            np.sum(-dxy[:, 1] / fy)])
        dev = dposes.device if dposes is not None else dpoints.device
        return (torch.from_numpy(duv).to(dev), torch.from_numpy(dK).to(dev), None, None, None, None, None, None)


def solve(pixels: torch.Tensor, intrinsics: torch.Tensor, template: BundleAdjustmentProblem,
          solver: SchurComplementBundlerGpu, info_uv=None, max_cg: int = 200, tol: float = 1e-12,
          newton_steps: int = 3):
    """Differentiable solve: returns (rig_from_world, points) as float64 tensors, the FP32 LM solution refined by
    ``newton_steps`` FP64 Newton steps. The implicit gradient assumes a stationary solution: give the solver
    enough iterations with fixed_iterations=True (the reference's step-size exit at sqrt(FLT_EPSILON) stops
    short of one) so that the Newton steps start inside their region of quadratic convergence."""
    if info_uv is None:
        info_uv = np.broadcast_to(np.eye(2), (pixels.shape[0], 2, 2)).copy()
    return SbaSolveFunction.apply(pixels, intrinsics, template, solver, np.asarray(info_uv, np.float64), max_cg, tol,
                                 newton_steps)
