"""Finite-difference references for the implicit gradients of a converged window (shared by the oracle and the
GPU tests). The loss is L = sum_j |T_j - G_j|^2 over the free key frames' [R | t] rows against the ground truth
G; the parameters are the observed pixels (u, v) and the pinhole intrinsics (fx, fy, cx, cy), mapped to the
solver's normalized observations by x = (u - c_x) / f_x, L_xy = diag(f) L_uv diag(f) with L_uv = I.

Each FD solve is the FP64 oracle run to convergence from the same initial state, then two full-Hessian Newton
steps (the oracle's own implicit-gradient conjugate gradients on grad F), so that the FD quotient sees the
stationary point rather than where LM stopped."""
from __future__ import annotations

import functools

import numpy as np

import cuvslam_amd as ca
from oracle import sba_oracle as O

FX, FY, CX, CY = 450.0, 450.0, 376.0, 240.0
K0 = np.array([FX, FY, CX, CY])


def pose_loss_grad(poses, gt, ncf):
    """dL/d(left twist) of L = sum_j |T_j - G_j|^2: d(Exp(d) T)/d omega_k = [e_k]x T, d/dv_k = e_k."""
    g = np.zeros(6 * ncf)
    for j in range(ncf):
        M = 2 * (poses[j, :3, :] - gt[j, :3, :])
        A = M @ poses[j, :3, :].T
        g[6 * j:6 * j + 3] = [A[2, 1] - A[1, 2], A[0, 2] - A[2, 0], A[1, 0] - A[0, 1]]
        g[6 * j + 3:6 * j + 6] = M[:, 3]
    return g


def obs(uv, K):
    fx, fy, cx, cy = K
    xy = np.stack([(uv[:, 0] - cx) / fx, (uv[:, 1] - cy) / fy], 1)
    return xy, np.broadcast_to(np.diag([fx * fx, fy * fy]), (uv.shape[0], 2, 2)).copy()


def chain_to_pixels(dxy, dinfo, uv, K):
    """(dL/dxy, dL/dinfo) -> (dL/duv (N_o, 2), dL/dK (4,)) for L_uv = I."""
    fx, fy = K[0], K[1]
    xy, info = obs(uv, K)
    duv = np.stack([dxy[:, 0] / fx, dxy[:, 1] / fy], 1)
    dK = np.array([
        np.sum(dxy[:, 0] * (-xy[:, 0] / fx) + dinfo[:, 0] * 2 * info[:, 0, 0] / fx),
        np.sum(dxy[:, 1] * (-xy[:, 1] / fy) + dinfo[:, 2] * 2 * info[:, 1, 1] / fy),
        np.sum(-dxy[:, 0] / fx), np.sum(-dxy[:, 1] / fy)])
    return duv, dK


class Case:
    """A window (by default 80 points, 5 key frames of which 2 fixed) with Student-t loss (smooth, so central
    differences are meaningful for every observation)."""

    def __init__(self, loss=ca.LOSS_STUDENT_T, seed=4, num_points=80, num_keyframes=5, max_iterations=50, **scene):
        self.loss = loss
        self.max_iterations = max_iterations
        self.problem, _, self.gt = ca.scene.generate(seed=seed, num_points=num_points, num_keyframes=num_keyframes,
                                                     num_fixed_keyframes=2, max_iterations=max_iterations, **scene)
        p = self.problem
        self.ncf = p.rig_from_world.shape[0] - p.num_fixed_key_frames
        self.npf = p.points.shape[0] - p.num_fixed_points
        xy0 = p.observation_xys.astype(np.float64)
        self.uv = np.stack([xy0[:, 0] * FX + CX, xy0[:, 1] * FY + CY], 1)
        self.init_pts = p.points.astype(np.float64)
        self.init_poses = p.rig_from_world.astype(np.float64)

    def converged(self, uv, K, newton=2):
        p = self.problem
        xy, info = obs(uv, K)
        r = O.solve_state(p, self.init_pts, self.init_poses, xy=xy, info=info, loss=self.loss, fixed_iterations=1)
        pts, poses = r["points"], r["poses"]
        for _ in range(newton):
            gg = O.grad(p, pts, poses, xy=xy, info=info, loss=self.loss)
            a, _, _, _ = O.implicit_grad(p, gg, pts, poses, xy=xy, info=info, loss=self.loss)
            poses = poses.copy()
            pts = pts.copy()
            for j in range(self.ncf):
                poses[j] = O.exp(-a[6 * j:6 * j + 6]) @ poses[j]
            pts[:self.npf] -= a[6 * self.ncf:].reshape(-1, 3)
        return pts, poses

    def L(self, poses):
        return float(np.sum((poses[:self.ncf, :3, :] - self.gt[:self.ncf, :3, :]) ** 2))

    def g_of(self, poses):
        return np.concatenate([pose_loss_grad(poses, self.gt, self.ncf), np.zeros(3 * self.npf)])

    @functools.cached_property
    def optimum(self):
        return self.converged(self.uv, K0)

    @functools.cached_property
    def sampled(self):
        """Observation components (o, c) sampled for the pixel differences."""
        rng = np.random.default_rng(1)
        return [(int(o), c) for o in rng.choice(self.problem.num_observations, 5, replace=False) for c in range(2)]

    @functools.cached_property
    def fd(self):
        """Central differences: dL/du at the sampled components (1e-3 px), dL/dK (1e-4)."""
        duv = []
        for o, c in self.sampled:
            up, um = self.uv.copy(), self.uv.copy()
            up[o, c] += 1e-3
            um[o, c] -= 1e-3
            duv.append((self.L(self.converged(up, K0)[1]) - self.L(self.converged(um, K0)[1])) / 2e-3)
        dK = []
        for k in range(4):
            Kp, Km = K0.copy(), K0.copy()
            Kp[k] += 1e-4
            Km[k] -= 1e-4
            dK.append((self.L(self.converged(self.uv, Kp)[1]) - self.L(self.converged(self.uv, Km)[1])) / 2e-4)
        return np.array(duv), np.array(dK)

    def pick(self, duv):
        return np.array([duv[o, c] for o, c in self.sampled])


@functools.lru_cache(maxsize=None)
def case(loss=ca.LOSS_STUDENT_T, seed=4):
    return Case(loss, seed)


# The differentiation workload (20 key frames, 10,000 points). Student-t is smooth, so central differences see the
# derivative itself. No outliers: with outliers the robust cost is not strictly convex at the solution (Student-t
# curvature turns negative past s = delta; a Huber point whose observations all lie in the linear region has a flat
# valley of optima), the solution is not a differentiable function of the observations and the implicit system is
# singular.
C4_SCENE = dict(loss=ca.LOSS_STUDENT_T, seed=4, num_points=10000, num_keyframes=20, max_iterations=50, outlier_rate=0.0)
