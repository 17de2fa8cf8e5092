"""Writes tests/golden/lm_step_fixtures.json: small SBA problems and one LM step per damping, computed with
numpy/scipy only -- an independent restatement used to pin the oracle (not the oracle's own code).

For each problem: the prediction of every observation (camera_from_rig * rig_from_world * point, divided
by depth), the robust weights (Huber or Student-t), the dense Jacobian of the prediction with respect to
every free parameter by FP64 central differences (a left-multiplied twist Exp(d) of each free key frame,
via scipy.linalg.expm; an additive step of each free point), the dense Gauss-Newton system
H = J^T diag(w Lambda) J, g = J^T diag(w Lambda) r, the damped step (H + lambda diag(H)) d = g by
numpy.linalg.solve, the reduced camera system of that damped system, the trial cost at the stepped state,
the predicted-reduction numerators d^T H d and d^T diag(H) d, and the decrease of the quadratic model
F(x + d) ~ F - 2 g^T d + d^T H d of the cost, 2 g^T d - d^T H d (F = sum rho(r^T L r) with no 1/2, so its
gradient is -2 g and its Gauss-Newton Hessian 2 H).

Run:  python tests/golden/make_fixtures.py
"""
import json
import os

import numpy as np
from scipy.linalg import expm

HERE = os.path.dirname(os.path.abspath(__file__))


def hat6(tw):
    w, v = tw[:3], tw[3:]
    M = np.zeros((4, 4))
    M[:3, :3] = [[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]]
    M[:3, 3] = v
    return M


def rho(kind, s, d):
    if kind == 0:
        return (d + 1) * np.log1p(s / d)
    return np.where(s < d * d, 0.5 * s, d * np.sqrt(s) - 0.5 * d * d)


def drho(kind, s, d):
    if kind == 0:
        return (d + 1) / (d + s)
    return np.where(s < d * d, 0.5, 0.5 * d / np.sqrt(np.maximum(s, 1e-300)))


def make_problem(seed, n_poses, n_points, n_fixed_poses, n_fixed_points):
    rng = np.random.default_rng(seed)
    cams = [np.eye(4), np.eye(4)]
    cams[1][:3, :3] = expm(hat6(np.array([0.02, -0.03, 0.01, 0, 0, 0])))[:3, :3]
    cams[1][:3, 3] = [0.125, 0.01, -0.02]
    poses = []
    for _ in range(n_poses):
        tw = np.concatenate([rng.uniform(-0.1, 0.1, 3), rng.uniform(-0.5, 0.5, 3)])
        poses.append(expm(hat6(tw)))
    pts = np.stack([rng.uniform(-2, 2, n_points), rng.uniform(-2, 2, n_points), rng.uniform(3, 8, n_points)], 1)
    xy, info, pid, qid, cid = [], [], [], [], []
    for j in range(n_poses):
        for i in range(n_points):
            for c in range(2):
                pc = cams[c] @ poses[j] @ np.append(pts[i], 1)
                if pc[2] <= 0.5:
                    continue
                u = pc[:2] / pc[2] + rng.normal(0, 0.004, 2)
                if rng.uniform() < 0.08:
                    u = u + rng.uniform(-0.1, 0.1, 2)
                L = rng.uniform(0.5, 2.0) * np.eye(2) * 450.0 ** 2
                L[0, 1] = L[1, 0] = 0.1 * 450.0 ** 2
                xy.append(u)
                info.append(L)
                pid.append(i)
                qid.append(j)
                cid.append(c)
    # perturb the state away from the truth (free entries only)
    poses_p = [p.copy() for p in poses]
    for j in range(n_poses - n_fixed_poses):
        tw = np.concatenate([rng.normal(0, 0.01, 3), rng.normal(0, 0.03, 3)])
        poses_p[j] = expm(hat6(tw)) @ poses[j]
    pts_p = pts.copy()
    pts_p[: n_points - n_fixed_points] += rng.normal(0, 0.05, (n_points - n_fixed_points, 3))
    # the solver works on FP32 inputs: store and compute with FP32-representable values
    f32 = lambda a: np.asarray(a, np.float32).astype(np.float64)  # noqa: E731
    return dict(points=f32(pts_p), xy=f32(xy), info=f32(info), point_ids=np.array(pid), pose_ids=np.array(qid),
                camera_ids=np.array(cid), poses=f32(poses_p), cams=f32(cams), n_fixed_poses=n_fixed_poses,
                n_fixed_points=n_fixed_points)


def predictions(P, poses, pts):
    T = np.stack([P["cams"][c] @ poses[j] for j, c in zip(P["pose_ids"], P["camera_ids"])])
    ph = np.concatenate([pts[P["point_ids"]], np.ones((len(P["point_ids"]), 1))], 1)
    pc = np.einsum("nij,nj->ni", T, ph)[:, :3]
    return pc[:, :2] / pc[:, 2:3], pc[:, 2]


def state_after(P, x):
    """Applies the free-parameter step x = (d_pose (6 per free key frame), d_point (3 per free point))."""
    nc, npnt = len(P["poses"]), len(P["points"])
    ncf, npf = nc - P["n_fixed_poses"], npnt - P["n_fixed_points"]
    poses = [p.copy() for p in P["poses"]]
    for j in range(ncf):
        poses[j] = expm(hat6(x[6 * j:6 * j + 6])) @ poses[j]
    pts = P["points"].copy()
    pts[:npf] += x[6 * ncf:].reshape(npf, 3)
    return poses, pts


def cost(P, poses, pts, kind, d):
    pred, z = predictions(P, poses, pts)
    r = P["xy"] - pred
    s = np.einsum("ni,nij,nj->n", r, P["info"], r)
    keep = z > np.float32(0.1)
    return float(np.sum(rho(kind, s[keep], d)))


def lm_step(P, kind, d, lams):
    nc, npnt = len(P["poses"]), len(P["points"])
    ncf, npf = nc - P["n_fixed_poses"], npnt - P["n_fixed_points"]
    nx = 6 * ncf + 3 * npf
    pred0, _ = predictions(P, P["poses"], P["points"])
    r = P["xy"] - pred0
    s = np.einsum("ni,nij,nj->n", r, P["info"], r)
    w = drho(kind, s, d)
    J = np.zeros((len(r), 2, nx))
    h = 1e-6
    for k in range(nx):
        e = np.zeros(nx)
        e[k] = h
        pp, _ = predictions(P, *state_after(P, e))
        pm, _ = predictions(P, *state_after(P, -e))
        J[:, :, k] = (pp - pm) / (2 * h)
    WL = w[:, None, None] * P["info"]
    H = np.einsum("nai,nab,nbj->ij", J, WL, J)
    g = np.einsum("nai,nab,nb->i", J, WL, r)
    F0 = cost(P, P["poses"], P["points"], kind, d)
    out = []
    n = 6 * ncf
    for lam in lams:
        Hd = H + lam * np.diag(np.diag(H))
        x = np.linalg.solve(Hd, g)
        S = Hd[:n, :n] - Hd[:n, n:] @ np.linalg.solve(Hd[n:, n:], Hd[n:, :n])
        e_red = g[:n] - Hd[:n, n:] @ np.linalg.solve(Hd[n:, n:], g[n:])
        trial = cost(P, *state_after(P, x), kind, d)
        out.append(dict(lam=lam, dc=x[:n].tolist(), dp=x[n:].tolist(), S=S.tolist(), e=e_red.tolist(),
                        trial_cost=trial, hess=float(x @ H @ x), scal=float(x @ np.diag(np.diag(H)) @ x),
                        model_reduction=float(2 * x @ g - x @ H @ x)))
    return F0, out


def main():
    cases = []
    for seed, (nposes, npts, nfp, nfpt), kind in [(1, (3, 12, 1, 0), 1), (2, (4, 10, 2, 1), 0), (3, (3, 8, 1, 2), 1)]:
        P = make_problem(seed, nposes, npts, nfp, nfpt)
        F0, steps = lm_step(P, kind, 1.0, [1e-3, 1e-1, 1e3])
        cases.append(dict(
            seed=seed, loss=kind, delta=1.0, initial_cost=F0, steps=steps,
            points=P["points"].tolist(), xy=P["xy"].tolist(), info=P["info"].tolist(),
            point_ids=P["point_ids"].tolist(), pose_ids=P["pose_ids"].tolist(), camera_ids=P["camera_ids"].tolist(),
            rig_from_world=[p.tolist() for p in P["poses"]], camera_from_rig=[c.tolist() for c in P["cams"]],
            num_fixed_key_frames=P["n_fixed_poses"], num_fixed_points=P["n_fixed_points"]))
    with open(os.path.join(HERE, "lm_step_fixtures.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_fixtures.py", "cases": cases}, f)
    print(f"wrote {len(cases)} cases")


if __name__ == "__main__":
    main()
