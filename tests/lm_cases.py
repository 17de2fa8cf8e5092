"""Windows that drive every branch of the Levenberg-Marquardt control loop
(schur_complement_bundler_gpu.cpp:175-253, schur_complement_bundler_cpu.cpp:435-529), shared by the oracle pins
and the GPU parity tests.

  reference_style_problem       the reference test generator's scene (sba_test.cpp:46-128)
  exact_observation_problem     its exact-observation variant (sba_test.cpp:1322-1355): converges, early exit
  stressed_window               large initial errors and 10% outliers: rejected steps, gain ratios on both sides
                                of 0.25 and 0.75
  window_with_unobserved_pose   a free key frame that no observation sees: its reduced-system block is zero, so
                                every factorization fails and every iteration multiplies lambda by 5
  window_with_negative_information
                                observations of fixed points made with a negated information matrix: the
                                reduced system is indefinite at small damping and positive definite at larger
                                damping, so factorizations fail, then succeed
"""
import numpy as np

import cuvslam_amd as ca


def _exp(tw):
    """Exp of a twist (omega, v) as a 4x4 matrix (Rodrigues' formula and its V matrix)."""
    w, v = np.asarray(tw[:3], float), np.asarray(tw[3:], float)
    th = np.linalg.norm(w)
    K = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    if th < 1e-12:
        R, V = np.eye(3) + K, np.eye(3) + 0.5 * K
    else:
        a, b = np.sin(th) / th, (1 - np.cos(th)) / th ** 2
        R = np.eye(3) + a * K + b * K @ K
        V = np.eye(3) + b * K + (1 - a) / th ** 2 * K @ K
    M = np.eye(4)
    M[:3, :3], M[:3, 3] = R, V @ v
    return M


def reference_style_problem(seed, num_points=200, num_poses=5):
    """The reference test generator's scene (sba_test.cpp:46-128): two-camera rig with a 0.125 m baseline,
    poses Exp(omega ~ U(+-0.25), t ~ U(+-2)), points xy ~ U(+-4), z ~ U(2, 16), every (pose, point, camera)
    with camera depth > 1 observed, identity information, one fixed pose."""
    rng = np.random.default_rng(seed)
    poses = []
    for _ in range(num_poses):
        tw = np.concatenate([rng.uniform(-0.25, 0.25, 3), rng.uniform(-2, 2, 3)])
        poses.append(_exp(tw).astype(np.float32))
    pts = np.stack([rng.uniform(-4, 4, num_points), rng.uniform(-4, 4, num_points), rng.uniform(2, 16, num_points)], 1)
    cams = np.stack([np.eye(4), np.eye(4)]).astype(np.float32)
    cams[1, 0, 3] = 0.125
    xy, pid, qid, cid = [], [], [], []
    for j in range(num_poses):
        for i in range(num_points):
            for c in range(2):
                q = cams[c].astype(np.float64) @ poses[j].astype(np.float64) @ np.append(pts[i], 1)
                if q[2] <= 1.0:
                    continue
                xy.append(q[:2] / q[2])
                pid.append(i)
                qid.append(j)
                cid.append(c)
    n = len(xy)
    return ca.BundleAdjustmentProblem(pts.astype(np.float32), np.array(xy, np.float32),
                                      np.tile(np.eye(2, dtype=np.float32), (n, 1, 1)), np.array(pid, np.int32),
                                      np.array(qid, np.int32), np.array(cid, np.int8), np.stack(poses), cams,
                                      num_fixed_points=0, num_fixed_key_frames=1, max_iterations=50, robustifier_scale=1.0)


def exact_observation_problem(seed=1):
    """sba_test.cpp:1322-1355: exact projections of the generated state, free key frames shifted by
    (0.01, -0.01, 0.01) m, 50 iterations."""
    p = reference_style_problem(seed)
    p.rig_from_world[:-1, :3, 3] += np.array([0.01, -0.01, 0.01], np.float32)
    return p


def stressed_window(max_iterations=15):
    """Six key frames (two fixed), 300 points, initial errors of 0.5 m / 10 deg (poses) and 3 m (points), 10%
    outliers. Under the Huber loss the FP64 solve rejects about half its steps and takes gain ratios below 0.25,
    between 0.25 and 0.75 and above 0.75."""
    p, _, _ = ca.scene.generate(seed=2, num_points=300, num_keyframes=6, num_fixed_keyframes=2,
                                max_iterations=max_iterations, pose_sigma_m=0.5, pose_sigma_deg=10.0,
                                point_sigma_m=3.0, outlier_rate=0.1)
    return p


def _keep_observations(p, keep):
    q = p.copy()
    for f in ("observation_xys", "observation_infos", "point_ids", "pose_ids", "camera_ids"):
        setattr(q, f, getattr(q, f)[keep])
    return q


def window_with_unobserved_pose(max_iterations=6):
    """Six key frames (two fixed), 300 points; every observation from free key frame 1 removed."""
    p, _, _ = ca.scene.generate(seed=3, num_points=300, num_keyframes=6, num_fixed_keyframes=2,
                                max_iterations=max_iterations)
    return _keep_observations(p, p.pose_ids != 1)


def window_with_negative_information(scale=3.0, seed=3, max_iterations=12):
    """Six key frames (two fixed), 300 points of which the last ten are fixed. The observations of those fixed
    points from free key frame 0 are replaced by their exact projections at the initial state and carry the
    information matrix times -scale: they add a negative term to that key frame's block of the reduced system
    (and nothing to the initial cost). With scale 3 and seed 3 the FP64 solve fails its first five
    factorizations, then accepts, fails again and accepts."""
    p, _, _ = ca.scene.generate(seed=seed, num_points=300, num_keyframes=6, num_fixed_keyframes=2,
                                max_iterations=max_iterations)
    q = p.copy()
    q.num_fixed_points = 10
    npnt = q.points.shape[0]
    sel = (q.point_ids >= npnt - 10) & (q.pose_ids == 0)
    T = q.camera_from_rig[0].astype(np.float64) @ q.rig_from_world[0].astype(np.float64)
    P = np.c_[q.points[q.point_ids[sel]].astype(np.float64), np.ones(int(sel.sum()))] @ T.T
    q.observation_xys[sel] = (P[:, :2] / P[:, 2:3]).astype(np.float32)
    q.observation_infos[sel] = (-scale * q.observation_infos[sel]).astype(np.float32)
    return q
