"""Synthetic monocular windows for the mono SBA solver (libs/sba/mono_sba_solver.cpp), in the form its caller builds
them (pipelines/sba_mono_wrapper.cpp:133-215): up to ten key frames, the oldest three constrained, tracks seen from at
least two frames, observations in normalized image coordinates."""
import numpy as np
from scipy.spatial.transform import Rotation


def mono_window(seed=0, num_frames=8, num_tracks=300, noise_px=0.5, pose_sigma_m=0.01, pose_sigma_deg=0.3,
                point_sigma_m=0.05, num_constrained=3, focal=450.0, max_iterations=30, min_residual=0.02,
                min_residual_speed=0.1, exact=False):
    rng = np.random.default_rng(seed)
    gt_poses = []
    for f in range(num_frames):
        T = np.eye(4)
        T[:3, :3] = Rotation.from_rotvec(np.radians(rng.normal(0, 1.0, 3))).as_matrix()
        T[:3, 3] = [0.1 * f, 0.02 * np.sin(f), 0.0]
        gt_poses.append(T)  # world_from_camera
    pts = np.stack([rng.uniform(-4, 4, num_tracks), rng.uniform(-3, 3, num_tracks), rng.uniform(4, 15, num_tracks)], 1)
    of, ot, xy = [], [], []
    for f in range(num_frames):
        cw = np.linalg.inv(gt_poses[f])
        for t in range(num_tracks):
            q = cw[:3, :3] @ pts[t] + cw[:3, 3]
            if q[2] <= 1.0:
                continue
            u = q[:2] / q[2]
            if abs(u[0]) > 0.8 or abs(u[1]) > 0.5:
                continue
            if not exact:
                u = u + rng.normal(0, noise_px / focal, 2)
            of.append(f)
            ot.append(t)
            xy.append(u)
    of, ot = np.array(of, np.int32), np.array(ot, np.int32)
    keep = np.bincount(ot, minlength=num_tracks) >= 2
    sel = keep[ot]
    of, ot, xy = of[sel], ot[sel], np.array(xy)[sel]
    remap = -np.ones(num_tracks, np.int64)
    remap[np.nonzero(keep)[0]] = np.arange(keep.sum())
    ot = remap[ot].astype(np.int32)
    gt_pts = pts[keep]
    poses = []
    for f, T in enumerate(gt_poses):
        P = T.copy()
        if f >= num_constrained:
            P[:3, :3] = Rotation.from_rotvec(np.radians(rng.normal(0, pose_sigma_deg, 3))).as_matrix() @ P[:3, :3]
            P[:3, 3] += rng.normal(0, pose_sigma_m, 3)
        poses.append(P)
    init_pts = gt_pts + rng.normal(0, point_sigma_m, gt_pts.shape)
    return dict(constrained=np.array([f < num_constrained for f in range(num_frames)], np.int32),
                frame_poses=np.array(poses, np.float32), points=init_pts.astype(np.float32), obs_frame=of,
                obs_track=ot, obs_xy=xy.astype(np.float32), max_iterations=max_iterations, min_residual=min_residual,
                min_residual_speed=min_residual_speed, gt_poses=np.array(gt_poses), gt_points=gt_pts)
