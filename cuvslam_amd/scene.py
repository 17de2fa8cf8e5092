"""Seeded synthetic SBA scenes (the benchmark generator of csrc/scene.cpp).

Presets follow the benchmark configurations: a pinhole camera fx=fy=450, cx=376, cy=240 (752x480), key frames
0.1 m apart along +x with 1 deg random rotations, points x, y in [-5, 5] m and z in [4, 20] m, N(0, 0.5 px)
pixel noise, uniform-pixel outliers, initial state perturbed by 2 cm / 0.5 deg (poses) and 5 cm (points),
information diag(fx^2, fy^2) so that a Huber scale of 1 is 1 px, the first key frame fixed and stored last.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .sba import BundleAdjustmentProblem

PRESETS = {
    # 10 key frames, 2,000 points (17.8k observations with every point seen wherever it projects in-image)
    "c1": dict(num_keyframes=10, num_points=2000),
    # 30 key frames, 20,000 points (528k observations)
    "c2": dict(num_keyframes=30, num_points=20000),
    # 20 key frames, 10,000 points
    "c4": dict(num_keyframes=20, num_points=10000),
}


def config(**overrides) -> _lib.SceneConfig:
    cfg = _lib.SceneConfig()
    _lib.lib().sba_scene_default_config(C.byref(cfg))
    for k, v in overrides.items():
        setattr(cfg, k, v)
    return cfg


def generate(preset: str | None = None, seed: int = 1, max_iterations: int = 10, **overrides):
    """Returns (problem, gt_points (N,3), gt_rig_from_world (N,4,4))."""
    kw = dict(PRESETS[preset]) if preset else {}
    kw.update(overrides)
    cfg = config(seed=seed, **kw)
    L = _lib.lib()
    sc = C.c_void_p()
    rc = L.sba_scene_generate(C.byref(cfg), C.byref(sc))
    if rc != 0:
        raise RuntimeError(f"sba_scene_generate failed ({rc})")
    return _take(L, sc, max_iterations)


def loop(seed: int = 5, max_iterations: int = 10, **overrides):
    """The large multi-camera map (defaults: 4 stereo pairs, 200 key frames on a 20 m loop, 200,000 points in a
    corridor, 3% outliers, one fixed key frame). Returns (problem, gt_points, gt_rig_from_world)."""
    cfg = _lib.LoopConfig()
    L = _lib.lib()
    L.sba_loop_default_config(C.byref(cfg))
    cfg.seed = seed
    for k, v in overrides.items():
        setattr(cfg, k, v)
    sc = C.c_void_p()
    rc = L.sba_scene_generate_loop(C.byref(cfg), C.byref(sc))
    if rc != 0:
        raise RuntimeError(f"sba_scene_generate_loop failed ({rc})")
    return _take(L, sc, max_iterations)


def _take(L, sc, max_iterations):
    try:
        v = _lib.ProblemView()
        L.sba_scene_view(sc, C.byref(v))
        n_p, n_o, n_c, n_cam = v.num_points, v.num_observations, v.num_poses, v.num_cameras

        def arr(ptr, count, dtype):
            if count == 0:
                return np.zeros(0, dtype)
            return np.ctypeslib.as_array(ptr, shape=(count,)).astype(dtype, copy=True)

        pts = arr(v.points, 3 * n_p, np.float32).reshape(n_p, 3)
        xy = arr(v.observation_xys, 2 * n_o, np.float32).reshape(n_o, 2)
        info = arr(v.observation_infos, 4 * n_o, np.float32).reshape(n_o, 2, 2).transpose(0, 2, 1).copy()
        pid = arr(v.point_ids, n_o, np.int32)
        qid = arr(v.pose_ids, n_o, np.int32)
        cid = arr(v.camera_ids, n_o, np.int8)
        poses = arr(v.rig_from_world, 16 * n_c, np.float32).reshape(n_c, 4, 4).transpose(0, 2, 1).copy()
        cams = arr(v.camera_from_rig, 16 * n_cam, np.float32).reshape(n_cam, 4, 4).transpose(0, 2, 1).copy()
        gp, gq = _lib._fp(), _lib._fp()
        L.sba_scene_ground_truth(sc, C.byref(gp), C.byref(gq))
        gt_pts = arr(gp, 3 * n_p, np.float32).reshape(n_p, 3)
        gt_poses = arr(gq, 16 * n_c, np.float32).reshape(n_c, 4, 4).transpose(0, 2, 1).copy()
        prob = BundleAdjustmentProblem(pts, xy, info, pid, qid, cid, poses, cams, num_fixed_points=0,
                                       num_fixed_key_frames=v.num_fixed_key_frames, max_iterations=max_iterations,
                                       robustifier_scale=1.0)
        return prob, gt_pts, gt_poses
    finally:
        L.sba_scene_free(sc)


def batched_windows(n_windows: int = 512, seed: int = 3, max_iterations: int = 10, kf_range=(5, 40), num_fixed_keyframes: int = 2,
                    pt_range=(500, 8000), outlier_range=(0.0, 0.10)):
    """The uneven batch: per window, key frames U{kf_range}, points U{pt_range}, outlier rate U(outlier_range),
    each point observed by U{2..key frames} of the key frames it projects in. Window w uses seed
    seed * 10**6 + w for its scene and a separate stream for its shape."""
    rng = np.random.default_rng(seed)
    out = []
    for w in range(n_windows):
        nk = int(rng.integers(kf_range[0], kf_range[1] + 1))
        npt = int(rng.integers(pt_range[0], pt_range[1] + 1))
        orate = float(rng.uniform(*outlier_range))
        p, _, _ = generate(seed=seed * 10**6 + w, max_iterations=max_iterations, num_keyframes=nk, num_points=npt,
                           outlier_rate=orate, sample_observations=1, num_fixed_keyframes=num_fixed_keyframes)
        out.append(p)
    return out
