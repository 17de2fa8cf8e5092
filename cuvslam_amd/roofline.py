"""Algorithmic work of one LM iteration per kernel class, from a window's structure.

Counts are what the algorithm must move or compute, not what an implementation happens to do:
  N_o observations, P (point, key frame) pairs, P_f pairs of a free point with a free key frame,
  N_p points (N_pf free), q_i = free pairs of free point i, n = 6 * free key frames per window.

k_schur (linearize + Schur complement)
  bytes  24 N_o (xy 8, information 12 + camera 4) + 8 P (first observation, key frame) + 72 N_p
         (pair offset 4, free count 4, point 16 read; V*^-1, V*^-1 e_b, diag V 48 written)
         + 4 (36 nb + 12 n/6) per window (the reduced system written once, FP32)
  flops  400 N_o (projection, 2x6 and 2x3 Jacobians, robust weight, U/e_a/W/V/e_b accumulation)
         + 600 N_pf (3x3 Jacobi pseudo-inverse) + 144 P_f (Z = W G, e update)
         + sum_i [216 q_i (q_i - 1) / 2 + 126 q_i] (6x6 Schur pair products, diagonal blocks symmetric)
k_backsub (back-substitution, trial cost, predicted-reduction terms)
  bytes  24 N_o + 8 P + 80 N_p (point 16 and V*^-1 data 48 read, trial point 16 written)
  flops  300 N_o FP32 (two linearizations) + 60 N_o FP64 (trial residual and robust cost)
k_factor (FP64 Cholesky + two triangular solves + trial poses), per window
  flops  n^3 / 3 + 2 n^2 (FP64)
"""
from __future__ import annotations

import numpy as np

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
FP32_PEAK_TFLOPS = 157.3       # MI355X FP32 (vector and matrix)
FP64_PEAK_TFLOPS = 78.6        # MI355X FP64 (vector and matrix)


def window_structure(problem):
    """Per-window counts used by the formulas above."""
    n_o = problem.num_observations
    n_p = problem.points.shape[0]
    n_pf = n_p - problem.num_fixed_points
    ncf = problem.rig_from_world.shape[0] - problem.num_fixed_key_frames
    pid = problem.point_ids.astype(np.int64)
    qid = problem.pose_ids.astype(np.int64)
    key = np.unique(pid * (problem.rig_from_world.shape[0] + 1) + qid)
    pair_pt = key // (problem.rig_from_world.shape[0] + 1)
    pair_pose = key % (problem.rig_from_world.shape[0] + 1)
    P = key.shape[0]
    free = (pair_pt < n_pf) & (pair_pose < ncf)
    q = np.bincount(pair_pt[free], minlength=n_p).astype(np.float64)
    return dict(n_o=n_o, n_p=n_p, n_pf=n_pf, ncf=ncf, P=P, P_f=int(free.sum()), q=q)


def work(problems):
    """Algorithmic bytes and flops per LM iteration of each kernel class, summed over windows."""
    out = {"schur": [0.0, 0.0, 0.0], "backsub": [0.0, 0.0, 0.0], "factor": [0.0, 0.0, 0.0]}  # bytes, f32, f64
    for p in problems:
        s = window_structure(p)
        nb = s["ncf"] * (s["ncf"] + 1) / 2
        n = 6 * s["ncf"]
        q = s["q"]
        out["schur"][0] += 24 * s["n_o"] + 8 * s["P"] + 72 * s["n_p"] + 4 * (36 * nb + 12 * s["ncf"])
        out["schur"][1] += (400 * s["n_o"] + 600 * s["n_pf"] + 144 * s["P_f"]
                            + float(np.sum(216 * q * (q - 1) / 2 + 126 * q)))
        out["backsub"][0] += 24 * s["n_o"] + 8 * s["P"] + 80 * s["n_p"]
        out["backsub"][1] += 300 * s["n_o"]
        out["backsub"][2] += 60 * s["n_o"]
        out["factor"][0] += 4 * (36 * nb + 12 * s["ncf"]) + 8 * n * (n + 1) / 2
        out["factor"][2] += n ** 3 / 3 + 2 * n ** 2
    return out
