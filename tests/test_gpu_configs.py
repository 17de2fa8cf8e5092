"""The benchmark configurations at full size through the HIP path, against the FP64 oracle.

The bar is 1e-6 relative in the final cost, with two stated exceptions whose evidence the tests compute
(tests/sensitivity.py):
  * a window whose FP64 solve takes a gain ratio within 0.01 of the 0.25 / 0.75 damping thresholds has two valid
    trajectories (FP32 rounding can put its ratio on the other side, and lambda differs by 8x or 40x from there on):
    it is compared with the nearest of the FP64 trajectories that take either side of each such threshold;
  * a window on which FP32 arithmetic itself moves the result: the oracle's own FP32 restatement of the method (an
    independent summation order) ends a distance s32 from the FP64 oracle, and the HIP path must end within
    max(1e-6, 10 s32), with an accepted-step count between the two solves. On the 200-key-frame map s32 is 2.6e-4
    (the FP32 restatement accepts all 10 steps, the FP64 solve 7: the steps at damping near 1e-7 are decided by
    rounding), while the FP64 solve moves by 5e-8 under one-ulp perturbations of the FP32 inputs and by 6e-10 under
    an FP32 rounding of the finished reduced system.
"""
import json
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import cuvslam_amd as ca  # noqa: E402
from oracle import sba_oracle as O  # noqa: E402
from sensitivity import branch_variants, fp32_method_spread, nearest_variant  # noqa: E402

pytestmark = pytest.mark.gpu

HUBER = ca.LOSS_HUBER
HERE = os.path.dirname(os.path.abspath(__file__))


def check_against_oracle(problem, solved, variants=None):
    variants = variants if variants is not None else branch_variants(problem, HUBER, fixed_iterations=1)
    err, flip, ref = nearest_variant(variants, solved.last_cost_f64)
    tol, accepted = 1e-6, (ref["accepted"], ref["accepted"])
    if err >= tol:  # the window's FP32 spread, measured by the oracle's FP32 restatement of the same trajectory
        s32, _, r32 = fp32_method_spread(problem, HUBER, reference=variants[0][1] if flip == 0 else None)
        tol = max(tol, 10.0 * s32)
        accepted = tuple(sorted((ref["accepted"], int(r32["accepted"]))))
    assert err < tol, (err, tol, flip, [(f, r["last_cost"]) for f, r in variants], solved.last_cost_f64)
    assert solved.iterations == ref["iterations"], flip
    # where FP32 arithmetic moves the result, it also moves single accept/reject decisions: the accepted count
    # lies between the FP64 solve's and the FP32 restatement's
    assert accepted[0] <= solved.accepted <= accepted[1], (solved.accepted, accepted, flip)
    # the returned FP32 state has the reported cost
    c, _ = O.cost(solved, loss=HUBER)
    assert abs(c - solved.last_cost_f64) <= 1e-9 * c
    return err, flip, ref


def test_c2_thirty_key_frames_twenty_thousand_points():
    # 30 key frames (2 fixed), 20,000 points, 528k observations, 10 iterations
    p, _, _ = ca.scene.generate("c2", seed=2, num_fixed_keyframes=2, max_iterations=10)
    q = p.copy()
    assert ca.SchurComplementBundlerGpu(loss=HUBER, fixed_iterations=True).solve(q)
    err, flip, ref = check_against_oracle(p, q)
    nf = p.rig_from_world.shape[0] - p.num_fixed_key_frames
    assert np.abs(q.rig_from_world[:nf, :3, 3] - ref["poses"][:nf, :3, 3]).max() < 1e-5  # metres
    perr = np.linalg.norm(q.points - ref["points"], axis=1)
    assert np.median(perr) < 1e-5


def test_c3_five_hundred_twelve_windows():
    probs = ca.scene.batched_windows(512, seed=3, max_iterations=10)
    assert sum(p.num_observations for p in probs) > 2.5e7
    solver = ca.SchurComplementBundlerGpu(loss=HUBER, fixed_iterations=True, chunk_pairs=3000)
    qs = [p.copy() for p in probs]
    assert all(solver.solve_batch(qs))
    for q in qs:
        assert q.status == 0 and q.iterations == 10
        assert np.isfinite(q.last_cost_f64) and q.last_cost_f64 <= q.initial_cost_f64
    sizes = np.array([p.num_observations for p in probs])
    kfs = np.array([p.rig_from_world.shape[0] for p in probs])
    rng = np.random.default_rng(0)
    sample = sorted({0, int(np.argmax(sizes)), int(np.argmin(sizes)), int(np.argmin(kfs)), int(np.argmax(kfs)),
                     *rng.choice(len(probs), 3, replace=False).tolist()})
    for w in sample:
        # the same window alone (same chunking): the batch changes nothing
        alone = probs[w].copy()
        assert solver.solve(alone)
        assert (alone.iterations, alone.accepted) == (qs[w].iterations, qs[w].accepted)
        assert abs(alone.last_cost_f64 - qs[w].last_cost_f64) <= 1e-12 * alone.last_cost_f64
        np.testing.assert_allclose(alone.points, qs[w].points, rtol=0, atol=1e-6)
        # and the FP64 oracle
        check_against_oracle(probs[w], qs[w])


def test_c5_two_hundred_key_frame_map():
    # 8 cameras, 200 key frames (1 fixed), 200,000 points, 3.0M observations; the reduced system is 1194 x 1194.
    # The oracle's solves are kept in tests/golden/c5_variants.json (about 25 s each on one core): the FP64 one, the
    # FP64 one that takes the other branch at iteration 8 (gain ratio 0.7457, 0.004 below the 0.75 threshold), the
    # spread under one-ulp input perturbations and the distance of the FP32 restatement.
    with open(os.path.join(HERE, "golden", "c5_variants.json")) as f:
        fx = json.load(f)
    p, _, _ = ca.scene.loop(seed=5, max_iterations=10)
    assert p.num_observations == fx["observations"]
    q = p.copy()
    assert ca.SchurComplementBundlerGpu(loss=HUBER, fixed_iterations=True).solve(q)
    variants = [(v["flip_iteration"], {"last_cost": v["last_cost"], "iterations": v["iterations"],
                                       "accepted": v["accepted"]}) for v in fx["variants"]]
    err, flip, ref = nearest_variant(variants, q.last_cost_f64)
    assert fx["fp32_input_spread"] < 1e-7 and fx["fp32_method_spread"] > 1e-4
    assert err < max(1e-6, 10.0 * fx["fp32_method_spread"]), (err, flip, fx["variants"], q.last_cost_f64)
    assert q.iterations == ref["iterations"]
    # Which steps are accepted is decided by FP32 rounding here: the FP64 solve accepts 7 of 10 steps, five of its
    # gain ratios lie within 0.1 of a threshold (0.21, 0.21, 0.18 and 0.34 against 0.25; 0.746 against 0.75), and
    # the FP32 restatement of the method accepts all 10 in every observation order tried (fp32_order_variants).
    # The count may differ from the FP64 one by at most those marginal decisions.
    marginal = sum(1 for r in fx["trace_rho"] if min(abs(r - 0.25), abs(r - 0.75)) < 0.1)
    assert marginal == 5 and all(v["accepted"] == 10 for v in fx["fp32_order_variants"])
    assert abs(q.accepted - ref["accepted"]) <= marginal, (q.accepted, ref["accepted"], marginal)
