"""Pins of the oracle's Levenberg-Marquardt control (runs without a GPU).

The step itself is pinned elsewhere (tests/test_oracle.py); these pin what the loop does with it:
  * the predicted reduction (ComputePredictedRelativeReduction, schur_complement_bundler_cpu.cpp:385-433) is the
    decrease of the quadratic model of the cost, 2 g^T d - d^T H d, written by numpy alone into the golden
    fixtures; with (H + lambda D) d = g that is d^T H d + 2 lambda d^T D d, the form the bundlers compute;
  * the gain ratio tends to 1 under heavy damping (the model is exact to first order), which fixes the scale
    of the prediction (a dropped factor 2 on the damping term gives 1/2);
  * the damping schedule and the accept rule over whole traces that take every branch: lambda_0 = 1e-3 (as an
    FP32 constant), accept iff rho > 0.25, lambda x 0.125 iff rho > 0.75, lambda x 5 on a rejected step or a
    failed factorization (schur_complement_bundler_cpu.cpp:497-526, schur_complement_bundler_gpu.cpp:204-253);
  * a vanished cost gives a prediction that is not finite (sba_test.cpp:1274-1317);
  * the reference's exact-observation convergence (sba_test.cpp:1322-1355) ends on the stop test, not on the
    iteration budget.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import cuvslam_amd as ca  # noqa: E402
from lm_cases import (exact_observation_problem, stressed_window, window_with_negative_information,  # noqa: E402
                      window_with_unobserved_pose)
from oracle import sba_oracle as O  # noqa: E402
from test_oracle import fixture_problems  # noqa: E402

ACCEPTED, FACTOR_FAILED, STOPPED = 1, 2, 4


@pytest.mark.parametrize("case", range(3))
def test_prediction_is_the_decrease_of_the_quadratic_model(case):
    p, c = fixture_problems()[case]
    for st in c["steps"]:
        _, _, _, _, sc, rc = O.step(p, st["lam"], loss=c["loss"])
        assert rc == 0
        pred_times_cost = sc[8] * sc[6]
        # limited by the fixture's finite-difference Jacobians
        assert abs(pred_times_cost - st["model_reduction"]) / st["model_reduction"] < 1e-6, (st["lam"], pred_times_cost)


@pytest.mark.parametrize("case", range(3))
def test_gain_ratio_tends_to_one_under_heavy_damping(case):
    p, c = fixture_problems()[case]
    st = [s for s in c["steps"] if s["lam"] >= 1e3][0]
    _, _, _, _, sc, _ = O.step(p, st["lam"], loss=c["loss"])
    # actual relative reduction (numpy's trial cost) over the oracle's predicted relative reduction
    rho = (1.0 - st["trial_cost"] / c["initial_cost"]) / sc[8]
    assert abs(rho - 1.0) < 0.01, rho


def check_schedule(r, max_iterations):
    """Asserts the damping schedule and the accept rule on an oracle trace; returns the branches seen."""
    lam, rho, flags = r["trace_lambda"], r["trace_rho"], r["trace_flags"]
    seen = set()
    assert lam[0] == np.float64(np.float32(1e-3))
    assert r["iterations"] == len(flags) <= max_iterations
    for k in range(len(flags)):
        nxt = lam[k + 1] if k + 1 < len(flags) else r["lam"]
        if flags[k] & FACTOR_FAILED:
            assert not flags[k] & ACCEPTED
            assert nxt == lam[k] * 5
            seen.add("factor failed")
            continue
        if flags[k] & STOPPED:
            assert k == len(flags) - 1 and nxt == lam[k]
            seen.add("stopped")
            continue
        accepted = bool(flags[k] & ACCEPTED)
        assert accepted == (rho[k] > 0.25), (k, rho[k], flags[k])
        if not accepted:
            assert nxt == lam[k] * 5
            seen.add("rejected")
        elif rho[k] > 0.75:
            assert nxt == lam[k] * 0.125
            seen.add("accepted, lambda / 8")
        else:
            assert nxt == lam[k]
            seen.add("accepted, lambda kept")
    assert r["accepted"] == int(np.sum(flags & ACCEPTED))
    return seen


def test_schedule_over_traces_that_take_every_branch():
    seen = set()
    for p, c in fixture_problems():
        seen |= check_schedule(O.solve(p, "f64", loss=c["loss"]), 10)
    p, _, _ = ca.scene.generate("c1", seed=1, num_fixed_keyframes=2, max_iterations=10)
    seen |= check_schedule(O.solve(p, "f64", loss=ca.LOSS_HUBER), 10)
    p = stressed_window()
    seen |= check_schedule(O.solve(p, "f64", loss=ca.LOSS_HUBER), 15)
    p = window_with_negative_information()
    seen |= check_schedule(O.solve(p, "f64", loss=ca.LOSS_HUBER), 12)
    p = exact_observation_problem()
    seen |= check_schedule(O.solve(p, "f64", loss=ca.LOSS_STUDENT_T), 50)
    assert seen == {"factor failed", "stopped", "rejected", "accepted, lambda / 8", "accepted, lambda kept"}, seen


def test_every_factorization_fails_when_a_free_key_frame_is_unobserved():
    p = window_with_unobserved_pose(max_iterations=6)
    r = O.solve(p, "f64", loss=ca.LOSS_HUBER)
    assert r["status"] == 0  # not an error: the solve ran (gpu.cpp:204-209 turns it into damping)
    assert r["iterations"] == 6 and r["accepted"] == 0
    assert all(f == FACTOR_FAILED for f in r["trace_flags"])
    assert r["lam"] == np.float64(np.float32(1e-3)) * 5 ** 6
    assert r["last_cost"] == r["initial_cost"]


def test_negative_information_fails_then_recovers():
    r = O.solve(window_with_negative_information(), "f64", loss=ca.LOSS_HUBER)
    f = r["trace_flags"]
    assert list(f[:6]) == [FACTOR_FAILED] * 5 + [ACCEPTED]
    assert 0 < r["accepted"] < r["iterations"] == 12


def test_prediction_needs_a_positive_cost():
    # sba_test.cpp:1274-1317: damping 0.01, current cost 10 -> finite; current cost 0 -> not a number
    p, c = fixture_problems()[0]
    assert np.isfinite(O.predicted(p, 0.01, 10.0, loss=c["loss"]))
    assert not np.isfinite(O.predicted(p, 0.01, 0.0, loss=c["loss"]))


def test_exact_observations_end_on_the_stop_test():
    r = O.solve(exact_observation_problem(), "f64", loss=ca.LOSS_STUDENT_T)
    assert r["status"] == 0 and r["iterations"] < 50
    assert r["trace_flags"][-1] == STOPPED
    assert r["last_cost"] < r["initial_cost"] * 1e-6
