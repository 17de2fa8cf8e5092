"""GPU parity of the Levenberg-Marquardt control (schur_complement_bundler_gpu.cpp:175-253) against the FP64 oracle:
the step maxima and the predicted reduction of one step, the final damping of whole solves, the reference's
early exit on exact observations (sba_test.cpp:1322-1355) and failed factorizations inside a full solve
(gpu.cpp:204-209; sba_test.cpp:984-1025)."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import cuvslam_amd as ca  # noqa: E402
from lm_cases import (exact_observation_problem, stressed_window, window_with_negative_information,  # noqa: E402
                      window_with_unobserved_pose)
from oracle import sba_oracle as O  # noqa: E402

pytestmark = pytest.mark.gpu

HUBER, STUDENT = ca.LOSS_HUBER, ca.LOSS_STUDENT_T


def rel(a, b):
    return abs(a - b) / abs(b)


@pytest.mark.parametrize("loss", [HUBER, STUDENT])
@pytest.mark.parametrize("lam", [1e-3, 1e-1, 10.0])
def test_step_maxima_and_prediction_match_oracle(loss, lam):
    p, _, _ = ca.scene.generate(seed=3, num_points=300, num_keyframes=6, num_fixed_keyframes=2)
    b = ca.SchurComplementBundlerGpu(loss=loss).batch([p])
    _, _, _, _, sc, rc = b.debug_step(lam)
    _, _, _, _, so, rco = O.step(p, lam, loss=loss)
    assert rc == 0 and rco == 0
    # the inputs of the step-size stop: max |dp| and max |dc| (the steps themselves agree to 1e-3)
    assert rel(sc[4], so[4]) < 1e-3
    assert rel(sc[5], so[5]) < 1e-3
    # the predicted relative reduction the gain ratio divides by, from the device's FP64 terms
    assert rel((sc[2] + 2.0 * np.float32(lam) * sc[3]) / sc[6], so[8]) < 1e-3


@pytest.mark.parametrize("loss", [HUBER, STUDENT])
def test_final_damping_matches_oracle(loss):
    p, _, _ = ca.scene.generate(seed=5, num_points=500, num_keyframes=8, num_fixed_keyframes=2, max_iterations=10)
    ref = O.solve(p, "f64", loss=loss)
    q = p.copy()
    assert ca.SchurComplementBundlerGpu(loss=loss).solve(q)
    assert (q.iterations, q.accepted) == (ref["iterations"], ref["accepted"])
    assert rel(q.final_lambda, ref["lam"]) < 1e-6


def trace_of(problem, loss=HUBER, **kw):
    solver = ca.SchurComplementBundlerGpu(loss=loss, **kw)
    b = solver.batch([problem])
    b.run()
    q = problem.copy()
    b.read([q])
    return q, b.trace(0)


def test_stressed_window_takes_the_oracle_damping_path():
    # initial errors of 0.5 m / 10 deg (poses) and 3 m (points), 10% outliers. Over the first eight iterations the
    # device takes the FP64 oracle's decisions: accepted with lambda / 8 twice, rejected (lambda x 5) three times,
    # accepted with lambda kept (gain ratio 0.64), accepted with lambda / 8 twice. From there the two trajectories
    # part: each FP32 step of this window differs from the FP64 one by about 1e-3 (its first trial cost by 2e-4), and
    # after eight large steps the states have moved apart enough to change later gain ratios by more than their
    # distance to the thresholds.
    p = stressed_window(max_iterations=15)
    ref = O.solve(p, "f64", loss=HUBER)
    q, t = trace_of(p)
    n = 8
    assert list(ref["trace_flags"][:n]) == [1, 1, 0, 0, 0, 1, 1, 1]
    assert list(t["flags"][:n]) == list(ref["trace_flags"][:n])
    np.testing.assert_allclose(t["lambda"][:n], ref["trace_lambda"][:n], rtol=1e-6)
    np.testing.assert_allclose(t["rho"][:6], ref["trace_rho"][:6], atol=0.02)
    assert rel(t["trial_cost"][0], ref["trace_cost"][0]) < 1e-3
    assert q.iterations == 15 and q.accepted >= 5


def test_exact_observations_converge_with_the_early_exit():
    # sba_test.cpp:1322-1355 through the HIP path with the reference's stop tests (no fixed iteration count)
    p = exact_observation_problem()
    ref = O.solve(p, "f64", loss=STUDENT)
    q = p.copy()
    assert ca.SchurComplementBundlerGpu(loss=STUDENT).solve(q)
    assert np.isfinite(q.initial_cost) and np.isfinite(q.last_cost)
    assert q.last_cost < q.initial_cost
    assert q.iterations < p.max_iterations
    assert q.iterations == ref["iterations"]
    assert q.accepted == ref["accepted"]


def test_every_factorization_fails_when_a_free_key_frame_is_unobserved():
    # the reduced system has a zero block: every factorization fails and becomes damping (gpu.cpp:204-209); the
    # solve still returns true, with the initial state
    p = window_with_unobserved_pose(max_iterations=6)
    ref = O.solve(p, "f64", loss=HUBER)
    q = p.copy()
    assert ca.SchurComplementBundlerGpu(loss=HUBER).solve(q)
    assert q.iterations == ref["iterations"] == 6
    assert q.accepted == ref["accepted"] == 0
    assert rel(q.final_lambda, ref["lam"]) < 1e-6
    assert q.last_cost_f64 == q.initial_cost_f64
    np.testing.assert_array_equal(q.points, p.points)


def test_failed_factorizations_then_recovery_match_oracle():
    # an indefinite reduced system at small damping: five failed factorizations (lambda x 5 each), then accepted
    # steps, then failures again. Every decision and every damping value is the oracle's. The first factorization
    # that succeeds does so at lambda = 3.1, just above where the damping overcomes the negative block, so the solves
    # of this barely positive-definite system carry the FP32 assembly error amplified by its condition number: the
    # final cost agrees to 1e-2, the gain ratios to 1e-2.
    p = window_with_negative_information()
    ref = O.solve(p, "f64", loss=HUBER)
    assert list(ref["trace_flags"][:6]) == [2, 2, 2, 2, 2, 1]
    q, t = trace_of(p)
    assert (q.iterations, q.accepted) == (ref["iterations"], ref["accepted"])
    assert list(t["flags"]) == list(ref["trace_flags"])
    np.testing.assert_allclose(t["lambda"], ref["trace_lambda"], rtol=1e-6)
    np.testing.assert_allclose(t["rho"], ref["trace_rho"], atol=1e-2)
    assert rel(q.final_lambda, ref["lam"]) < 1e-6
    assert rel(q.last_cost_f64, ref["last_cost"]) < 1e-2
