"""GPU parity of the batched mono SBA solve (sba_mono_solve_batch, the path of MonoSBASolver::solve,
libs/sba/mono_sba_solver.cpp:205-301) against the FP64 oracle: both compute in FP64, so the iteration counts, the
damping, the failure flags and the residuals agree (residuals to 1e-9 relative); the returned FP32 state agrees to
its rounding and the covariance to 1e-6 relative."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import mono_cases as MC  # noqa: E402
from cuvslam_amd import _lib, mono  # noqa: E402
from oracle import sba_oracle as O  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def solver():
    return mono.MonoSolverGpu()


def check(g, r):
    assert g["feasible"] == r["feasible"] and g["failed"] == r["failed"]
    assert g["iterations"] == r["iterations"] and g["lam"] == r["lam"]
    assert abs(g["residual"] - r["residual"]) <= 1e-9 * abs(r["residual"])
    assert abs(g["residual_initial"] - r["residual_initial"]) <= 1e-12 * abs(r["residual_initial"])
    np.testing.assert_allclose(g["frame_poses"], r["frame_poses"], rtol=0, atol=2e-6)
    np.testing.assert_allclose(g["points"], r["points"], rtol=0, atol=2e-5)
    np.testing.assert_allclose(g["covariance"], r["covariance"], rtol=0, atol=1e-6 * np.abs(r["covariance"]).max())


@pytest.mark.parametrize("seed,exact", [(0, True), (1, True), (0, False), (1, False), (2, False)])
def test_window_matches_oracle(solver, seed, exact):
    w = MC.mono_window(seed, exact=exact)
    check(solver.solve(w), O.mono_solve(w))


def test_batch_of_uneven_windows_matches_single_solves(solver):
    ws = [MC.mono_window(10 + k, num_frames=4 + k, num_tracks=80 + 40 * k) for k in range(6)]
    got = solver.solve_batch(ws)
    for w, g in zip(ws, got):
        check(g, O.mono_solve(w))


def test_window_without_usable_observations_is_infeasible(solver):
    w = MC.mono_window(3, num_tracks=40)
    # move every point behind the cameras
    w["points"] = (w["points"] * np.array([1, 1, -1], np.float32)).astype(np.float32)
    g = solver.solve(w)
    r = O.mono_solve(w)
    assert not g["feasible"] and not r["feasible"] and g["status"] == _lib.SBA_E_INFEASIBLE
    assert g["iterations"] == r["iterations"] == 0
    np.testing.assert_allclose(g["covariance"], np.eye(6))
