"""GPU test of sba_solve_noalloc, the solve whose device memory is the library's fixed scratch array (no device
allocation): the same results as sba_solve on two calls in a row, the second on a smaller window (the scratch keeps
the first call's contents, so everything that must start at zero is cleared), and a clear capacity error for a
window that does not fit."""
import ctypes as C

import numpy as np
import pytest

import cuvslam_amd as ca
from cuvslam_amd import _lib
from cuvslam_amd.sba import _result_for

pytestmark = pytest.mark.gpu


def solve_noalloc(solver, problem):
    fn = solver._L.sba_solve_noalloc
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.POINTER(_lib.ProblemView), C.POINTER(_lib.Result)]
    v = problem.view()
    r, pts, poses = _result_for(problem)
    rc = fn(solver._h, C.byref(v), C.byref(r))
    return rc, r, pts, poses.transpose(0, 2, 1)


def test_noalloc_matches_solve_on_two_calls_the_second_smaller():
    big, _, _ = ca.scene.generate(seed=21, num_points=3000, num_keyframes=14, num_fixed_keyframes=2, max_iterations=8)
    small, _, _ = ca.scene.generate(seed=22, num_points=600, num_keyframes=6, num_fixed_keyframes=2, max_iterations=8)
    assert small.num_observations < big.num_observations // 3
    solver = ca.SchurComplementBundlerGpu(loss=ca.LOSS_HUBER)
    for p in (big, small):
        want = p.copy()
        assert solver.solve(want)
        rc, r, pts, poses = solve_noalloc(solver, p)
        assert rc == _lib.SBA_OK, solver.last_error()
        assert (r.status, r.iterations, r.accepted) == (_lib.SBA_OK, want.iterations, want.accepted)
        # the same kernels over the same layout with fixed-order reductions: the results are sba_solve's bit for bit
        assert (r.initial_cost_f64, r.last_cost_f64, r.lambda_) == (want.initial_cost_f64, want.last_cost_f64,
                                                                    want.final_lambda)
        np.testing.assert_array_equal(pts, want.points)
        np.testing.assert_array_equal(poses, want.rig_from_world)


def test_noalloc_rejects_a_window_beyond_its_scratch():
    p, _, _ = ca.scene.generate(seed=23, num_points=40000, num_keyframes=40, num_fixed_keyframes=2, max_iterations=2)
    solver = ca.SchurComplementBundlerGpu(loss=ca.LOSS_HUBER)
    rc, r, _, _ = solve_noalloc(solver, p)
    assert rc == _lib.SBA_E_CAPACITY
    assert "sba_solve_noalloc holds at most" in solver.last_error()
    # the handle still solves through its own workspace, and the scratch path still works after the refusal
    q, _, _ = ca.scene.generate(seed=24, num_points=500, num_keyframes=6, num_fixed_keyframes=2, max_iterations=4)
    want = q.copy()
    assert solver.solve(want)
    rc, r, _, _ = solve_noalloc(solver, q)
    assert rc == _lib.SBA_OK and r.iterations == want.iterations
