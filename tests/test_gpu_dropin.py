"""The C++ drop-in class (include/sba/schur_complement_bundler_gpu.h) run by a plain g++ program of the reference's
API, tests/cpp/dropin_solve.cpp, as a child process: its solve of a synthetic window against the FP64 oracle, with
the reference bundler's defaults (Student-t or Huber loss, the early exit on)."""
import json
import os
import subprocess

import numpy as np
import pytest

import cuvslam_amd as ca
from oracle import sba_oracle as O

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "dropin_solve")


@pytest.mark.parametrize("seed,loss", [(1, 1), (2, 0)])
def test_dropin_class_matches_oracle(seed, loss):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (__graft_entry__.build() builds it)")
    out = subprocess.run([EXE, str(seed), str(loss)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    got = json.loads(out.stdout.strip().splitlines()[-1])
    # the same window in Python: the program's scene (400 points, 6 key frames, 2 fixed, 10 iterations)
    p, _, _ = ca.scene.generate(seed=seed, num_points=400, num_keyframes=6, num_fixed_keyframes=2, max_iterations=10)
    ref = O.solve(p, "f64", loss=loss)
    assert got["ok"] is True
    assert got["iterations"] == ref["iterations"]
    assert abs(got["initial_cost"] - ref["initial_cost"]) <= 1e-6 * ref["initial_cost"]
    assert abs(got["last_cost"] - ref["last_cost"]) <= 1e-6 * ref["last_cost"]
    np.testing.assert_allclose(got["p0"], ref["points"][0], rtol=0, atol=1e-4)
