"""The C++ drop-in MonoSBASolver (include/sba/mono_sba_solver.h) run by a plain g++ program of the reference's API,
tests/cpp/dropin_mono.cpp, as a child process: addTrack / addFrame / solve / getters on a window, against the FP64
oracle's solve of the same window."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import mono_cases as MC  # noqa: E402
from oracle import sba_oracle as O  # noqa: E402
from test_mono_host import write_window  # noqa: E402

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "dropin_mono")


@pytest.mark.parametrize("seed,exact", [(0, True), (1, False)])
def test_dropin_class_matches_oracle(tmp_path, seed, exact):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (__graft_entry__.build() builds it)")
    w = MC.mono_window(seed, exact=exact)
    path = os.path.join(tmp_path, "w.bin")
    write_window(path, w)
    out = subprocess.run([EXE, path], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    g = json.loads(out.stdout.strip().splitlines()[-1])
    r = O.mono_solve(w)
    assert g["iterations"] == r["iterations"] and g["failed"] == r["failed"]
    assert abs(g["residual"] - r["residual"]) <= 1e-6 * r["residual"]
    poses = np.array(g["poses"]).reshape(-1, 4, 4).transpose(0, 2, 1)
    np.testing.assert_allclose(poses, r["frame_poses"], rtol=0, atol=2e-6)
    np.testing.assert_allclose(np.array(g["points"]).reshape(-1, 3), r["points"], rtol=0, atol=2e-5)
    np.testing.assert_allclose(np.array(g["cov"]).reshape(6, 6), r["covariance"], rtol=0,
                               atol=1e-6 * np.abs(r["covariance"]).max())
