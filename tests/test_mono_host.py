"""The mono solver body (cuvslam_amd/csrc/sba_mono_body.h, the code each GPU workgroup runs) compiled for the host and
run with one thread (tests/cpp/mono_host_check.cpp) against the FP64 oracle: the same LM trajectory (iterations,
damping, failure flags) and residuals to 1e-9 relative. No GPU."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import mono_cases as MC  # noqa: E402
from oracle import sba_oracle as O  # noqa: E402

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "mono_host_check")


def write_window(path, w):
    with open(path, "wb") as f:
        f.write(np.array([len(w["constrained"]), len(w["points"]), len(w["obs_frame"])], np.int32).tobytes())
        f.write(np.asarray(w["constrained"], np.int32).tobytes())
        f.write(np.asarray(w["frame_poses"], np.float32).transpose(0, 2, 1).tobytes())
        f.write(np.asarray(w["points"], np.float32).tobytes())
        f.write(np.asarray(w["obs_frame"], np.int32).tobytes())
        f.write(np.asarray(w["obs_track"], np.int32).tobytes())
        f.write(np.asarray(w["obs_xy"], np.float32).tobytes())
        f.write(np.int32(w.get("max_iterations", 1000)).tobytes())
        f.write(np.array([w.get("min_residual", 0.001), w.get("min_residual_speed", 0.001)], np.float32).tobytes())


@pytest.mark.parametrize("seed,exact", [(0, True), (1, False), (4, False)])
def test_solver_body_on_the_cpu_matches_oracle(tmp_path, seed, exact):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (__graft_entry__.build() builds it)")
    w = MC.mono_window(seed, exact=exact)
    path = os.path.join(tmp_path, "w.bin")
    write_window(path, w)
    out = subprocess.run([EXE, path], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    g = json.loads(out.stdout)
    r = O.mono_solve(w)
    assert (g["iterations"], bool(g["failed"]), bool(g["feasible"])) == (r["iterations"], r["failed"], r["feasible"])
    assert g["lambda"] == r["lam"]
    assert abs(g["residual"] - r["residual"]) <= 1e-9 * r["residual"]
    poses = np.array(g["poses"]).reshape(-1, 4, 4).transpose(0, 2, 1)
    np.testing.assert_allclose(poses, r["frame_poses"], rtol=0, atol=2e-6)
    np.testing.assert_allclose(np.array(g["points"]).reshape(-1, 3), r["points"], rtol=0, atol=2e-5)
    C = np.array(g["cov"]).reshape(6, 6)
    np.testing.assert_allclose(C, r["covariance"], rtol=0, atol=1e-9 * np.abs(r["covariance"]).max())
