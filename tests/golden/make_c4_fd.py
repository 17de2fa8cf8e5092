"""Writes tests/golden/c4_fd.json: FP64 central finite differences of the differentiable-solve loss on the
20 key frame / 10,000 point window (seed 4, Student-t, no outliers), for the GPU gradient test and the
benchmark's gradient check.

The loss is L = sum_j |T_j - G_j|^2 over the free key frames after convergence (tests/adjoint_fd.py). Each
difference quotient solves the perturbed window with the FP64 oracle from the same initial state (30 LM
iterations, then two full-Hessian Newton steps); steps are 1e-3 px for pixels and 1e-4 for the intrinsics.
The sampled observation components are those of `Case.sampled` (numpy default_rng(1)). It takes a few
minutes on one core.

Run:  python tests/golden/make_c4_fd.py
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402

import cuvslam_amd as ca  # noqa: E402
from adjoint_fd import C4_SCENE, Case  # noqa: E402


def main():
    t0 = time.time()
    c = Case(**C4_SCENE)
    fd_uv, fd_K = c.fd
    out = {
        "scene": {k: (v if not isinstance(v, float) else float(v)) for k, v in C4_SCENE.items()},
        "num_observations": c.problem.num_observations,
        "sampled": c.sampled,
        "fd_pixels": fd_uv.tolist(),
        "fd_intrinsics": fd_K.tolist(),
        "loss_at_optimum": c.L(c.optimum[1]),
    }
    with open(os.path.join(HERE, "c4_fd.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote c4_fd.json in {time.time() - t0:.0f}s")


if __name__ == "__main__":
    main()
