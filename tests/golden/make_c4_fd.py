"""Writes tests/golden/c4_fd.json: FP64 central finite differences of the differentiable-solve loss on the
20 key frame / 10,000 point window (seed 4, Student-t, no outliers), for the GPU gradient test and the
benchmark's gradient check.

The loss is L = sum_j |T_j - G_j|^2 over the free key frames after convergence (tests/adjoint_fd.py). Each
difference quotient solves the perturbed window with the FP64 oracle from the same initial state (50 LM
iterations, then two full-Hessian Newton steps); steps are 1e-3 px for pixels and 1e-4 for the intrinsics.
Both pixel components of 60 observations drawn with numpy default_rng(1) are differenced (120 components), and
the four intrinsics. The 248 solves run in a pool of worker processes (about 10 minutes on 8 cores).

Run:  python tests/golden/make_c4_fd.py [workers]
"""
import json
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402

NUM_OBSERVATIONS = 60  # both pixel components of each: 120 differenced components
H_PX, H_K = 1e-3, 1e-4

_case = None


def _get_case():
    global _case
    if _case is None:
        from adjoint_fd import C4_SCENE, Case
        _case = Case(**C4_SCENE)
    return _case


def _loss_at(task):
    """L at one perturbed input: ("uv", o, c, sign) or ("K", k, sign)."""
    from adjoint_fd import K0
    c = _get_case()
    if task[0] == "uv":
        _, o, comp, sign = task
        uv = c.uv.copy()
        uv[o, comp] += sign * H_PX
        return c.L(c.converged(uv, K0)[1])
    _, k, sign = task
    K = K0.copy()
    K[k] += sign * H_K
    return c.L(c.converged(c.uv, K)[1])


def main():
    workers = int(sys.argv[1]) if len(sys.argv) > 1 else min(8, os.cpu_count() or 1)
    t0 = time.time()
    c = _get_case()
    rng = np.random.default_rng(1)
    sampled = [(int(o), comp) for o in rng.choice(c.problem.num_observations, NUM_OBSERVATIONS, replace=False)
               for comp in range(2)]
    tasks = [("uv", o, comp, s) for o, comp in sampled for s in (1, -1)] + [("K", k, s) for k in range(4) for s in (1, -1)]
    with ProcessPoolExecutor(max_workers=workers) as ex:
        L = list(ex.map(_loss_at, tasks, chunksize=1))
    n = 2 * len(sampled)
    fd_uv = [(L[2 * k] - L[2 * k + 1]) / (2 * H_PX) for k in range(len(sampled))]
    fd_K = [(L[n + 2 * k] - L[n + 2 * k + 1]) / (2 * H_K) for k in range(4)]
    from adjoint_fd import C4_SCENE
    out = {
        "scene": {k: (v if not isinstance(v, float) else float(v)) for k, v in C4_SCENE.items()},
        "num_observations": c.problem.num_observations,
        "sampled": sampled,
        "fd_pixels": fd_uv,
        "fd_intrinsics": fd_K,
        "loss_at_optimum": c.L(c.optimum[1]),
    }
    with open(os.path.join(HERE, "c4_fd.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote c4_fd.json ({len(sampled)} pixel components) in {time.time() - t0:.0f}s")


if __name__ == "__main__":
    main()
