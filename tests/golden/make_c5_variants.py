"""Writes tests/golden/c5_variants.json: the FP64 oracle's solve of the 200-key-frame multi-camera map (seed 5,
10 LM iterations, Huber), the FP64 solves that take the other side of each damping threshold its gain ratio comes
within 0.01 of (tests/sensitivity.py: branch_variants), the spread of FP64 solves under one-ulp perturbations of the
FP32 inputs, the distance of the oracle's FP32 restatement of the method (fp32_input_sensitivity,
fp32_method_spread), and the same FP32 restatement with the observations in three other orders
(fp32_order_variant: distance and accepted steps). The oracle takes about 25 s per solve, so the results are kept
here rather than recomputed.

Run:  python tests/golden/make_c5_variants.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

import cuvslam_amd as ca  # noqa: E402
from sensitivity import branch_variants, fp32_input_sensitivity, fp32_method_spread, fp32_order_variant  # noqa: E402


def main():
    p, _, _ = ca.scene.loop(seed=5, max_iterations=10)
    variants = branch_variants(p, ca.LOSS_HUBER, fixed_iterations=1)
    spread, _ = fp32_input_sensitivity(p, ca.LOSS_HUBER, trials=2, reference=variants[0][1])
    spread32, _, r32 = fp32_method_spread(p, ca.LOSS_HUBER, reference=variants[0][1])
    out = {"generator": "tests/golden/make_c5_variants.py", "scene": "ca.scene.loop(seed=5, max_iterations=10)",
           "observations": p.num_observations, "fp32_input_spread": spread, "fp32_method_spread": spread32,
           "fp32_method_accepted": int(r32["accepted"]),
           "fp32_order_variants": [{"seed": seed, "spread": sp, "accepted": acc} for seed, (sp, acc) in
                                   ((sd, fp32_order_variant(p, ca.LOSS_HUBER, sd, variants[0][1])) for sd in (1, 2, 3))],
           "trace_rho": [float(x) for x in variants[0][1]["trace_rho"]],
           "trace_flags": [int(x) for x in variants[0][1]["trace_flags"]],
           "variants": [{"flip_iteration": f, "last_cost": r["last_cost"], "iterations": r["iterations"],
                         "accepted": r["accepted"], "lambda": r["lam"]} for f, r in variants]}
    with open(os.path.join(HERE, "c5_variants.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
