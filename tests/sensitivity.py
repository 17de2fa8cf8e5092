"""FP64 sensitivity of a solve to the rounding of its FP32 inputs, and to its damping decisions (test infrastructure).

The solver reads FP32 inputs and computes its bulk in FP32, so its result can differ from the FP64 oracle's by
what the oracle's own trajectory does under perturbations of the size of FP32 rounding. This measures that
directly: the FP64 oracle is rerun on copies of the problem whose FP32 inputs (points, observation coordinates,
pose translations) are each moved by one FP32 unit in the last place in a random direction, and the spread of
the final costs is reported relative to the unperturbed FP64 solve. A window whose spread is large has a
trajectory that amplifies last-bit differences (few key frames, points seen from two or three of them, not
converged): there the distance between the FP32 path and the FP64 oracle is explained by conditioning and is
bounded by a multiple of this spread.

The spread stays at the 1e-8 level on the benchmark windows (5e-8 on the 200-key-frame map), and so does an FP32
rounding of the finished reduced system (the oracle's round_reduced_f32 option: 6e-10 on the map): neither explains
a distance of 1e-6 or more. Two things do:
  * a damping decision: a gain ratio within about 1e-3 of 0.25 or 0.75 can land on the other side of the threshold
    in FP32, and from there the two solvers follow different but equally valid LM trajectories (lambda 8x or 40x
    apart). branch_variants() gives the FP64 trajectories of every such flip, and a solver's result is compared
    with the nearest of them;
  * FP32 arithmetic itself on an ill-conditioned window: its accumulated rounding moves each step at small damping
    by far more than a final rounding does. fp32_method_spread() measures it with an independent FP32
    implementation of the same method, the oracle's FP32 restatement (block-sparse Schur complement, Cholesky): how
    far its final cost lands from the FP64 oracle's. A solver whose own FP32 arithmetic (another summation order)
    lands within a small multiple of that distance is as close as FP32 allows on that window.
"""
import numpy as np

from oracle import sba_oracle as O


def nudge(problem, rng):
    """A copy with every point coordinate, observation coordinate and pose translation moved by one FP32 ulp."""
    q = problem.copy()

    def ulp(a):
        a = np.asarray(a, np.float32)
        up = rng.integers(0, 2, a.shape).astype(bool)
        return np.where(up, np.nextafter(a, np.float32(np.inf)), np.nextafter(a, np.float32(-np.inf))).astype(np.float32)

    q.points = ulp(q.points)
    q.observation_xys = ulp(q.observation_xys)
    free = q.rig_from_world.shape[0] - q.num_fixed_key_frames
    q.rig_from_world[:free, :3, 3] = ulp(q.rig_from_world[:free, :3, 3])
    return q


def fp32_input_sensitivity(problem, loss, trials=3, seed=0, fixed_iterations=1, reference=None):
    """(max over trials of |F_perturbed - F| / F, the unperturbed FP64 result) for the FP64 oracle's solve."""
    ref = reference if reference is not None else O.solve(problem, "f64", loss=loss, fixed_iterations=fixed_iterations)
    rng = np.random.default_rng(seed)
    spread = 0.0
    for _ in range(trials):
        r = O.solve(nudge(problem, rng), "f64", loss=loss, fixed_iterations=fixed_iterations)
        spread = max(spread, abs(r["last_cost"] - ref["last_cost"]) / ref["last_cost"])
    return spread, ref


def branch_variants(problem, loss, margin=0.01, fixed_iterations=1, reference=None, **opt):
    """The FP64 oracle's solve and, for every LM iteration whose gain ratio lies within `margin` of 0.25 or 0.75,
    the FP64 solve that takes the other side of that threshold there. Returns [(flip iteration (0: none), result)]."""
    ref = reference if reference is not None else O.solve(problem, "f64", loss=loss, fixed_iterations=fixed_iterations,
                                                          **opt)
    out = [(0, ref)]
    ok = (ref["trace_flags"] & 6) == 0  # a factorization ran and no stop test ended the iteration
    for k in range(ref["iterations"]):
        rho = ref["trace_rho"][k]
        if ok[k] and min(abs(rho - 0.25), abs(rho - 0.75)) < margin:
            out.append((k + 1, O.solve(problem, "f64", loss=loss, fixed_iterations=fixed_iterations,
                                       flip_iteration=k + 1, **opt)))
    return out


def nearest_variant(variants, cost):
    """(relative cost distance, flip iteration, result) of the variant whose final cost is nearest `cost`."""
    best = min(variants, key=lambda v: abs(cost - v[1]["last_cost"]))
    return abs(cost - best[1]["last_cost"]) / best[1]["last_cost"], best[0], best[1]


def fp32_method_spread(problem, loss, fixed_iterations=1, reference=None):
    """(|F32 - F64| / F64, FP64 result, FP32 result): the final cost of the oracle's FP32 restatement of the method
    (its FP32 state re-evaluated in FP64) against the FP64 oracle's."""
    ref = reference if reference is not None else O.solve(problem, "f64", loss=loss, fixed_iterations=fixed_iterations)
    r32 = O.solve(problem, "f32", loss=loss, fixed_iterations=fixed_iterations, rcs_solver=0, dense_schur=0)
    q = problem.copy()
    q.points = r32["points"].astype(np.float32)
    q.rig_from_world = r32["poses"].astype(np.float32)
    c, _ = O.cost(q, loss=loss)
    return abs(c - ref["last_cost"]) / ref["last_cost"], ref, r32


def permuted(problem, seed):
    """A copy whose observations are listed in a random order (the same problem; every sum over observations is
    taken in another order)."""
    q = problem.copy()
    perm = np.random.default_rng(seed).permutation(q.point_ids.shape[0])
    q.observation_xys = q.observation_xys[perm].copy()
    q.observation_infos = q.observation_infos[perm].copy()
    q.point_ids = q.point_ids[perm].copy()
    q.pose_ids = q.pose_ids[perm].copy()
    q.camera_ids = q.camera_ids[perm].copy()
    return q


def fp32_order_variant(problem, loss, seed, reference, fixed_iterations=1):
    """(|F32 - F64| / F64, accepted steps) of the oracle's FP32 restatement on the problem with its observations in
    another order: how far, and along which damping decisions, FP32 rounding alone moves the method."""
    q = permuted(problem, seed)
    r32 = O.solve(q, "f32", loss=loss, fixed_iterations=fixed_iterations, rcs_solver=0, dense_schur=0)
    q.points = r32["points"].astype(np.float32)
    q.rig_from_world = r32["poses"].astype(np.float32)
    c, _ = O.cost(q, loss=loss)
    return abs(c - reference["last_cost"]) / reference["last_cost"], int(r32["accepted"])
