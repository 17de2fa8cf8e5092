"""GPU parity: the HIP path through the C ABI against the FP64 oracle on the same seeded inputs."""
import os
import sys

import numpy as np
import pytest

import cuvslam_amd as ca
from oracle import sba_oracle as O

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.gpu

HUBER, STUDENT = ca.LOSS_HUBER, ca.LOSS_STUDENT_T


def small_scene(seed=1, n_points=300, n_kf=6, nf=2, **kw):
    p, g, q = ca.scene.generate(seed=seed, num_points=n_points, num_keyframes=n_kf, num_fixed_keyframes=nf, **kw)
    return p, g, q


def test_mfma_lane_layout(solver_huber):
    # the v_mfma_f32_16x16x4_f32 lane layout the Schur kernel relies on, checked against a plain product
    import ctypes as C
    rng = np.random.default_rng(0)
    A = rng.standard_normal((16, 4)).astype(np.float32)
    B = rng.standard_normal((4, 16)).astype(np.float32)
    D = np.zeros((16, 16), np.float32)
    f = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
    assert solver_huber._L.sba_debug_mfma_layout(solver_huber._h, f(A), f(B), f(D)) == 0
    np.testing.assert_allclose(D, A.astype(np.float64) @ B.astype(np.float64), rtol=1e-5, atol=1e-5)


def test_factor_known_answer(solver_huber):
    # The reference's Cholesky known answer (libs/cuda_modules/test/sba_test.cpp:700-719): A x = b with
    # A = [[4,1,0],[1,3,1],[0,1,2]], b = (6,10,8) -> x = (1,2,3); embedded in a 6x6 block-diagonal system.
    A = np.eye(6)
    A[:3, :3] = [[4, 1, 0], [1, 3, 1], [0, 1, 2]]
    b = np.array([6, 10, 8, 0, 0, 0], float)
    x, rc = solver_huber.debug_factor(A, b)
    assert rc == 0
    np.testing.assert_allclose(x, [1, 2, 3, 0, 0, 0], atol=1e-12)
    # indefinite: leading 2x2 minor negative -> factorization fails; the next call recovers
    B = np.eye(6)
    B[:3, :3] = [[1, 2, 0], [2, 1, 0], [0, 0, 1]]
    _, rc = solver_huber.debug_factor(B, b)
    assert rc == ca.sba._lib.SBA_E_INFEASIBLE
    x, rc = solver_huber.debug_factor(A, b)
    assert rc == 0
    np.testing.assert_allclose(x, [1, 2, 3, 0, 0, 0], atol=1e-12)


@pytest.mark.parametrize("n", [6, 54, 174, 186, 192, 234, 240])
def test_factor_random_spd(solver_huber, n):
    rng = np.random.default_rng(n)
    M = rng.standard_normal((n, n))
    A = M @ M.T + n * np.eye(n)
    b = rng.standard_normal(n)
    x, rc = solver_huber.debug_factor(A, b)
    assert rc == 0
    ref = np.linalg.solve(A, b)
    # A and b pass through the FP32 accumulator of the reduced system, so compare against the FP32-rounded system
    A32 = A.astype(np.float32).astype(np.float64)
    b32 = b.astype(np.float32).astype(np.float64)
    ref32 = np.linalg.solve(np.tril(A32) + np.tril(A32, -1).T, b32)
    np.testing.assert_allclose(x, ref32, rtol=1e-10, atol=1e-12)
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-5


@pytest.mark.parametrize("loss", [HUBER, STUDENT])
@pytest.mark.parametrize("lam", [1e-3, 1e-1])
def test_debug_step_matches_oracle(loss, lam):
    p, _, _ = small_scene(seed=3)
    solver = ca.SchurComplementBundlerGpu(loss=loss)
    b = solver.batch([p])
    S, e, dc, dp, sc, rc = b.debug_step(lam)
    So, eo, dco, dpo, sco, rco = O.step(p, lam, loss=loss)
    assert rc == 0 and rco == 0
    # FP32 linearization and Schur accumulation against FP64
    assert np.linalg.norm(S - So) / np.linalg.norm(So) < 2e-5
    assert np.linalg.norm(e - eo) / np.linalg.norm(eo) < 1e-4
    assert np.linalg.norm(dc - dco) / np.linalg.norm(dco) < 1e-3
    assert np.linalg.norm(dp - dpo) / np.linalg.norm(dpo) < 1e-3
    # trial cost in FP64 from the FP32 trial state; predicted-reduction terms
    assert abs(sc[0] - sco[0]) / sco[0] < 1e-5
    assert sc[1] == sco[1]
    assert abs(sc[2] - sco[2]) / abs(sco[2]) < 1e-3
    assert abs(sc[3] - sco[3]) / abs(sco[3]) < 1e-3
    assert abs(sc[6] - sco[6]) / sco[6] < 1e-6


@pytest.mark.parametrize("loss", [HUBER, STUDENT])
def test_solve_matches_oracle_small(loss):
    p, _, _ = small_scene(seed=5, n_points=500, n_kf=8, max_iterations=10)
    ref = O.solve(p, "f64", loss=loss)
    solver = ca.SchurComplementBundlerGpu(loss=loss)
    q = p.copy()
    assert solver.solve(q)
    assert q.iterations == ref["iterations"]
    assert q.accepted == ref["accepted"]
    assert abs(q.initial_cost_f64 - ref["initial_cost"]) / ref["initial_cost"] < 1e-7
    assert abs(q.last_cost_f64 - ref["last_cost"]) / ref["last_cost"] < 1e-6
    # the FP32 state the GPU returns, re-evaluated by the oracle in FP64
    c, _ = O.cost(q, loss=loss)
    assert abs(c - ref["last_cost"]) / ref["last_cost"] < 1e-6


def test_solve_c1_huber():
    p, _, _ = ca.scene.generate("c1", seed=1, num_fixed_keyframes=2, max_iterations=10)
    ref = O.solve(p, "f64", loss=HUBER)
    solver = ca.SchurComplementBundlerGpu(loss=HUBER)
    q = p.copy()
    assert solver.solve(q)
    assert q.iterations == ref["iterations"] == 10
    assert abs(q.last_cost_f64 - ref["last_cost"]) / ref["last_cost"] < 1e-6


def test_near_plane_all_skipped_is_infeasible(solver_huber):
    # every point between 0.02 and 0.08 m in front of the cameras: every observation is skipped, the start
    # is infeasible and solve() returns false (schur_complement_bundler_gpu.cpp:165-168; sba_test.cpp:1402-1462)
    p, _, _ = small_scene(seed=7)
    p.rig_from_world[:] = np.eye(4, dtype=np.float32)
    p.points[:, 2] = np.linspace(0.02, 0.08, p.points.shape[0], dtype=np.float32)
    p.points[:, :2] *= 0.001
    q = p.copy()
    assert not solver_huber.solve(q)
    assert q.status == ca.sba._lib.SBA_E_INFEASIBLE
    assert np.isinf(q.initial_cost)


def test_near_field_observations_carry_weight(solver_huber):
    # depths 0.3-0.8 m are in front of the near plane: kept, finite cost (sba_test.cpp:1361-1396)
    p, _, _ = small_scene(seed=9)
    p.rig_from_world[:] = np.eye(4, dtype=np.float32)
    p.points[:, 2] = np.linspace(0.3, 0.8, p.points.shape[0], dtype=np.float32)
    p.points[:, :2] *= 0.01
    q = p.copy()
    assert solver_huber.solve(q)
    assert np.isfinite(q.initial_cost)


def test_empty_problem_returns_false(solver_huber):
    p, _, _ = small_scene(seed=8)
    q = p.copy()
    q.observation_xys = q.observation_xys[:0]
    q.observation_infos = q.observation_infos[:0]
    q.point_ids = q.point_ids[:0]
    q.pose_ids = q.pose_ids[:0]
    q.camera_ids = q.camera_ids[:0]
    assert not solver_huber.solve(q)


def max_points_per_back_substitution_tile(p, max_pairs=256, max_points=256):
    """Largest number of points in one back-substitution tile: consecutive points holding at most 256
    (point, key frame) pairs and at most 256 points."""
    key = np.unique(np.asarray(p.point_ids).astype(np.int64) * 4096 + np.asarray(p.pose_ids))
    pairs = np.bincount(key // 4096, minlength=p.points.shape[0])
    best, i, n = 0, 0, len(pairs)
    while i < n:
        b, k = i, 0
        while i < n and i - b < max_points and (k + pairs[i] <= max_pairs or i == b):
            k += pairs[i]
            i += 1
        best = max(best, i - b)
    return best


def test_back_substitution_tiles_of_many_points_match_oracle():
    # six key frames: points are seen from at most six of them, so a back-substitution tile holds about 50 points
    # (far more than a Schur tile's 16), and every point's camera coupling sum_j W_ij^T dc_j must reach its step.
    # Every point is seen from all the key frames it projects in: the window is well constrained, so the FP32 path
    # stays within 1e-7 of the FP64 oracle and the 1e-6 bound tests the step, not rounding.
    p, _, _ = small_scene(seed=42, n_points=1500, n_kf=6, max_iterations=10)
    assert max_points_per_back_substitution_tile(p) > 32
    ref = O.solve(p, "f64", loss=HUBER)
    q = p.copy()
    assert ca.SchurComplementBundlerGpu(loss=HUBER).solve(q)
    assert q.iterations == ref["iterations"]
    assert q.accepted == ref["accepted"]
    assert abs(q.last_cost_f64 - ref["last_cost"]) / ref["last_cost"] < 1e-6
    c, _ = O.cost(q, loss=HUBER)
    assert abs(c - ref["last_cost"]) / ref["last_cost"] < 1e-6


def test_batch_matches_single(solver_huber):
    probs = [small_scene(seed=20 + k, n_points=200 + 100 * k, n_kf=5 + 3 * k)[0] for k in range(4)]
    singles = []
    for p in probs:
        q = p.copy()
        assert solver_huber.solve(q)
        singles.append(q)
    batch = [p.copy() for p in probs]
    assert all(solver_huber.solve_batch(batch))
    for a, s in zip(batch, singles):
        assert a.iterations == s.iterations
        assert abs(a.last_cost_f64 - s.last_cost_f64) / s.last_cost_f64 < 1e-6


def test_every_window_of_an_uneven_batch_matches_oracle():
    # windows of 5 to 40 key frames, points seen from two key frames up to all of them, outliers up to 10%, exactly 8
    # iterations: each window of the batch against its own FP64 oracle solve at 1e-6. A window whose FP64 gain ratio
    # comes within 0.01 of a damping threshold (here window 9: 0.7512 at iteration 3) is compared with the nearest
    # of the FP64 trajectories that take either side of it (tests/sensitivity.py).
    from sensitivity import branch_variants, nearest_variant
    probs = ca.scene.batched_windows(10, seed=5, max_iterations=8, pt_range=(300, 2500))
    qs = [p.copy() for p in probs]
    assert all(ca.SchurComplementBundlerGpu(loss=HUBER, fixed_iterations=True).solve_batch(qs))
    flips = []
    for p, q in zip(probs, qs):
        err, flip, ref = nearest_variant(branch_variants(p, HUBER, fixed_iterations=1), q.last_cost_f64)
        assert err < 1e-6, (err, flip)
        assert (q.iterations, q.accepted) == (ref["iterations"], ref["accepted"])
        c, _ = O.cost(q, loss=HUBER)
        assert abs(c - q.last_cost_f64) / ref["last_cost"] < 1e-9  # the returned state has the reported cost
        flips.append(flip)
    assert sum(f != 0 for f in flips) <= 2, flips


def test_autotuner_and_chunking_leave_results_unchanged():
    # the work granularity changes only the order of the FP32 record sums; after 5 LM iterations (not yet
    # converged) that rounding difference shows in the cost at the 1e-6 level
    probs = ca.scene.batched_windows(6, seed=7, max_iterations=5)
    s = ca.SchurComplementBundlerGpu(loss=ca.LOSS_HUBER, fixed_iterations=True)
    best, ms = s.autotune(probs, [512, 2048, 8192], repeats=1)
    assert best in (512, 2048, 8192) and len(ms) == 3 and all(m > 0 for m in ms)
    costs = []
    for cp in (512, 8192):
        qs = [p.copy() for p in probs]
        sv = ca.SchurComplementBundlerGpu(loss=ca.LOSS_HUBER, fixed_iterations=True, chunk_pairs=cp)
        assert all(sv.solve_batch(qs))
        costs.append(np.array([q.last_cost_f64 for q in qs]))
    np.testing.assert_allclose(costs[0], costs[1], rtol=1e-5)


def test_large_multicamera_window_matches_oracle():
    # 59 free key frames (reduced system of order 354, above the 40 key frames one chunk stages): the mapped
    # path (per-chunk key frame maps, chunk-ordered FP64 gather, blocked FP64 Cholesky over several workgroups)
    p, _, _ = ca.scene.loop(num_keyframes=60, num_points=20000, loop_length_m=12.0, max_iterations=10)
    assert p.rig_from_world.shape[0] - p.num_fixed_key_frames > 40 and p.camera_from_rig.shape[0] == 8
    q = p.copy()
    s = ca.SchurComplementBundlerGpu(loss=ca.LOSS_HUBER, fixed_iterations=True)
    assert s.solve(q)
    # the gather sums every block in chunk order: a second solve gives the same bits
    q2 = p.copy()
    assert s.solve(q2)
    assert np.array_equal(q.rig_from_world, q2.rig_from_world) and np.array_equal(q.points, q2.points)
    assert q.last_cost_f64 == q2.last_cost_f64
    r = O.solve(p, "f64", loss=ca.LOSS_HUBER, fixed_iterations=1)
    assert q.iterations == r["iterations"] == 10
    assert abs(q.last_cost_f64 - r["last_cost"]) / r["last_cost"] < 1e-6
    nf = p.rig_from_world.shape[0] - 1
    assert np.abs(q.rig_from_world[:nf, :3, 3] - r["poses"][:nf, :3, 3]).max() < 1e-4
    # the window's points are packed in the order of their first key frame; they come back in the caller's order
    assert np.median(np.linalg.norm(q.points - r["points"], axis=1)) < 1e-5


def _add_far_points(p, gt_poses, n, seed):
    """p with n more points 30-60 m out in the plane of the loop, observed (without noise) by every camera of every
    key frame that sees them in front within |x/z|, |y/z| < 1; on the multi-camera loop each is seen from nearly
    every key frame."""
    rng = np.random.default_rng(seed)
    centers = np.array([-(q[:3, :3].T @ q[:3, 3]) for q in gt_poses.astype(np.float64)])
    mid = centers.mean(0)
    up = np.argmin(centers.var(0))  # the loop's vertical axis: the one its key frames vary least along
    flat = [k for k in range(3) if k != up]
    ang = rng.uniform(0, 2 * np.pi, n)
    rad = rng.uniform(30, 60, n)
    X = np.tile(mid, (n, 1))
    X[:, flat[0]] += rad * np.cos(ang)
    X[:, flat[1]] += rad * np.sin(ang)
    X[:, up] += rng.uniform(-1, 1, n)
    pid, qid, cid, xy = [], [], [], []
    base = p.points.shape[0]
    for i in range(n):
        for j in range(gt_poses.shape[0]):
            for cam in range(p.camera_from_rig.shape[0]):
                T = p.camera_from_rig[cam].astype(np.float64) @ gt_poses[j].astype(np.float64)
                Xc = T[:3, :3] @ X[i] + T[:3, 3]
                if Xc[2] > 0.5 and abs(Xc[0] / Xc[2]) < 1 and abs(Xc[1] / Xc[2]) < 1:
                    pid.append(base + i)
                    qid.append(j)
                    cid.append(cam)
                    xy.append(Xc[:2] / Xc[2])
    q = p.copy()
    q.points = np.concatenate([p.points, (X + rng.normal(0, 0.05, X.shape)).astype(np.float32)])
    q.observation_xys = np.concatenate([p.observation_xys, np.array(xy, np.float32)])
    q.observation_infos = np.concatenate([p.observation_infos, np.repeat(p.observation_infos[:1], len(xy), 0)])
    q.point_ids = np.concatenate([p.point_ids, np.array(pid, np.int32)])
    q.pose_ids = np.concatenate([p.pose_ids, np.array(qid, np.int32)])
    q.camera_ids = np.concatenate([p.camera_ids, np.array(cid, np.int8)])
    return q


def test_large_window_with_points_seen_from_more_than_40_key_frames():
    # the mapped window above plus 20 far points seen from (nearly) all 59 free key frames: each such point is
    # staged by tiles of its own (its key frames in groups of at most 20, every two groups together)
    p0, _, gt_poses = ca.scene.loop(num_keyframes=60, num_points=20000, loop_length_m=12.0, max_iterations=10)
    p = _add_far_points(p0, gt_poses, 20, seed=3)
    nf = p.rig_from_world.shape[0] - p.num_fixed_key_frames
    new = p.point_ids >= p0.points.shape[0]
    seen = [np.unique(p.pose_ids[new & (p.point_ids == i)]) for i in range(p0.points.shape[0], p.points.shape[0])]
    assert sum((s < nf).sum() > 40 for s in seen) >= 15
    q = p.copy()
    s = ca.SchurComplementBundlerGpu(loss=ca.LOSS_HUBER, fixed_iterations=True)
    assert s.solve(q)
    r = O.solve(p, "f64", loss=ca.LOSS_HUBER, fixed_iterations=1)
    assert q.iterations == r["iterations"] == 10
    assert abs(q.last_cost_f64 - r["last_cost"]) / r["last_cost"] < 1e-6
    assert np.abs(q.rig_from_world[:nf, :3, 3] - r["poses"][:nf, :3, 3]).max() < 1e-4
    assert np.median(np.linalg.norm(q.points - r["points"], axis=1)) < 1e-5
    q2 = p.copy()
    assert s.solve(q2)
    assert np.array_equal(q.rig_from_world, q2.rig_from_world) and np.array_equal(q.points, q2.points)


def test_window_of_300_key_frames_matches_oracle():
    # more key frames than a 256-entry table would hold: 299 free key frames (reduced system of order 1794); the
    # key frame index of a pair record and the back-substitution's key frame staging are sized for 1024. The
    # 60 m loop with one fixed key frame is weakly constrained along it: the oracle's FP32 restatement of the
    # method ends 5.7e-7 from the FP64 solve (with 3% outliers 1e-5), so the bound is ten times that distance,
    # computed here as for the 200-key-frame map.
    p, _, _ = ca.scene.loop(num_keyframes=300, num_points=40000, loop_length_m=60.0, max_iterations=5,
                            outlier_rate=0.0)
    assert p.rig_from_world.shape[0] == 300
    q = p.copy()
    s = ca.SchurComplementBundlerGpu(loss=ca.LOSS_HUBER, fixed_iterations=True)
    assert s.solve(q)
    r = O.solve(p, "f64", loss=ca.LOSS_HUBER, fixed_iterations=1)
    from sensitivity import fp32_method_spread
    spread, _, _ = fp32_method_spread(p, ca.LOSS_HUBER, reference=r)
    assert q.iterations == r["iterations"] == 5
    assert abs(q.last_cost_f64 - r["last_cost"]) / r["last_cost"] < max(1e-6, 10 * spread), spread
    nf = p.rig_from_world.shape[0] - p.num_fixed_key_frames
    assert np.abs(q.rig_from_world[:nf, :3, 3] - r["poses"][:nf, :3, 3]).max() < 1e-4


def test_large_window_in_a_batch_with_small_windows():
    big, _, _ = ca.scene.loop(num_keyframes=48, num_points=6000, loop_length_m=9.6, max_iterations=6)
    small = ca.scene.batched_windows(3, seed=11, max_iterations=6)
    probs = [small[0], big, small[1], small[2]]
    s = ca.SchurComplementBundlerGpu(loss=ca.LOSS_HUBER, fixed_iterations=True)
    together = [p.copy() for p in probs]
    assert all(s.solve_batch(together))
    for p, t in zip(probs, together):
        alone = p.copy()
        assert s.solve(alone)
        assert abs(alone.last_cost_f64 - t.last_cost_f64) <= 1e-6 * alone.last_cost_f64


def test_winsorize_matches_oracle_and_the_second_run_continues_without_the_reset_points():
    # between the two bundler runs (service_sba.h:317-339): winsorize the solved window, then solve again
    p, _, _ = small_scene(seed=12, n_points=1500, n_kf=8, max_iterations=8, outlier_rate=0.04)
    s = ca.SchurComplementBundlerGpu(loss=HUBER, fixed_iterations=True)
    b = s.batch([p])
    b.run()
    q = p.copy()
    b.read([q])
    state = (q.points.astype(np.float64), q.rig_from_world.astype(np.float64))
    want0, n0 = O.winsorize(p, 0, *state)
    mask0, m0 = b.winsorize(0)
    assert m0 == n0 > 0
    np.testing.assert_array_equal(mask0, want0)
    # the second run starts from the first solve's state, without the reset points' observations
    r = q.copy()
    r.observation_infos[np.isin(r.point_ids, np.nonzero(mask0)[0])] = 0
    ref = O.solve(r, "f64", loss=HUBER, fixed_iterations=1)
    b.run()
    t = p.copy()
    b.read([t])
    assert t.iterations == ref["iterations"]
    assert abs(t.initial_cost_f64 - ref["initial_cost"]) / ref["initial_cost"] < 1e-7
    assert abs(t.last_cost_f64 - ref["last_cost"]) / ref["last_cost"] < 1e-6


def test_winsorize_of_a_mapped_window_matches_oracle():
    # a window of more than 40 free key frames packs its points in another order: the reset marks come back in
    # the caller's point order
    p, _, _ = ca.scene.loop(num_keyframes=48, num_points=6000, loop_length_m=9.6, max_iterations=6)
    s = ca.SchurComplementBundlerGpu(loss=HUBER, fixed_iterations=True)
    b = s.batch([p])
    b.run()
    q = p.copy()
    b.read([q])
    state = (q.points.astype(np.float64), q.rig_from_world.astype(np.float64))
    want0, n0 = O.winsorize(p, 0, *state)
    mask0, m0 = b.winsorize(0)
    assert m0 == n0 > 0
    np.testing.assert_array_equal(mask0, want0)


def test_winsorize_needs_fifty_observations_and_checks_arguments(solver_huber):
    p, _, _ = small_scene(seed=13, n_points=40, n_kf=4, max_iterations=3)
    b = solver_huber.batch([p])
    b.run()
    mask, n = b.winsorize(0)
    assert n == 0 and not mask.any()
    with pytest.raises(ca.SbaError):
        b.winsorize(4)
    b.close()


def test_repeated_solves_are_bitwise_identical():
    # every sum of a solve has a fixed order (chunk records, per-point pair sums, per-chunk cost slots), so repeated
    # runs of a batch and repeated single-window solves return the same bits
    probs = ca.scene.batched_windows(6, seed=9, max_iterations=6, pt_range=(300, 2000))
    s = ca.SchurComplementBundlerGpu(loss=HUBER, fixed_iterations=True)
    b = s.batch(probs)
    runs = []
    for _ in range(3):
        b.run()
        qs = [p.copy() for p in probs]
        b.read(qs)
        runs.append(qs)
    for qs in runs[1:]:
        for a, c in zip(runs[0], qs):
            assert a.last_cost_f64 == c.last_cost_f64 and a.final_lambda == c.final_lambda
            np.testing.assert_array_equal(a.points, c.points)
            np.testing.assert_array_equal(a.rig_from_world, c.rig_from_world)
    b.close()
    singles = []
    for _ in range(2):
        q = probs[0].copy()
        assert s.solve(q)
        singles.append(q)
    assert singles[0].last_cost_f64 == singles[1].last_cost_f64
    np.testing.assert_array_equal(singles[0].points, singles[1].points)
