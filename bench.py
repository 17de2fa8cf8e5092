"""Benchmark: batched SBA solves/s (10 LM iterations per window) on one MI355X, with ms per solve of the single
30-key-frame window and the relative cost error vs the FP64 oracle.

A step is one whole solve of every window of the batch from its initial state, with the inputs resident in HBM:
the hipGraph of one solve (initial cost, 10 LM iterations, gather) replayed on the solver's stream.

Default workload (`--workload c3`): 512 independent windows, per window key frames U[5,40] (2 fixed), points
U[500,8000], outlier rate U[0,10]%, each point observed by U{2..key frames} of the key frames it projects in,
Huber 1 px, exactly 10 LM iterations; the work granularity is autotuned before the timed region. The same
windows are also solved one window per launch sequence (the plain form) for the batched/plain ratio, and the
30-key-frame / 20,000-point window is timed and checked against the FP64 oracle as an extra.
`--workload c2` times that single window alone, `--workload c5` the 200-key-frame multi-camera map, and
`--workload c4` the differentiable solve (20 key frames, 10,000 points, Student-t): forward LM solve, FP64
Newton refinement, and the implicit backward to pixel and intrinsics gradients, checked against the FP64
oracle's adjoint at the same state and against FP64 central finite differences of converged oracle solves.

The CPU baseline is the oracle's FP32 restatement of the reference CPU bundler (one thread per window, like the
reference), timed on this host: `--cpu-warmups` untimed runs, then the median of `--cpu-runs` timed runs, on a
fixed sample of the workload's windows, both on one core and with one window per core on up to 16 cores.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_cores():
    """Host threads for the one-window-per-core baseline: the logical CPUs of this process, at most 16 (the CPU
    share of one GPU on an 8-GPU node)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def time_cpu(fn, warmups, runs):
    """Median, min and max seconds of fn() over `runs` timed runs after `warmups` untimed ones."""
    for _ in range(warmups):
        fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), float(np.min(ts)), float(np.max(ts))


def cpu_baseline_windows(ca, problems, args):
    """The FP32 restatement of SchurComplementBundlerCpu::solve on a fixed sample of the windows: (i) one core,
    the sample's windows one after another; (ii) one window per core, `cores` windows at once."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import sba_oracle as O

    def solve(p):
        O.solve(p, "f32", loss=ca.LOSS_HUBER, rcs_solver=1, dense_schur=1, fixed_iterations=1)

    rng = np.random.default_rng(12345)
    n_single = min(args.cpu_sample, len(problems))
    cores = cpu_cores()
    n_multi = min(cores, len(problems))
    single = [problems[i].copy() for i in sorted(rng.choice(len(problems), n_single, replace=False))]
    multi = [problems[i].copy() for i in sorted(rng.choice(len(problems), n_multi, replace=False))]
    t1 = time_cpu(lambda: [solve(p) for p in single], args.cpu_warmups, args.cpu_runs)
    with ThreadPoolExecutor(max_workers=cores) as pool:  # ctypes releases the GIL inside the oracle
        tn = time_cpu(lambda: list(pool.map(solve, multi)), args.cpu_warmups, args.cpu_runs)
    obs1 = sum(p.num_observations for p in single)
    obsn = sum(p.num_observations for p in multi)
    method = f"{args.cpu_warmups} warm-up runs, median of {args.cpu_runs} timed runs"
    return {
        "value": round(n_multi / tn[0], 4), "unit": "solves/s", "cores": cores, "kind": "port",
        "sample": (f"FP32 restatement of the reference CPU bundler (SchurComplementBundlerCpu::solve: dense Schur "
                   f"product, pseudo-inverse of the reduced system), 10 LM iterations per window, one window per "
                   f"thread on {cores} threads: {n_multi} windows ({obsn} observations) of the batch at once; "
                   f"{method}; {cpu_model()}, {os.cpu_count()} logical CPUs"),
        "seconds_median_min_max": [round(x, 3) for x in tn],
        "single_core": {"value": round(n_single / t1[0], 4), "unit": "solves/s", "cores": 1,
                        "sample": f"{n_single} windows ({obs1} observations) one after another on one thread; {method}",
                        "seconds_median_min_max": [round(x, 3) for x in t1]},
    }


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed counter passes (profiles/hbm_traffic_pmc.txt), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_pmc.txt")
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 4 and parts[0] == "traffic" and parts[1] == workload and parts[2] == kernel:
                    return float(parts[3])
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c3", choices=["c2", "c3", "c4", "c5", "inertial", "mono"])
    ap.add_argument("--cpu-warmups", type=int, default=3)
    ap.add_argument("--cpu-runs", type=int, default=10)
    ap.add_argument("--cpu-sample", type=int, default=4, help="windows of the one-core CPU baseline (c3)")
    ap.add_argument("--no-extra-c2", action="store_true", help="c3: skip the single 30-key-frame window extra")
    ap.add_argument("--no-extra-inertial", action="store_true", help="skip the inertial-SBA timings")
    ap.add_argument("--no-extra-production", action="store_true", help="skip the production-window timings")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-oracle", action="store_true")
    ap.add_argument("--chunk-pairs", type=int, default=0)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--autotune-table", default=os.path.join(ROOT, "cuvslam_amd", "autotune_table.txt"))
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a 256 MiB device copy kernel (k_copy_calibrate), the reference of HBM byte counters")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the SBA batch is sharded across GPUs by window with no exchange, "
              "and multi-GPU runs are not part of this build (use --gpus 1).", file=sys.stderr)
        sys.exit(2)

    import cuvslam_amd as ca
    from cuvslam_amd import roofline as RF

    if args.workload == "c4":
        return run_c4(args, ca)
    if args.workload == "inertial":
        return run_inertial(args)
    if args.workload == "mono":
        return run_mono(args)
    t0 = time.time()
    if args.workload == "c2":
        problems = [ca.scene.generate("c2", seed=2, num_fixed_keyframes=2, max_iterations=10)[0]]
        workload = "30 key frames (2 fixed), 20000 points, pinhole 450 px, 0.5 px noise, 5% outliers, Huber 1 px"
    elif args.workload == "c5":
        problems = [ca.scene.loop(seed=5, max_iterations=10)[0]]
        workload = ("large multi-camera map: 4 stereo pairs (8 cameras, 0.12 m baseline, yaw 0/90/180/270 deg), "
                    "200 key frames (1 fixed) on a 20 m loop, 200000 points on corridor walls/floor/ceiling plus "
                    "clutter, 3% outliers, Huber 1 px; reduced system 1194 x 1194")
    else:
        problems = ca.scene.batched_windows(512, seed=3, max_iterations=10)
        workload = ("512 windows in one batch: key frames U[5,40] (2 fixed), points U[500,8000], each point observed by "
                    "U{2..key frames} of the key frames it projects in, outliers U[0,10]%, 0.5 px noise, Huber 1 px, "
                    "10 LM iterations")
    pristine = [p.copy() for p in problems]
    n_obs = sum(p.num_observations for p in problems)
    log(f"[bench] generated {len(problems)} window(s), {n_obs} observations in {time.time() - t0:.1f}s")

    chunk_pairs, bchunk_pairs = args.chunk_pairs, 0
    tuned = None
    if args.workload == "c3" and chunk_pairs == 0 and not args.no_autotune:
        # work granularity of the batched launch, from the autotune table (tuned before the timed region and added
        # to the table when this batch's shape class is not in it)
        t0 = time.time()
        tuner = ca.SchurComplementBundlerGpu(loss=ca.LOSS_HUBER, fixed_iterations=True)
        chunk_pairs, bchunk_pairs, hit = tuner.autotune_table(problems, args.autotune_table, repeats=1)
        tuner.close()
        tuned = {"chunk_pairs": chunk_pairs, "bchunk_pairs": bchunk_pairs, "from_table": hit,
                 "table": os.path.relpath(args.autotune_table, ROOT)}
        log(f"[bench] autotune in {time.time() - t0:.1f}s: {tuned}")
    solver = ca.SchurComplementBundlerGpu(loss=ca.LOSS_HUBER, fixed_iterations=True, chunk_pairs=chunk_pairs,
                                          bchunk_pairs=bchunk_pairs)
    t0 = time.time()
    batch = solver.batch(problems)
    info = batch.info()
    log(f"[bench] packed and uploaded in {time.time() - t0:.1f}s: {info}")

    if args.calibrate_copy:
        ms_copy = solver.debug_copy(256 << 20, repeats=3)
        log(f"[bench] calibration copy of 256 MiB: {ms_copy:.3f} ms ({2 * 256 * 1.048576 / ms_copy:.0f} GB/s)")
    for _ in range(args.warmup):
        batch.run()
    batch.sync()
    log("[bench] warmup done")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.run()
    batch.sync()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    solves_per_s = len(problems) * args.steps / elapsed
    log(f"[bench] {args.steps} steps in {elapsed:.3f}s: {ms_per_step:.3f} ms/step, {solves_per_s:.1f} solves/s")

    # per-kernel durations: the same solve launched eagerly with HIP events around every kernel
    ms, launches = batch.kernel_times()
    names = ["k_schur", "k_reduce", "k_factor", "k_backsub", "k_decide"]
    per_launch = {n: (m / max(l, 1)) for n, m, l in zip(names, ms, launches)}
    log(f"[bench] per-launch ms: {per_launch}")

    work = RF.work(problems)
    # the dominant kernel and the roofline that bounds it
    def roofline(kern):
        bytes_, f32, f64 = work[{"k_schur": "schur", "k_factor": "factor", "k_backsub": "backsub"}[kern]]
        t_launch = per_launch[kern] * 1e-3
        if f64 > f32:
            ach, peak, unit, bound = f64 / t_launch / 1e12, RF.FP64_PEAK_TFLOPS, "TFLOP/s", "mfma"
        elif f32 / max(bytes_, 1) > RF.FP32_PEAK_TFLOPS * 1e3 / RF.HBM_PEAK_GBS:
            ach, peak, unit, bound = f32 / t_launch / 1e12, RF.FP32_PEAK_TFLOPS, "TFLOP/s", "mfma"
        else:
            ach, peak, unit, bound = bytes_ / t_launch / 1e9, RF.HBM_PEAK_GBS, "GB/s", "hbm"
        return {"bound": bound, "achieved": round(ach, 4), "peak": peak, "unit": unit, "frac": round(ach / peak, 6),
                "traffic": measured_traffic(args.workload, kern), "kernel": kern,
                "launch_ms": round(per_launch[kern], 5),
                "algorithmic": {"bytes": bytes_, "flops_f32": f32, "flops_f64": f64}}

    dom = max(["k_schur", "k_factor", "k_backsub"], key=lambda n: per_launch[n])
    roof = roofline(dom)
    # the same figures for each of the three main kernels
    roof_all = {k: {f: v for f, v in roofline(k).items() if f in ("bound", "achieved", "unit", "frac", "launch_ms")}
                for k in ("k_schur", "k_factor", "k_backsub")}

    # accuracy against the FP64 oracle (window 0)
    res = batch.read()
    extra = {}
    p0 = problems[0]
    if not args.no_oracle:
        from oracle import sba_oracle as O
        t0 = time.time()
        ref = O.solve(pristine[0], "f64", loss=ca.LOSS_HUBER, fixed_iterations=1)
        extra["rel_cost_err_vs_fp64_oracle"] = abs(p0.last_cost_f64 - ref["last_cost"]) / ref["last_cost"]
        q = pristine[0].copy()
        q.points, q.rig_from_world = p0.points, p0.rig_from_world
        c_eval, _ = O.cost(q, loss=ca.LOSS_HUBER)
        extra["rel_cost_err_state_reevaluated"] = abs(c_eval - ref["last_cost"]) / ref["last_cost"]
        nf = p0.rig_from_world.shape[0] - p0.num_fixed_key_frames
        dt = np.abs(p0.rig_from_world[:nf, :3, 3] - ref["poses"][:nf, :3, 3]).max()
        dR = p0.rig_from_world[:nf, :3, :3] @ np.transpose(ref["poses"][:nf, :3, :3], (0, 2, 1))
        # rotation angle from the skew part (arcsin |vee(dR - dR^T)| / 2), exact for small angles where arccos of
        # the trace loses half the digits
        sk = np.stack([dR[:, 2, 1] - dR[:, 1, 2], dR[:, 0, 2] - dR[:, 2, 0], dR[:, 1, 0] - dR[:, 0, 1]], 1)
        ang = np.degrees(np.arcsin(np.clip(np.linalg.norm(sk, axis=1) / 2, 0, 1))).max()
        extra["pose_err_m"] = float(dt)
        extra["pose_err_deg"] = float(ang)
        perr = np.linalg.norm(p0.points - ref["points"], axis=1)
        extra["point_rms_vs_oracle_m"] = float(np.sqrt((perr ** 2).mean()))
        extra["point_median_vs_oracle_m"] = float(np.median(perr))
        extra["point_p99_vs_oracle_m"] = float(np.percentile(perr, 99))
        extra["point_max_vs_oracle_m"] = float(perr.max())
        # points with an unobservable depth (two or three observations, one of them an outlier) drift along their
        # ray in both FP32 and FP64 solvers; count those that end more than 1 m apart
        extra["points_over_1m_apart"] = int((perr > 1.0).sum())
        extra["oracle_iterations"] = ref["iterations"]
        log(f"[bench] oracle FP64 solve {time.time() - t0:.1f}s: {extra}")
        if args.workload == "c5":
            # the map is not converged after 10 iterations: FP32 and FP64 trajectories differ at the 1e-4 level
            # there (the reference's own FP32 path included); compare with the FP32 restatement of the reference
            # path at 10 iterations, and with the FP64 oracle after 30 iterations of both
            t0 = time.time()
            r32 = O.solve(pristine[0], "f32", loss=ca.LOSS_HUBER, rcs_solver=0, dense_schur=0, fixed_iterations=1)
            q32 = pristine[0].copy()
            q32.points, q32.rig_from_world = r32["points"].astype(np.float32), r32["poses"].astype(np.float32)
            c32, _ = O.cost(q32, loss=ca.LOSS_HUBER)
            extra["rel_cost_err_vs_fp32_reference_path"] = abs(p0.last_cost_f64 - c32) / c32
            log(f"[bench] FP32 reference path {time.time() - t0:.1f}s")
            t0 = time.time()
            p30 = pristine[0].copy()
            p30.max_iterations = 30
            g30 = ca.SchurComplementBundlerGpu(loss=ca.LOSS_HUBER, fixed_iterations=True)
            q30 = p30.copy()
            g30.solve(q30)
            ref30 = O.solve(p30, "f64", loss=ca.LOSS_HUBER, fixed_iterations=1)
            extra["rel_cost_err_vs_fp64_oracle_30_iterations"] = abs(q30.last_cost_f64 - ref30["last_cost"]) / ref30["last_cost"]
            log(f"[bench] 30-iteration comparison {time.time() - t0:.1f}s: {extra}")
    extra["iterations"] = int(p0.iterations)
    extra["accepted"] = int(p0.accepted)
    extra["final_cost"] = p0.last_cost_f64

    # end to end through the host boundary: host buffers in (sort, pack, upload), solve, results back to the host
    # (PCIe-inclusive; not the headline value, whose inputs are already resident in HBM)
    reps = 3 if args.workload != "c3" else 1
    host_sets = [[q.copy() for q in pristine] for _ in range(reps)]
    t0 = time.perf_counter()
    for hosts in host_sets:
        if len(hosts) > 1:
            solver.solve_batch(hosts)
        else:
            solver.solve(hosts[0])
    del host_sets
    extra["end_to_end_ms_per_batch_host_buffers"] = round(1e3 * (time.perf_counter() - t0) / reps, 3)
    log(f"[bench] end to end from host buffers: {extra['end_to_end_ms_per_batch_host_buffers']} ms per batch")

    if args.workload == "c3" and not args.no_plain:
        # the plain form: the same windows solved one window per launch sequence, one after another
        plain = [solver.batch([p]) for p in pristine]
        for pb in plain[:4]:
            pb.run()
        plain[0].sync()
        t0 = time.perf_counter()
        for pb in plain:
            pb.run()
            pb.sync()
        t_plain = time.perf_counter() - t0
        extra["plain_solves_per_s"] = len(plain) / t_plain
        extra["batched_over_plain"] = solves_per_s / extra["plain_solves_per_s"]
        log(f"[bench] plain one-window-per-launch: {t_plain:.3f}s for {len(plain)} windows, "
            f"{extra['plain_solves_per_s']:.1f} solves/s; batched/plain = {extra['batched_over_plain']:.1f}x")
        for pb in plain:
            pb.close()

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import sba_oracle as O
        sample = pristine[0]
        if args.workload == "c3":
            cpu = cpu_baseline_windows(ca, pristine, args)
            log(f"[bench] CPU baseline: {cpu}")
        elif args.workload == "c5":
            # the reference's dense Schur product is 1.7 TFLOP per iteration at this size: the baseline is the
            # FP32 restatement with the block-sparse Schur product, one LM iteration timed and scaled to ten
            one = sample.copy()
            one.max_iterations = 1
            t0 = time.perf_counter()
            O.solve(one, "f32", loss=ca.LOSS_HUBER, rcs_solver=0, dense_schur=0, fixed_iterations=1)
            t_cpu = 10 * (time.perf_counter() - t0)
            what = ("one LM iteration of the map with the FP32 restatement of the reference CPU bundler "
                    "(block-sparse Schur product, Cholesky solve of the reduced system), x10")
            cpu = {"value": round(1.0 / t_cpu, 6), "unit": "solves/s", "cores": 1, "kind": "port",
                   "sample": f"{what}, single thread, {cpu_model()}", "seconds": round(t_cpu, 3)}
        else:
            med = time_cpu(lambda: O.solve(sample, "f32", loss=ca.LOSS_HUBER, rcs_solver=1, dense_schur=1,
                                           fixed_iterations=1), args.cpu_warmups, args.cpu_runs)
            what = (f"full solve (10 LM iterations) of the window ({sample.num_observations} observations) with the "
                    f"FP32 restatement of the reference CPU bundler, single thread, {args.cpu_warmups} warm-up runs, "
                    f"median of {args.cpu_runs} timed runs, {cpu_model()}")
            cpu = {"value": round(1.0 / med[0], 6), "unit": "solves/s", "cores": 1, "kind": "port",
                   "sample": what, "seconds_median_min_max": [round(x, 3) for x in med]}
        log(f"[bench] CPU baseline {cpu['value']} solves/s")

    if args.workload == "c3" and not args.no_extra_c2:
        extra["single_window_30kf"] = single_window_extra(ca, RF, args)
    if args.workload in ("c2", "c3") and not args.no_extra_production:
        extra["production_window_7kf"] = production_window_extra(ca, args)
    if args.workload in ("c2", "c3") and not args.no_extra_inertial:
        extra["inertial_sba"] = inertial_extra(args)

    line = {
        "metric": "ms per BA solve (10 LM iters) + batched solves/s; rel. cost err vs FP64 oracle",
        "value": round(solves_per_s, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": workload, "windows": len(problems), "observations": n_obs,
                   "pairs": info["pairs"], "lm_iterations": 10, "chunks": info["chunks"], "autotune": tuned},
        "roofline": roof,
        "roofline_by_kernel": roof_all,
        "cpu_baseline": cpu,
        "kernels_ms_per_launch": {k: round(v, 5) for k, v in per_launch.items()},
        "accuracy": extra,
    }
    print(json.dumps(line), flush=True)


def single_window_extra(ca, RF, args):
    """The 30-key-frame / 20,000-point window (528k observations) alone: ms per solve with inputs resident in HBM,
    per-kernel launch times, and the relative cost error vs the FP64 oracle after exactly 10 LM iterations."""
    p, _, _ = ca.scene.generate("c2", seed=2, num_fixed_keyframes=2, max_iterations=10)
    pristine = p.copy()
    solver = ca.SchurComplementBundlerGpu(loss=ca.LOSS_HUBER, fixed_iterations=True)
    batch = solver.batch([p])
    for _ in range(args.warmup):
        batch.run()
    batch.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.run()
    batch.sync()
    ms_solve = 1e3 * (time.perf_counter() - t0) / args.steps
    ms, launches = batch.kernel_times()
    names = ["k_schur", "k_reduce", "k_factor", "k_backsub", "k_decide"]
    per_launch = {n: round(m / max(l, 1), 5) for n, m, l in zip(names, ms, launches)}
    batch.read()
    out = {"workload": "30 key frames (2 fixed), 20000 points, 0.5 px noise, 5% outliers, Huber 1 px",
           "observations": p.num_observations, "ms_per_solve": round(ms_solve, 4),
           "kernels_ms_per_launch": per_launch, "iterations": int(p.iterations), "accepted": int(p.accepted)}
    f64 = RF.work([pristine])["factor"][2]
    out["k_factor_fp64_frac"] = round(f64 / (per_launch["k_factor"] * 1e-3) / 1e12 / RF.FP64_PEAK_TFLOPS, 6)
    if not args.no_oracle:
        from oracle import sba_oracle as O
        ref = O.solve(pristine, "f64", loss=ca.LOSS_HUBER, fixed_iterations=1)
        out["rel_cost_err_vs_fp64_oracle"] = abs(p.last_cost_f64 - ref["last_cost"]) / ref["last_cost"]
    # the drop-in call from host buffers in steady state (the handle keeps its device workspace, as the reference
    # keeps its buffers across solve() calls): two warm-up calls, then the median of five
    host = []
    for k in range(7):
        q = pristine.copy()
        t0 = time.perf_counter()
        solver.solve(q)
        if k >= 2:
            host.append(time.perf_counter() - t0)
    out["end_to_end_ms_host_buffers"] = round(1e3 * float(np.median(host)), 3)
    batch.close()
    solver.close()
    log(f"[bench] single 30-key-frame window: {out}")
    return out


def production_window_extra(ca, args):
    """The reference's production window (sba_config.h:27-46: 7 key frames of which 3 fixed, 7 LM iterations, the
    Student-t loss with scale 0.5, the early exit on), 2,000 points: the drop-in solve from host buffers (pack,
    upload, solve, read back through the handle's kept workspace) and the device-only solve, median of 50 runs."""
    p, _, _ = ca.scene.generate(seed=7, num_points=2000, num_keyframes=7, num_fixed_keyframes=3, max_iterations=7)
    p.robustifier_scale = 0.5
    solver = ca.SchurComplementBundlerGpu(loss=ca.LOSS_STUDENT_T)
    host = []
    for k in range(55):
        q = p.copy()
        t0 = time.perf_counter()
        solver.solve(q)
        if k >= 5:
            host.append(time.perf_counter() - t0)
    batch = solver.batch([p])
    dev = []
    for k in range(55):
        t0 = time.perf_counter()
        batch.run()
        batch.sync()
        if k >= 5:
            dev.append(time.perf_counter() - t0)
    batch.close()
    solver.close()
    out = {"workload": "7 key frames (3 fixed), 2000 points, Student-t 0.5, 7 LM iterations, early exit on",
           "observations": p.num_observations, "iterations": int(q.iterations),
           "ms_host_buffers_median": round(1e3 * float(np.median(host)), 4),
           "ms_device_resident_median": round(1e3 * float(np.median(dev)), 4)}
    log(f"[bench] production window: {out}")
    return out


def inertial_extra(args):
    """The inertial solve (IMUBundlerGpuFixedVel::solve, 10 LM iterations, the settings of run_imu_sba()) from host
    buffers on the reference's own benchmark shapes (libs/cuda_modules/test/imu_sba_test.cpp:426-429: 10 key frames
    of which 1 fixed, 450 / 800 / 1000 points on 4-frame tracks; the 20-key-frame, 5900-point initialization window),
    median of 20 runs; beside it the FP64 oracle's solve of the same window on one core (the CPU baseline)."""
    from cuvslam_amd import imu, imu_scene
    out = {}
    for name, shape, init in (("10kf_450pts", (10, 1, 450, 4), False), ("10kf_800pts", (10, 1, 800, 4), False),
                              ("10kf_1000pts", (10, 1, 1000, 4), False), ("init_20kf_5900pts", imu_scene.INIT_SHAPE, True)):
        p, calib, _ = imu_scene.make_scenario(1, shape=shape)
        if init:
            imu_scene.use_init_settings(p)
        bundler = imu.IMUBundlerGpuFixedVel(calib)
        ms = []
        for k in range(25):
            q = p.copy()
            t0 = time.perf_counter()
            bundler.solve(q)
            if k >= 5:
                ms.append(1e3 * (time.perf_counter() - t0))
        bundler.close()
        row = {"observations": p.num_observations, "iterations": int(q.iterations), "accepted": int(q.accepted),
               "ms_per_solve_median": round(float(np.median(ms)), 3)}
        if not args.no_cpu_baseline:
            from oracle import imu_oracle
            t = []
            for _ in range(5):
                t0 = time.perf_counter()
                ref = imu_oracle.solve(p, calib)
                t.append(1e3 * (time.perf_counter() - t0))
            row["cpu_baseline_ms_fp64_oracle_one_core"] = round(float(np.median(t)), 3)
            row["rel_cost_err_vs_fp64_oracle"] = abs(q.last_cost_f64 - ref["last_cost"]) / ref["last_cost"]
        out[name] = row
        log(f"[bench] inertial {name}: {row}")
    return out


def run_inertial(args):
    """The inertial solve alone on the reference's initialization window (20 key frames, 5900 points; the settings of
    the reference's init path): `steps` solves from host buffers after `warmup`, with the FP64 oracle's solve of the
    same window on one core beside it."""
    from cuvslam_amd import imu, imu_scene
    p, calib, _ = imu_scene.make_scenario(1, shape=imu_scene.INIT_SHAPE)
    imu_scene.use_init_settings(p)
    bundler = imu.IMUBundlerGpuFixedVel(calib)
    t = []
    for k in range(args.warmup + args.steps):
        q = p.copy()
        t0 = time.perf_counter()
        bundler.solve(q)
        if k >= args.warmup:
            t.append(time.perf_counter() - t0)
    bundler.close()
    med = float(np.median(t))
    line = {
        "metric": "ms per inertial SBA solve (10 LM iters); rel. cost err vs FP64 oracle",
        "value": round(1.0 / med, 4), "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * med, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "inertial initialization window: 20 key frames (1 fixed), 5900 points, IMU at 200 Hz, "
                               "10 LM iterations, from host buffers", "observations": p.num_observations,
                   "iterations": int(q.iterations), "accepted": int(q.accepted)},
    }
    if not args.no_cpu_baseline:
        from oracle import imu_oracle
        c = []
        for _ in range(3):
            t0 = time.perf_counter()
            ref = imu_oracle.solve(p, calib)
            c.append(time.perf_counter() - t0)
        line["cpu_baseline"] = {"value": round(1.0 / float(np.median(c)), 4), "unit": "solves/s", "cores": 1,
                                "kind": "port", "sample": "the same window, FP64 oracle, median of 3"}
        line["rel_cost_err_vs_fp64_oracle"] = abs(q.last_cost_f64 - ref["last_cost"]) / ref["last_cost"]
    print(json.dumps(line), flush=True)


def run_mono(args):
    """Mono SBA windows (MonoSBASolver::solve) in batches of 512, one workgroup per window, from host buffers."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import mono_cases as MC
    from cuvslam_amd import mono
    solver = mono.MonoSolverGpu()
    base = [MC.mono_window(100 + k) for k in range(64)]
    ws = (base * 8)[:512]
    t = []
    for k in range(args.warmup + args.steps):
        t0 = time.perf_counter()
        solver.solve_batch(ws)
        if k >= args.warmup:
            t.append(time.perf_counter() - t0)
    med = float(np.median(t))
    single = []  # one window per call, as the reference's MonoSBASolver::solve
    for k in range(12):
        t0 = time.perf_counter()
        solver.solve_batch([base[k % len(base)]])
        if k >= 2:
            single.append(time.perf_counter() - t0)
    line = {
        "metric": "mono SBA windows/s", "value": round(len(ws) / med, 2), "unit": "windows/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * med, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "512 monocular windows per batch (8 frames, 3 constrained, 300 tracks, 0.5 px "
                               "noise), up to 30 LM iterations with the reference's stop tests, from host buffers"},
        "single_window_ms_median": round(1e3 * float(np.median(single)), 4),
    }
    if not args.no_cpu_baseline:
        from oracle import sba_oracle as O
        c = []
        for w in base[:8]:
            t0 = time.perf_counter()
            O.mono_solve(w)
            c.append(time.perf_counter() - t0)
        line["cpu_baseline"] = {"value": round(1.0 / float(np.median(c)), 2), "unit": "windows/s", "cores": 1,
                                "kind": "port", "sample": "8 of the windows, FP64 oracle, median"}
    print(json.dumps(line), flush=True)


def run_c4(args, ca):
    """Differentiable solve on the C4 window: L = sum |T_j - G_j|^2 over the free key frames after convergence."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from adjoint_fd import C4_SCENE, K0, Case, chain_to_pixels

    c = Case(**C4_SCENE)
    p = c.problem
    log(f"[bench] c4 window: {p.num_observations} observations, {p.points.shape[0]} points")
    solver = ca.SchurComplementBundlerGpu(loss=c.loss, fixed_iterations=True)
    batch = solver.batch([p])
    t_fwd, t_ref, t_adj = [], [], []
    for k in range(args.warmup + args.steps):
        t0 = time.perf_counter()
        batch.run()
        batch.sync()
        t1 = time.perf_counter()
        poses, pts, gn = batch.refine(gauss_newton_steps=8, newton_steps=3, max_cg=300, tol=1e-12)
        t2 = time.perf_counter()
        dxy, dinfo, it, _ = batch.adjoint(c.g_of(poses), max_cg=300, tol=1e-12)
        t3 = time.perf_counter()
        if k >= args.warmup:
            t_fwd.append(t1 - t0)
            t_ref.append(t2 - t1)
            t_adj.append(t3 - t2)
    ms = lambda v: round(1e3 * float(np.median(v)), 4)  # noqa: E731
    log(f"[bench] forward {ms(t_fwd)} ms, refine {ms(t_ref)} ms, backward {ms(t_adj)} ms, |grad F| {gn:.3e}")
    duv, dK = chain_to_pixels(dxy, dinfo, c.uv, K0)
    acc = {"grad_norm_after_refine": gn, "cg_iterations": it}
    with open(os.path.join(ROOT, "tests", "golden", "c4_fd.json")) as f:
        fx = json.load(f)
    fd_uv, fd_K = np.array(fx["fd_pixels"]), np.array(fx["fd_intrinsics"])
    picked = np.array([duv[o, cc] for o, cc in fx["sampled"]])
    acc["rel_err_pixels_vs_fd"] = float(np.linalg.norm(picked - fd_uv) / np.linalg.norm(fd_uv))
    acc["rel_err_intrinsics_vs_fd"] = [float(abs(a - b) / abs(b)) for a, b in zip(dK, fd_K)]
    acc["fd"] = "FP64 central differences of converged oracle solves (tests/golden/c4_fd.json): " \
                f"{len(fd_uv)} pixel components at 1e-3 px, 4 intrinsics at 1e-4"
    if not args.no_oracle:
        from oracle import sba_oracle as O
        t0 = time.time()
        _, rdxy, rdinfo, _ = O.implicit_grad(p, c.g_of(poses), pts, poses, loss=c.loss)
        acc["rel_err_vs_oracle_adjoint_same_state"] = float(
            max(np.linalg.norm(dxy - rdxy) / np.linalg.norm(rdxy), np.linalg.norm(dinfo - rdinfo) / np.linalg.norm(rdinfo)))
        log(f"[bench] oracle adjoint {time.time() - t0:.1f}s")
    log(f"[bench] gradient checks: {acc}")
    total = np.array(t_fwd) + np.array(t_ref) + np.array(t_adj)
    line = {
        "metric": "ms per BA solve (10 LM iters) + batched solves/s; rel. cost err vs FP64 oracle",
        "value": round(1.0 / float(np.median(total)), 4),
        "unit": "differentiable solves/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms(total),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": "20 key frames (2 fixed), 10000 points, no outliers, Student-t, 50 LM iterations, then "
                               "8 Gauss-Newton + 3 Newton steps in FP64; L = sum of squared pose error vs ground truth",
                   "observations": p.num_observations},
        "ms": {"forward_lm": ms(t_fwd), "refine": ms(t_ref), "backward": ms(t_adj)},
        "gradients": acc,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
