"""Python mirror of cuVSLAM's SBA problem/solve API over the MI355X C ABI.

``BundleAdjustmentProblem`` has the fields of libs/sba/bundle_adjustment_problem.h:34-68 (as numpy arrays) and
``SchurComplementBundlerGpu.solve`` the contract of libs/sba/schur_complement_bundler_gpu.h:29-38: the problem is
updated in place (points, rig_from_world, iterations, initial_cost, last_cost) and the return value is True when
the solve ran, False for an empty problem or an infeasible start (every observation inside the near plane).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import LOSS_HUBER, LOSS_STUDENT_T, SBA_OK  # noqa: F401


class SbaError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"SBA error {code}: {msg}")
        self.code = code


@dataclass
class BundleAdjustmentProblem:
    """In/out problem of one window. Poses are 4x4 matrices indexed [row, col]; the last
    ``num_fixed_key_frames`` poses and the last ``num_fixed_points`` points do not move."""
    points: np.ndarray                      # (N_p, 3) float32
    observation_xys: np.ndarray             # (N_o, 2) float32, normalized image coordinates
    observation_infos: np.ndarray           # (N_o, 2, 2) float32 information matrices
    point_ids: np.ndarray                   # (N_o,) int32
    pose_ids: np.ndarray                    # (N_o,) int32
    camera_ids: np.ndarray                  # (N_o,) int8
    rig_from_world: np.ndarray              # (N_c, 4, 4) float32
    camera_from_rig: np.ndarray             # (C, 4, 4) float32 (camera::Rig::camera_from_rig)
    num_fixed_points: int = 0
    num_fixed_key_frames: int = 0
    max_iterations: int = 20
    robustifier_scale: float = 1.0
    iterations: int = 0
    initial_cost: float = 0.0
    last_cost: float = 0.0
    # extra statistics of the MI355X solver (not in the reference struct)
    initial_cost_f64: float = 0.0
    last_cost_f64: float = 0.0
    status: int = 0
    accepted: int = 0
    final_lambda: float = 0.0               # the LM damping when the solve ended
    _keep: list = field(default_factory=list, repr=False)

    def copy(self) -> "BundleAdjustmentProblem":
        return BundleAdjustmentProblem(
            self.points.copy(), self.observation_xys.copy(), self.observation_infos.copy(), self.point_ids.copy(),
            self.pose_ids.copy(), self.camera_ids.copy(), self.rig_from_world.copy(), self.camera_from_rig.copy(),
            self.num_fixed_points, self.num_fixed_key_frames, self.max_iterations, self.robustifier_scale)

    @property
    def num_observations(self) -> int:
        return int(self.point_ids.shape[0])

    def view(self) -> _lib.ProblemView:
        """A C view whose buffers stay alive as long as this problem (column-major copies of the matrices)."""
        pts = np.ascontiguousarray(self.points, dtype=np.float32).reshape(-1, 3)
        xy = np.ascontiguousarray(self.observation_xys, dtype=np.float32).reshape(-1, 2)
        # column-major 2x2 blocks; symmetric blocks (the usual information matrices) read the same either way, so
        # only asymmetric ones are copied with their off-diagonal entries swapped
        info = np.ascontiguousarray(self.observation_infos, dtype=np.float32).reshape(-1, 4)
        if not np.array_equal(info[:, 1], info[:, 2]):
            info = np.ascontiguousarray(info.reshape(-1, 2, 2).transpose(0, 2, 1)).reshape(-1, 4)
        pid = np.ascontiguousarray(self.point_ids, dtype=np.int32)
        qid = np.ascontiguousarray(self.pose_ids, dtype=np.int32)
        cid = np.ascontiguousarray(self.camera_ids, dtype=np.int8)
        poses = np.ascontiguousarray(np.asarray(self.rig_from_world, dtype=np.float32).reshape(-1, 4, 4)
                                     .transpose(0, 2, 1))
        cams = np.ascontiguousarray(np.asarray(self.camera_from_rig, dtype=np.float32).reshape(-1, 4, 4)
                                    .transpose(0, 2, 1))
        self._keep = [pts, xy, info, pid, qid, cid, poses, cams]
        v = _lib.ProblemView()
        v.num_points = pts.shape[0]
        v.num_fixed_points = int(self.num_fixed_points)
        v.num_observations = pid.shape[0]
        v.num_poses = poses.shape[0]
        v.num_fixed_key_frames = int(self.num_fixed_key_frames)
        v.num_cameras = cams.shape[0]
        f = lambda a, t=C.c_float: a.ctypes.data_as(C.POINTER(t))  # noqa: E731
        v.points, v.observation_xys, v.observation_infos = f(pts), f(xy), f(info)
        v.point_ids, v.pose_ids, v.camera_ids = f(pid, C.c_int32), f(qid, C.c_int32), f(cid, C.c_int8)
        v.rig_from_world, v.camera_from_rig = f(poses), f(cams)
        v.max_iterations = int(self.max_iterations)
        v.robustifier_scale = float(self.robustifier_scale)
        return v


def _result_for(p: BundleAdjustmentProblem):
    out_pts = np.zeros((p.points.shape[0], 3), np.float32)
    out_poses = np.zeros((p.rig_from_world.shape[0], 4, 4), np.float32)
    r = _lib.Result()
    r.points = out_pts.ctypes.data_as(C.POINTER(C.c_float))
    r.rig_from_world = out_poses.ctypes.data_as(C.POINTER(C.c_float))
    return r, out_pts, out_poses


def _apply(p: BundleAdjustmentProblem, r: _lib.Result, out_pts, out_poses) -> bool:
    p.status = int(r.status)
    if r.status in (_lib.SBA_OK, _lib.SBA_E_INFEASIBLE):
        p.iterations = int(r.iterations)
        p.initial_cost = float(r.initial_cost)
        p.last_cost = float(r.last_cost)
        p.initial_cost_f64 = float(r.initial_cost_f64)
        p.last_cost_f64 = float(r.last_cost_f64)
        p.accepted = int(r.accepted)
        p.final_lambda = float(r.lambda_)
    if r.status == _lib.SBA_OK:
        p.points = out_pts
        p.rig_from_world = out_poses.transpose(0, 2, 1).copy()
        return True
    return False


class SchurComplementBundlerGpu:
    """cuvslam::sba::SchurComplementBundlerGpu on MI355X. ``loss``: LOSS_STUDENT_T (the reference's) or LOSS_HUBER."""

    def __init__(self, loss: int = LOSS_STUDENT_T, device: int = 0, fixed_iterations: bool = False,
                 use_graph: bool = True, chunk_pairs: int = 0, bchunk_pairs: int = 0, pack_threads: int = 0):
        L = _lib.lib()
        opt = _lib.Options()
        L.sba_default_options(C.byref(opt))
        opt.device, opt.loss = device, loss
        opt.fixed_iterations, opt.use_graph, opt.chunk_pairs = int(fixed_iterations), int(use_graph), chunk_pairs
        opt.bchunk_pairs, opt.pack_threads = bchunk_pairs, pack_threads
        h = C.c_void_p()
        rc = L.sba_create(C.byref(opt), C.byref(h))
        if rc != SBA_OK:
            raise SbaError(rc, "sba_create failed (is a GPU visible?)")
        self._L, self._h = L, h

    def close(self):
        if getattr(self, "_h", None):
            self._L.sba_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def last_error(self) -> str:
        return self._L.sba_last_error(self._h).decode()

    def autotune(self, problems: Sequence[BundleAdjustmentProblem], candidates: Optional[Sequence[int]] = None,
                 repeats: int = 2):
        """Times the batched solve of ``problems`` for each chunk size (pairs of work per Schur chunk) and returns
        (best chunk_pairs, [ms per solve of each given candidate]). Use it as ``chunk_pairs`` of a new solver."""
        views = (_lib.ProblemView * len(problems))(*[p.view() for p in problems])
        cand = np.ascontiguousarray(candidates if candidates is not None else [], dtype=np.int32)
        ms = np.zeros(max(len(cand), 1), np.float32)
        best = C.c_int32()
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))  # noqa: E731
        rc = self._L.sba_autotune_chunks(self._h, views, len(problems), ip(cand) if len(cand) else None, len(cand),
                                         repeats, C.byref(best), ms.ctypes.data_as(C.POINTER(C.c_float)))
        if rc != SBA_OK:
            raise SbaError(rc, self.last_error())
        return int(best.value), [float(x) for x in ms[:len(cand)]]

    def autotune_table(self, problems: Sequence[BundleAdjustmentProblem], path: str, repeats: int = 1):
        """Chunk sizes for ``problems`` from the autotune table at ``path`` (tuned and appended when the batch's shape
        class is not there): (chunk_pairs, bchunk_pairs, taken from the table)."""
        views = (_lib.ProblemView * len(problems))(*[p.view() for p in problems])
        c, bc, hit = C.c_int32(), C.c_int32(), C.c_int32()
        rc = self._L.sba_autotune_table(self._h, os.fsencode(path), views, len(problems), repeats, C.byref(c),
                                        C.byref(bc), C.byref(hit))
        if rc != SBA_OK:
            raise SbaError(rc, self.last_error())
        return int(c.value), int(bc.value), bool(hit.value)

    def debug_copy(self, nbytes: int, repeats: int = 3) -> float:
        """Device-to-device copy of nbytes by a plain kernel (the calibration of HBM byte counters); ms per copy."""
        ms = C.c_float()
        rc = self._L.sba_debug_copy(self._h, int(nbytes), int(repeats), C.byref(ms))
        if rc != SBA_OK:
            raise SbaError(rc, self.last_error())
        return ms.value

    def solve(self, problem: BundleAdjustmentProblem) -> bool:
        v = problem.view()
        r, op, oq = _result_for(problem)
        rc = self._L.sba_solve(self._h, C.byref(v), C.byref(r))
        if rc in (_lib.SBA_E_DEVICE, _lib.SBA_E_CAPACITY, _lib.SBA_E_INVALID):
            raise SbaError(rc, self.last_error())
        if rc == _lib.SBA_E_EMPTY:
            problem.status = rc
            return False
        return _apply(problem, r, op, oq)

    def solve_batch(self, problems: Sequence[BundleAdjustmentProblem]) -> List[bool]:
        views = (_lib.ProblemView * len(problems))(*[p.view() for p in problems])
        res = (_lib.Result * len(problems))()
        outs = []
        for k, p in enumerate(problems):
            r, op, oq = _result_for(p)
            res[k] = r
            outs.append((op, oq))
        rc = self._L.sba_solve_batch(self._h, views, len(problems), res)
        if rc != SBA_OK:
            raise SbaError(rc, self.last_error())
        return [_apply(p, res[k], *outs[k]) for k, p in enumerate(problems)]

    def debug_factor(self, A, b):
        """Factor/solve of a symmetric system with the reduced-camera-system kernel (n % 6 == 0): (x, rc)."""
        A = np.ascontiguousarray(A, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b)
        d = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
        rc = self._L.sba_debug_factor(self._h, A.shape[0], d(A), d(b), d(x))
        if rc not in (SBA_OK, _lib.SBA_E_INFEASIBLE):
            raise SbaError(rc, self.last_error())
        return x, rc

    def batch(self, problems: Sequence[BundleAdjustmentProblem]) -> "DeviceBatch":
        return DeviceBatch(self, problems)


class DeviceBatch:
    """Windows packed and resident in HBM; ``run()`` replays the whole LM solve (one hipGraph) from the
    uploaded initial state without host transfers."""

    def __init__(self, solver: SchurComplementBundlerGpu, problems: Sequence[BundleAdjustmentProblem]):
        self.solver, self.problems = solver, list(problems)
        self._views = (_lib.ProblemView * len(self.problems))(*[p.view() for p in self.problems])
        b = C.c_void_p()
        rc = solver._L.sba_batch_create(solver._h, self._views, len(self.problems), C.byref(b))
        if rc != SBA_OK:
            raise SbaError(rc, solver.last_error())
        self._b = b

    def run(self):
        rc = self.solver._L.sba_batch_run(self.solver._h, self._b)
        if rc != SBA_OK:
            raise SbaError(rc, self.solver.last_error())

    def sync(self):
        rc = self.solver._L.sba_batch_sync(self.solver._h)
        if rc != SBA_OK:
            raise SbaError(rc, self.solver.last_error())

    def read(self, problems: Optional[Sequence[BundleAdjustmentProblem]] = None) -> List[bool]:
        problems = self.problems if problems is None else problems
        res = (_lib.Result * len(problems))()
        outs = []
        for k, p in enumerate(problems):
            r, op, oq = _result_for(p)
            res[k] = r
            outs.append((op, oq))
        rc = self.solver._L.sba_batch_read(self.solver._h, self._b, res)
        if rc != SBA_OK:
            raise SbaError(rc, self.solver.last_error())
        return [_apply(p, res[k], *outs[k]) for k, p in enumerate(problems)]

    def info(self):
        no, npair = C.c_int64(), C.c_int64()
        nch, nw = C.c_int32(), C.c_int32()
        self.solver._L.sba_batch_info(self.solver._h, self._b, C.byref(no), C.byref(npair), C.byref(nch), C.byref(nw))
        return {"observations": no.value, "pairs": npair.value, "chunks": nch.value, "windows": nw.value}

    def kernel_times(self):
        """One eager solve with HIP events around every launch: total ms and launches per kernel class
        (linearize+Schur, record reduction, factor, back-substitute+cost, decide)."""
        ms = (C.c_float * 5)()
        n = (C.c_int32 * 5)()
        rc = self.solver._L.sba_batch_kernel_times(self.solver._h, self._b, ms, n)
        if rc != SBA_OK:
            raise SbaError(rc, self.solver.last_error())
        return list(ms), list(n)

    def debug_step(self, lam: float):
        """One LM step of window 0 from its initial state with damping ``lam``: (S, e, dc, dp, scalars, rc)."""
        p = self.problems[0]
        ncf = p.rig_from_world.shape[0] - p.num_fixed_key_frames
        npf = p.points.shape[0] - p.num_fixed_points
        n = 6 * ncf
        S = np.zeros((n, n)); e = np.zeros(n); dc = np.zeros(n); dp = np.zeros(3 * npf); sc = np.zeros(8)
        d = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
        rc = self.solver._L.sba_batch_debug_step(self.solver._h, self._b, lam, d(S), d(e), d(dc), d(dp), d(sc))
        if rc not in (SBA_OK, _lib.SBA_E_INFEASIBLE):
            raise SbaError(rc, self.solver.last_error())
        return S, e, dc, dp, sc, rc

    def trace(self, window: int = 0):
        """The LM iterations of ``window`` in the last run: dict of arrays rho, trial_cost, lambda, flags (1 accepted,
        2 factorization failed, 4 stopped), one entry per iteration (the first 64)."""
        cap = 64
        rho, cost = np.zeros(cap), np.zeros(cap)
        lam = np.zeros(cap, np.float32)
        flags = np.zeros(cap, np.int32)
        n = C.c_int32()
        d = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
        rc = self.solver._L.sba_batch_trace(self.solver._h, self._b, window, cap, d(rho), d(cost),
                                            lam.ctypes.data_as(C.POINTER(C.c_float)),
                                            flags.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(n))
        if rc != SBA_OK:
            raise SbaError(rc, self.solver.last_error())
        k = n.value
        return {"rho": rho[:k], "trial_cost": cost[:k], "lambda": lam[:k], "flags": flags[:k]}

    def winsorize(self, key_frame: int, window: int = 0):
        """Landmark winsorizing of a solved window (pipelines/winsorizer.cpp:23-73): resets the points whose
        observations from ``key_frame`` are among the 5% largest reprojection errors (above 0.0025^2, at least 50
        observations). The next ``run`` continues from this solve without them. Returns (reset mask (N_p,) bool,
        number reset)."""
        p = self.problems[window]
        mask = np.zeros(p.points.shape[0], np.uint8)
        n = C.c_int32()
        rc = self.solver._L.sba_batch_winsorize(self.solver._h, self._b, window, key_frame,
                                                mask.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(n))
        if rc != SBA_OK:
            raise SbaError(rc, self.solver.last_error())
        return mask.astype(bool), n.value

    def refine(self, newton_steps: int = 3, max_cg: int = 200, tol: float = 1e-12, gauss_newton_steps: int = 4):
        """FP64 refinement of the solved state of window 0 (the state ``adjoint`` then differentiates at):
        ``gauss_newton_steps`` steps with the factored Gauss-Newton system, then ``newton_steps`` full-Hessian
        Newton steps. Returns (rig_from_world (N_c, 4, 4) float64, points (N_p, 3) float64, |grad F| after the
        last step)."""
        p = self.problems[0]
        pts = np.zeros((p.points.shape[0], 3))
        T = np.zeros((p.rig_from_world.shape[0], 16))
        gn = C.c_double()
        d = lambda x: x.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
        rc = self.solver._L.sba_batch_refine(self.solver._h, self._b, gauss_newton_steps, newton_steps, max_cg, tol,
                                             d(pts), d(T), C.byref(gn))
        if rc != SBA_OK:
            raise SbaError(rc, self.solver.last_error())
        return T.reshape(-1, 4, 4).transpose(0, 2, 1).copy(), pts, gn.value

    def adjoint(self, g, max_cg: int = 200, tol: float = 1e-10):
        """Implicit gradient through the solved state of window 0: g is dL/dx in tangent coordinates (left twist
        (omega, v) per free key frame, then each free point). Returns (dL/dxy (N_o, 2), dL/dinfo (N_o, 3) as
        (L00, L01 shared, L11), conjugate-gradient iterations, adjoint a)."""
        g = np.ascontiguousarray(g, dtype=np.float64)
        p = self.problems[0]
        n = p.num_observations
        dxy, dinfo, a = np.zeros((n, 2)), np.zeros((n, 3)), np.zeros_like(g)
        it = C.c_int32()
        d = lambda x: x.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
        rc = self.solver._L.sba_batch_adjoint(self.solver._h, self._b, d(g), d(dxy), d(dinfo), max_cg, tol,
                                              C.byref(it), d(a))
        if rc != SBA_OK:
            raise SbaError(rc, self.solver.last_error())
        if it.value >= max_cg:
            import warnings
            warnings.warn(f"adjoint conjugate gradients stopped at max_cg={max_cg} before reaching tol={tol}: the "
                          "Hessian at the solution is (nearly) singular or the solve has not converged", RuntimeWarning)
        return dxy, dinfo, it.value, a

    def close(self):
        if getattr(self, "_b", None):
            self.solver._L.sba_batch_destroy(self.solver._h, self._b)
            self._b = None

    def __del__(self):
        self.close()
