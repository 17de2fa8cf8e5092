"""Python mirror of cuVSLAM's MonoSBASolver (libs/sba/mono_sba_solver.h:90-191) over the batched MI355X C ABI of
include/cuvslam_sba_mono.h. A window is the solver's input after addFrame / addTrack: a dict with constrained (F,),
frame_poses (F, 4, 4) world_from_camera, points (T, 3), obs_frame, obs_track (O,), obs_xy (O, 2) and the
constructor's max_iterations, min_residual, min_residual_speed."""
from __future__ import annotations

import ctypes as C
from typing import Dict, List, Sequence

import numpy as np

from . import _lib
from .sba import SbaError


class MonoSolverGpu:
    """Solves batches of mono windows, one workgroup per window with the whole LM loop on the device."""

    def __init__(self, device: int = 0):
        L = _lib.lib()
        opt = _lib.Options()
        L.sba_default_options(C.byref(opt))
        opt.device = device
        h = C.c_void_p()
        rc = L.sba_create(C.byref(opt), C.byref(h))
        if rc != _lib.SBA_OK:
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

    def solve_batch(self, windows: Sequence[Dict]) -> List[Dict]:
        keep, views, outs = [], (_lib.MonoWindow * len(windows))(), []
        res = (_lib.MonoResult * len(windows))()
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))  # noqa: E731
        fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
        for k, w in enumerate(windows):
            con = np.ascontiguousarray(w["constrained"], np.int32)
            poses = np.ascontiguousarray(np.asarray(w["frame_poses"], np.float32).transpose(0, 2, 1))
            pts = np.ascontiguousarray(w["points"], np.float32).reshape(-1, 3)
            of = np.ascontiguousarray(w["obs_frame"], np.int32)
            ot = np.ascontiguousarray(w["obs_track"], np.int32)
            xy = np.ascontiguousarray(w["obs_xy"], np.float32).reshape(-1, 2)
            keep += [con, poses, pts, of, ot, xy]
            views[k] = _lib.MonoWindow(len(con), len(pts), len(of), ip(con), fp(poses), fp(pts), ip(of), ip(ot), fp(xy),
                                       int(w.get("max_iterations", 1000)), float(w.get("min_residual", 0.001)),
                                       float(w.get("min_residual_speed", 0.001)))
            out_poses = np.zeros((len(con), 16), np.float32)
            out_pts = np.zeros((len(pts), 3), np.float32)
            outs.append((out_poses, out_pts))
            res[k].frame_poses, res[k].points = fp(out_poses), fp(out_pts)
        rc = self._L.sba_mono_solve_batch(self._h, views, len(windows), res)
        if rc != _lib.SBA_OK:
            raise SbaError(rc, self.last_error())
        out = []
        for k, (op, oq) in enumerate(outs):
            r = res[k]
            out.append(dict(status=int(r.status), frame_poses=op.reshape(-1, 4, 4).transpose(0, 2, 1), points=oq,
                            covariance=np.array(r.covariance, np.float64).reshape(6, 6), residual=r.residual,
                            residual_initial=r.residual_initial, lam=r.lambda_, iterations=int(r.iterations),
                            failed=bool(r.failed), feasible=bool(r.feasible)))
        return out

    def solve(self, window: Dict) -> Dict:
        return self.solve_batch([window])[0]
