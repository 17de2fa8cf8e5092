"""ctypes binding of libcuvslam_sba.so (the C ABI declared in include/cuvslam_sba.h)."""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcuvslam_sba.so")

SBA_OK, SBA_E_EMPTY, SBA_E_INFEASIBLE, SBA_E_DEVICE, SBA_E_CAPACITY, SBA_E_INVALID, SBA_E_NO_PROGRESS = range(7)
LOSS_STUDENT_T, LOSS_HUBER = 0, 1

_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class ProblemView(C.Structure):
    _fields_ = [
        ("num_points", C.c_int32), ("num_fixed_points", C.c_int32), ("num_observations", C.c_int32),
        ("num_poses", C.c_int32), ("num_fixed_key_frames", C.c_int32), ("num_cameras", C.c_int32),
        ("points", _fp), ("observation_xys", _fp), ("observation_infos", _fp),
        ("point_ids", _ip), ("pose_ids", _ip), ("camera_ids", C.POINTER(C.c_int8)),
        ("rig_from_world", _fp), ("camera_from_rig", _fp),
        ("max_iterations", C.c_int32), ("robustifier_scale", C.c_float),
    ]


class Result(C.Structure):
    _fields_ = [
        ("points", _fp), ("rig_from_world", _fp), ("iterations", C.c_int32), ("initial_cost", C.c_float),
        ("last_cost", C.c_float), ("status", C.c_int32), ("initial_cost_f64", C.c_double),
        ("last_cost_f64", C.c_double), ("lambda_", C.c_float), ("accepted", C.c_int32),
    ]


class Options(C.Structure):
    _fields_ = [
        ("device", C.c_int32), ("loss", C.c_int32), ("fixed_iterations", C.c_int32), ("use_graph", C.c_int32),
        ("chunk_pairs", C.c_int32), ("pack_threads", C.c_int32), ("bchunk_pairs", C.c_int32),
        ("reserved", C.c_int32 * 5),
    ]


class SceneConfig(C.Structure):
    _fields_ = [
        ("num_keyframes", C.c_int32), ("num_points", C.c_int32), ("sample_observations", C.c_int32),
        ("num_cameras", C.c_int32), ("num_fixed_keyframes", C.c_int32), ("fx", C.c_double), ("fy", C.c_double), ("cx", C.c_double), ("cy", C.c_double),
        ("width", C.c_int32), ("height", C.c_int32), ("keyframe_spacing", C.c_double),
        ("keyframe_rotation_deg", C.c_double), ("pixel_noise", C.c_double), ("outlier_rate", C.c_double),
        ("pose_sigma_m", C.c_double), ("pose_sigma_deg", C.c_double), ("point_sigma_m", C.c_double),
        ("huber_px", C.c_double), ("xy_extent", C.c_double), ("z_min", C.c_double), ("z_max", C.c_double),
        ("seed", C.c_uint64),
    ]


class LoopConfig(C.Structure):
    _fields_ = [
        ("num_keyframes", C.c_int32), ("num_points", C.c_int32), ("num_stereo_pairs", C.c_int32),
        ("num_fixed_keyframes", C.c_int32), ("loop_length_m", C.c_double), ("stereo_baseline_m", C.c_double),
        ("corridor_half_m", C.c_double), ("floor_y_m", C.c_double), ("ceiling_y_m", C.c_double),
        ("clutter_fraction", C.c_double), ("max_range_m", C.c_double), ("fx", C.c_double), ("fy", C.c_double),
        ("cx", C.c_double), ("cy", C.c_double), ("width", C.c_int32), ("height", C.c_int32),
        ("keyframe_rotation_deg", C.c_double), ("pixel_noise", C.c_double), ("outlier_rate", C.c_double),
        ("pose_sigma_m", C.c_double), ("pose_sigma_deg", C.c_double), ("point_sigma_m", C.c_double),
        ("huber_px", C.c_double), ("seed", C.c_uint64),
    ]


# include/cuvslam_sba_imu.h
class ImuMeasurement(C.Structure):
    _fields_ = [("time_ns", C.c_int64), ("linear_acceleration", C.c_float * 3), ("angular_velocity", C.c_float * 3)]


class ImuCalibration(C.Structure):
    _fields_ = [("rig_from_imu", C.c_float * 16), ("gyroscope_noise_density", C.c_float),
                ("gyroscope_random_walk", C.c_float), ("accelerometer_noise_density", C.c_float),
                ("accelerometer_random_walk", C.c_float), ("frequency", C.c_float)]


class ImuPreintegration(C.Structure):
    _fields_ = [("cov", C.c_float * 81), ("acc_rw_cov", C.c_float * 9), ("gyro_rw_cov", C.c_float * 9),
                ("dR", C.c_float * 9), ("dV", C.c_float * 3), ("dP", C.c_float * 3), ("JRg", C.c_float * 9),
                ("JVg", C.c_float * 9), ("JVa", C.c_float * 9), ("JPg", C.c_float * 9), ("JPa", C.c_float * 9),
                ("dT_s", C.c_float), ("gyro_bias", C.c_float * 3), ("acc_bias", C.c_float * 3),
                ("gyro_bias_updated", C.c_float * 3), ("acc_bias_updated", C.c_float * 3),
                ("first_measurement", C.c_int32), ("num_measurements", C.c_int32)]


class ImuProblemView(C.Structure):
    _fields_ = [
        ("num_points", C.c_int32), ("num_observations", C.c_int32), ("num_poses", C.c_int32),
        ("num_fixed_key_frames", C.c_int32), ("num_cameras", C.c_int32), ("num_measurements", C.c_int32),
        ("points", _fp), ("observation_xys", _fp), ("observation_infos", _fp),
        ("point_ids", _ip), ("pose_ids", _ip), ("camera_ids", C.POINTER(C.c_int8)),
        ("camera_from_rig", _fp), ("w_from_imu", _fp), ("velocity", _fp), ("gyro_bias", _fp), ("acc_bias", _fp),
        ("preintegration", C.POINTER(ImuPreintegration)), ("measurements", C.POINTER(ImuMeasurement)),
        ("gravity", C.c_float * 3), ("max_iterations", C.c_int32), ("robustifier_scale", C.c_float),
        ("robustifier_scale_pose", C.c_float), ("prior_gyro", C.c_float), ("prior_acc", C.c_float),
        ("imu_penalty", C.c_float), ("acc_rw_penalty", C.c_float), ("boundary_imu_penalty", C.c_float),
        ("reintegration_thresh", C.c_float),
    ]


class ImuResult(C.Structure):
    _fields_ = [
        ("points", _fp), ("w_from_imu", _fp), ("velocity", _fp), ("gyro_bias", _fp), ("acc_bias", _fp),
        ("preintegration", C.POINTER(ImuPreintegration)), ("status", C.c_int32), ("iterations", C.c_int32),
        ("accepted", C.c_int32), ("initial_cost", C.c_float), ("lambda_", C.c_float), ("pad", C.c_int32),
        ("initial_cost_f64", C.c_double), ("last_cost_f64", C.c_double),
    ]


# include/cuvslam_sba_mono.h
class MonoWindow(C.Structure):
    _fields_ = [("num_frames", C.c_int32), ("num_tracks", C.c_int32), ("num_obs", C.c_int32),
                ("constrained", _ip), ("frame_poses", _fp), ("points", _fp), ("obs_frame", _ip),
                ("obs_track", _ip), ("obs_xy", _fp), ("max_iterations", C.c_int32), ("min_residual", C.c_float),
                ("min_residual_speed", C.c_float)]


class MonoResult(C.Structure):
    _fields_ = [("frame_poses", _fp), ("points", _fp), ("covariance", C.c_float * 36), ("residual", C.c_double),
                ("residual_initial", C.c_double), ("lambda_", C.c_double), ("iterations", C.c_int32),
                ("failed", C.c_int32), ("feasible", C.c_int32), ("status", C.c_int32)]


# Every exported symbol of include/cuvslam_sba.h, cuvslam_sba_imu.h and cuvslam_sba_mono.h with its signature.
SIGNATURES = {
    "sba_default_options": (None, [C.POINTER(Options)]),
    "sba_create": (C.c_int, [C.POINTER(Options), C.POINTER(C.c_void_p)]),
    "sba_destroy": (C.c_int, [C.c_void_p]),
    "sba_last_error": (C.c_char_p, [C.c_void_p]),
    "sba_abi_version": (C.c_int, []),
    "sba_solve": (C.c_int, [C.c_void_p, C.POINTER(ProblemView), C.POINTER(Result)]),
    "sba_solve_noalloc": (C.c_int, [C.c_void_p, C.POINTER(ProblemView), C.POINTER(Result)]),
    "sba_solve_batch": (C.c_int, [C.c_void_p, C.POINTER(ProblemView), C.c_int32, C.POINTER(Result)]),
    "sba_batch_create": (C.c_int, [C.c_void_p, C.POINTER(ProblemView), C.c_int32, C.POINTER(C.c_void_p)]),
    "sba_batch_run": (C.c_int, [C.c_void_p, C.c_void_p]),
    "sba_batch_sync": (C.c_int, [C.c_void_p]),
    "sba_batch_read": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Result)]),
    "sba_batch_destroy": (C.c_int, [C.c_void_p, C.c_void_p]),
    "sba_batch_info": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64), _ip, _ip]),
    "sba_batch_kernel_times": (C.c_int, [C.c_void_p, C.c_void_p, _fp, _ip]),
    "sba_batch_debug_step": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, _dp, _dp, _dp, _dp, _dp]),
    "sba_batch_winsorize": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_uint8),
                                      C.POINTER(C.c_int32)]),
    "sba_debug_factor": (C.c_int, [C.c_void_p, C.c_int32, _dp, _dp, _dp]),
    "sba_batch_trace": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, _dp, _dp, _fp, _ip, _ip]),
    "sba_batch_adjoint": (C.c_int, [C.c_void_p, C.c_void_p, _dp, _dp, _dp, C.c_int32, C.c_double, _ip, _dp]),
    "sba_autotune_chunks": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, _ip, C.c_int32, C.c_int32, _ip, _fp]),
    "sba_autotune_table": (C.c_int, [C.c_void_p, C.c_char_p, C.c_void_p, C.c_int32, C.c_int32, _ip, _ip, _ip]),
    "sba_loop_default_config": (None, [C.POINTER(LoopConfig)]),
    "sba_scene_generate_loop": (C.c_int, [C.POINTER(LoopConfig), C.POINTER(C.c_void_p)]),
    "sba_batch_refine": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_double, _dp, _dp,
                                   _dp]),
    "sba_debug_copy": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, _fp]),
    "sba_debug_mfma_layout": (C.c_int, [C.c_void_p, _fp, _fp, _fp]),
    "sba_scene_default_config": (None, [C.POINTER(SceneConfig)]),
    "sba_scene_generate": (C.c_int, [C.POINTER(SceneConfig), C.POINTER(C.c_void_p)]),
    "sba_scene_view": (C.c_int, [C.c_void_p, C.POINTER(ProblemView)]),
    "sba_scene_ground_truth": (C.c_int, [C.c_void_p, C.POINTER(_fp), C.POINTER(_fp)]),
    "sba_scene_free": (None, [C.c_void_p]),
    "sba_imu_solve": (C.c_int, [C.c_void_p, C.POINTER(ImuCalibration), C.POINTER(ImuProblemView),
                                C.POINTER(ImuResult)]),
    "sba_imu_integrate": (C.c_int, [C.POINTER(ImuCalibration), C.POINTER(ImuMeasurement), C.c_int32, _fp, _fp,
                                    C.POINTER(ImuPreintegration)]),
    "sba_imu_information": (C.c_int, [C.POINTER(ImuPreintegration), _fp, _fp, _fp]),
    "sba_mono_solve_batch": (C.c_int, [C.c_void_p, C.POINTER(MonoWindow), C.c_int32, C.POINTER(MonoResult)]),
}

_lib = None


def lib() -> C.CDLL:
    """Loads the native library; raises if it has not been built (there is no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "(the HIP extension is the only implementation)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib
