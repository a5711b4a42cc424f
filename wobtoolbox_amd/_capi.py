"""ctypes binding of libwob_hip.so (include/wob_c.h).

The product path: there is no CPU fallback.  If the library is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libwob_hip.so")

WOB_OK = 0
STATUS_NAMES = {0: "WOB_OK", 1: "WOB_INVALID_ARG", 2: "WOB_INVALID_ENUM", 3: "WOB_UNSUPPORTED_COMBO", 4: "WOB_OOM",
                5: "WOB_DEVICE", 6: "WOB_CANCELLED"}

# every symbol declared in include/wob_c.h
EXPORTS = ["wob_scene_create", "wob_scene_create_ex", "wob_scene_destroy", "wob_scene_get_info", "wob_scene_set_volume_grid",
           "wob_scene_set_volume_gaussians",
           "wob_estimate_domain", "wob_estimate_domain_noalloc", "wob_estimate_boundary", "wob_intersect_lines", "wob_eval_kernels",
           "wob_eval_weights", "wob_philox4x32_10", "wob_string_to_enum", "wob_enum_to_string",
           "wob_last_error_message", "wob_version", "wob_batch_create", "wob_batch_run", "wob_batch_tune",
           "wob_batch_destroy", "wob_cv_create", "wob_cv_destroy", "wob_cv_get_info", "wob_cv_get_params",
           "wob_cv_set_params", "wob_cv_train", "wob_cv_gradient", "wob_cv_infer", "wob_scene_set_control_variate",
           "wob_sample_boundary", "wob_sample_lines", "wob_vpl_paths", "wob_vpl_aggregate",
           "wob_closest_points", "wob_wos_estimate", "wob_scene_fit_harmonic_cv", "wob_harmonic_cv_targets",
           "wob_scene_set_harmonic_cv", "wob_scene_get_harmonic_cv", "wob_eval_harmonic_cv", "wob_tune_bvh_width"]


class WobError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class SceneDesc(C.Structure):
    _fields_ = [("dim", C.c_uint32), ("n_elements", C.c_uint32), ("vertices", C.c_void_p), ("normals", C.c_void_p),
                ("boundary_values", C.c_void_p), ("boundary_types", C.c_void_p), ("robin_alphas", C.c_void_p)]


class SceneInfo(C.Structure):
    _fields_ = [("dim", C.c_uint32), ("n_elements", C.c_uint32), ("n_nodes", C.c_uint32), ("depth", C.c_uint32),
                ("total_area", C.c_float), ("total_boundary_value", C.c_float), ("device_bytes", C.c_uint64),
                ("stack_need", C.c_uint32), ("bvh_width", C.c_uint32)]


class EstimatorDesc(C.Structure):
    _fields_ = [("equation_kernel", C.c_int32), ("bie_model", C.c_int32), ("domain", C.c_int32),
                ("estimator", C.c_int32), ("problem", C.c_int32), ("path_length", C.c_uint32),
                ("num_resampling_candidates", C.c_uint32), ("num_volume_samples", C.c_uint32), ("eps", C.c_float),
                ("first_kind_equation_scaling_constant", C.c_float), ("flags", C.c_uint32),
                ("cv_samples", C.c_uint32)]


class HcvDesc(C.Structure):
    _fields_ = [("n_sources", C.c_uint32), ("radius_scale", C.c_float), ("n_train", C.c_uint32),
                ("volume_samples", C.c_uint32), ("eps", C.c_float), ("ridge", C.c_float), ("seed", C.c_uint64),
                ("stream_id", C.c_uint32), ("flags", C.c_uint32)]


HCV_EXTERIOR = 1
HCV_ZERO_CHARGE = 2
EST_EXCLUDE_ORIGIN_ELEMENT = 1
EST_CONTROL_VARIATE = 2
EST_BOUNDARY_DENSITY = 4
EST_INDEPENDENT_VOLUME_SAMPLES = 8
EST_HARMONIC_CV = 16
VOLUME_NO_MASK = 1
SCENE_DEVICE_BVH = 1
SCENE_BVH_WIDTH_SHIFT = 4


def scene_bvh_width(w):
    """WOB_SCENE_BVH_WIDTH(w)"""
    return int(w) << SCENE_BVH_WIDTH_SHIFT


class BatchEntry(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("est", EstimatorDesc), ("d_points", C.c_void_p), ("n_points", C.c_uint32),
                ("paths_per_point", C.c_uint32), ("point_base", C.c_uint32), ("stream_id", C.c_uint32),
                ("path_base", C.c_uint64), ("d_sum", C.c_void_p), ("d_sumsq", C.c_void_p)]


class BatchConfig(C.Structure):
    _fields_ = [("boundary", C.c_uint32), ("paths_per_lane", C.c_uint32), ("block_size", C.c_uint32),
                ("persistent", C.c_uint32), ("use_graph", C.c_uint32), ("schedule", C.c_uint32),
                ("regen_threshold", C.c_uint32)]


class LaunchConfig(C.Structure):
    _fields_ = [("paths_per_lane", C.c_uint32), ("block_size", C.c_uint32), ("flags", C.c_uint32),
                ("walk_ms", C.c_float), ("used_paths_per_lane", C.c_uint32), ("used_work_items", C.c_uint32),
                ("schedule", C.c_uint32), ("regen_threshold", C.c_uint32), ("used_schedule", C.c_uint32)]


LAUNCH_TIME_WALK = 1
LAUNCH_ONE_LANE_PER_ITEM = 2
SCHEDULE_AUTO, SCHEDULE_LANE, SCHEDULE_WAVEFRONT = 0, 1, 2


_lib = None


def lib():
    """Load libwob_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m wobtoolbox_amd.build` "
                           "(the walk-on-boundary engine has no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    P, U32, U64, I32, F = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int32, C.c_float
    sig = {
        "wob_scene_create": ([C.POINTER(SceneDesc), C.POINTER(P)], C.c_int),
        "wob_scene_create_ex": ([C.POINTER(SceneDesc), U32, C.POINTER(P)], C.c_int),
        "wob_scene_destroy": ([P], C.c_int),
        "wob_scene_get_info": ([P, C.POINTER(SceneInfo)], C.c_int),
        "wob_scene_set_volume_grid": ([P, U32, F, P], C.c_int),
        "wob_scene_set_volume_gaussians": ([P, U32, P, P, P, U32], C.c_int),
        "wob_estimate_domain": ([P, C.POINTER(EstimatorDesc), P, U32, U32, U32, U64, U64, U32, P, P,
                                 C.POINTER(LaunchConfig), P], C.c_int),
        "wob_estimate_boundary": ([P, C.POINTER(EstimatorDesc), P, U32, U32, U32, U64, U64, U32, P, P,
                                   C.POINTER(LaunchConfig), P], C.c_int),
        "wob_estimate_domain_noalloc": ([P, C.POINTER(EstimatorDesc), P, U32, U32, U32, U64, U64, U32, P, P,
                                         C.POINTER(LaunchConfig), P], C.c_int),
        "wob_intersect_lines": ([P, P, P, U32, U32, P, P, P, P], C.c_int),
        "wob_eval_kernels": ([U32, U32, P, P, P, P, F, P, P], C.c_int),
        "wob_eval_weights": ([U32, I32, I32, F, F, U32, P, P, P, P, P, P, P], C.c_int),
        "wob_philox4x32_10": ([U32, P, P, P, P], C.c_int),
        "wob_string_to_enum": ([C.c_char_p], I32),
        "wob_enum_to_string": ([C.c_char_p, I32], C.c_char_p),
        "wob_last_error_message": ([], C.c_char_p),
        "wob_version": ([], C.c_char_p),
        "wob_batch_create": ([C.POINTER(BatchEntry), U32, C.POINTER(BatchConfig), C.POINTER(P)], C.c_int),
        "wob_batch_run": ([P, U64, U32, P], C.c_int),
        "wob_batch_tune": ([P, U64, C.POINTER(BatchConfig), C.POINTER(C.c_float), P], C.c_int),
        "wob_batch_destroy": ([P], C.c_int),
        "wob_cv_create": ([U32, U32, F, U64, C.POINTER(P)], C.c_int),
        "wob_cv_destroy": ([P], C.c_int),
        "wob_cv_get_info": ([P, C.POINTER(U32), C.POINTER(U32)], C.c_int),
        "wob_cv_get_params": ([P, P, P], C.c_int),
        "wob_cv_set_params": ([P, P, P], C.c_int),
        "wob_cv_train": ([P, P, P, P, U32, U32, U32, F, P, P], C.c_int),
        "wob_cv_gradient": ([P, P, P, P, U32, P, C.POINTER(F), P], C.c_int),
        "wob_cv_infer": ([P, P, U32, P, P], C.c_int),
        "wob_scene_set_control_variate": ([P, P, P, U32, P], C.c_int),
        "wob_sample_boundary": ([P, U32, U64, U32, P, P], C.c_int),
        "wob_sample_lines": ([P, P, U32, U64, U32, U32, P, P, P], C.c_int),
        "wob_vpl_paths": ([P, C.POINTER(EstimatorDesc), U32, U64, U64, U32, P, P, C.POINTER(LaunchConfig), P],
                          C.c_int),
        "wob_vpl_aggregate": ([P, C.POINTER(EstimatorDesc), P, P, U32, P, U32, P, P], C.c_int),
        "wob_closest_points": ([P, P, U32, P, P, P, P], C.c_int),
        "wob_wos_estimate": ([P, P, U32, U32, U32, U64, U64, U32, F, U32, P, P, P], C.c_int),
        "wob_scene_fit_harmonic_cv": ([P, C.POINTER(HcvDesc), C.POINTER(F), P], C.c_int),
        "wob_harmonic_cv_targets": ([P, C.POINTER(HcvDesc), P, P], C.c_int),
        "wob_scene_set_harmonic_cv": ([P, U32, P, F], C.c_int),
        "wob_scene_get_harmonic_cv": ([P, C.POINTER(U32), P, C.POINTER(F)], C.c_int),
        "wob_eval_harmonic_cv": ([P, P, U32, P, P], C.c_int),
        "wob_tune_bvh_width": ([C.POINTER(SceneDesc), U32, C.POINTER(EstimatorDesc), P, U32, U32, U64,
                                C.POINTER(U32), P, P], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def check(status):
    if status != WOB_OK:
        msg = lib().wob_last_error_message().decode(errors="replace")
        raise WobError(status, msg)
