"""ctypes mirror of include/sensemap_pm.h and the loader of libsensemap_pm.so.

The product path is the in-tree HIP library; there is no CPU fallback.  If
the library is missing, ``load_engine()`` raises instead of degrading.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")

PM_MAX_NUM_SRC = 32
PM_POSE_STRIDE = 44
PM_DEPTH_RANGES_NOALLOC_MAX_IMAGES = 65535

PM_OK = 0
(PM_ERR_INVALID_ARG, PM_ERR_OUT_OF_MEMORY, PM_ERR_DEVICE, PM_ERR_TOO_MANY_SOURCES, PM_ERR_NOT_READY,
 PM_ERR_OVER_BUDGET, PM_ERR_INTERNAL) = range(1, 8)
STATUS_NAMES = {
    0: "PM_OK", 1: "PM_ERR_INVALID_ARG", 2: "PM_ERR_OUT_OF_MEMORY", 3: "PM_ERR_DEVICE",
    4: "PM_ERR_TOO_MANY_SOURCES", 5: "PM_ERR_NOT_READY", 6: "PM_ERR_OVER_BUDGET", 7: "PM_ERR_INTERNAL",
}
PM_STORAGE_U8, PM_STORAGE_F16 = 0, 1
PM_ACCUM_PLAIN, PM_ACCUM_COMPENSATED = 0, 1
PM_RCP_EXACT, PM_RCP_FAST = 0, 1


class pm_options(C.Structure):
    _fields_ = [
        ("window_radius", C.c_int32),
        ("window_step", C.c_int32),
        ("sigma_spatial", C.c_float),
        ("sigma_color", C.c_float),
        ("num_samples", C.c_int32),
        ("num_iterations", C.c_int32),
        ("geom_consistency", C.c_int32),
        ("geom_consistency_regularizer", C.c_float),
        ("geom_consistency_max_cost", C.c_float),
        ("est_curvature", C.c_int32),
        ("random_optimization", C.c_int32),
        ("local_optimization", C.c_int32),
        ("random_depth_ratio", C.c_float),
        ("random_angle1_range", C.c_float),
        ("random_angle2_range", C.c_float),
        ("init_ncc_matching_cost", C.c_float),
        ("max_ncc_matching_cost", C.c_float),
        ("alpha", C.c_float),
        ("beta", C.c_float),
        ("thk", C.c_int32),
        ("num_good_hypothesis", C.c_int32),
        ("num_bad_hypothesis", C.c_int32),
        ("init_depth_random", C.c_int32),
        ("filter", C.c_int32),
        ("conf_filter", C.c_int32),
        ("median_filter", C.c_int32),
        ("conf_threshold", C.c_float),
        ("depth_diff_threshold", C.c_float),
        ("filter_geom_consistency_max_cost", C.c_float),
        ("filter_min_num_consistent", C.c_int32),
        ("propagate_depth", C.c_int32),
        ("has_prior_depth", C.c_int32),
        ("init_samples", C.c_int32),
        ("compat_lo_geo_bug", C.c_int32),
        ("image_storage", C.c_int32),
        ("accum_mode", C.c_int32),
        ("rcp_mode", C.c_int32),
        ("reserved0", C.c_int32),
        ("reserved1", C.c_int32),
        ("seed", C.c_uint64),
    ]


class pm_camera(C.Structure):
    _fields_ = [("K", C.c_float * 4), ("R", C.c_float * 9), ("T", C.c_float * 3)]


class pm_image(C.Structure):
    _fields_ = [
        ("data", C.POINTER(C.c_uint8)),
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("stride", C.c_int32),
        ("cache_id", C.c_int32),
    ]


class pm_view_desc(C.Structure):
    _fields_ = [
        ("view_id", C.c_uint32),
        ("num_src", C.c_int32),
        ("ref", pm_image),
        ("ref_cam", pm_camera),
        ("src", C.POINTER(pm_image)),
        ("src_cams", C.POINTER(pm_camera)),
        ("depth_min", C.c_float),
        ("depth_max", C.c_float),
        ("src_depth_maps", C.POINTER(C.POINTER(C.c_float))),
        ("init_depth", C.POINTER(C.c_float)),
        ("init_normal", C.POINTER(C.c_float)),
        ("ref_mask", pm_image),
    ]


class pm_outputs(C.Structure):
    _fields_ = [
        ("depth", C.POINTER(C.c_float)),
        ("normal", C.POINTER(C.c_float)),
        ("conf", C.POINTER(C.c_float)),
        ("cost", C.POINTER(C.c_float)),
        ("view_sel", C.POINTER(C.c_uint32)),
        ("curvature", C.POINTER(C.c_uint16)),
    ]


class pm_run_stats(C.Structure):
    _fields_ = [
        ("total_ms", C.c_double),
        ("sweep_ms", C.c_double),
        ("local_opt_ms", C.c_double),
        ("init_ms", C.c_double),
        ("sweep_launches", C.c_int32),
        ("local_opt_launches", C.c_int32),
        ("ncc_evaluations", C.c_uint64),
        ("ncc_samples", C.c_uint64),
        ("geo_evaluations", C.c_uint64),
        ("sweep_ncc_evaluations", C.c_uint64),
        ("sweep_ncc_samples", C.c_uint64),
        ("sweep_geo_evaluations", C.c_uint64),
        ("post_ms", C.c_double),
    ]


# Every symbol the header declares, with its ctypes signature.
(PM_PHASE_INIT, PM_PHASE_SWEEP, PM_PHASE_LOCAL, PM_PHASE_FILTER, PM_PHASE_CURVATURE, PM_PHASE_MASK,
 PM_PHASE_GEOM_GRADIENT, PM_PHASE_MEDIAN) = range(8)


class pm_tuning(C.Structure):
    _fields_ = [("tile_lw", C.c_int32), ("blocks_per_cu", C.c_int32)]


class pm_tune_report(C.Structure):
    _fields_ = [
        ("sweep", pm_tuning),
        ("local_opt", pm_tuning),
        ("sweep_ms_default", C.c_double),
        ("sweep_ms_best", C.c_double),
        ("local_opt_ms_default", C.c_double),
        ("local_opt_ms_best", C.c_double),
        ("candidates", C.c_int32),
        ("reserved0", C.c_int32),
    ]


class pm_server_config(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_active_views", C.c_int32), ("reserved0", C.c_int32)]


class pm_job_stats(C.Structure):
    _fields_ = [("queue_ms", C.c_double), ("latency_ms", C.c_double), ("status", C.c_int32),
                ("reserved0", C.c_int32)]


class pm_server_stats(C.Structure):
    _fields_ = [
        ("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
        ("cache_hits", C.c_uint64), ("cache_misses", C.c_uint64), ("cache_evictions", C.c_uint64),
        ("steps", C.c_uint64), ("launches", C.c_uint64),
        ("arena_capacity", C.c_uint64), ("arena_used", C.c_uint64), ("cache_bytes", C.c_uint64),
        ("active", C.c_int32), ("pending", C.c_int32), ("max_active_seen", C.c_int32), ("reserved0", C.c_int32),
    ]


class pm_cross_options(C.Structure):
    _fields_ = [("filter_geom_consistency_max_cost", C.c_float), ("max_normal_error", C.c_float),
                ("depth_diff_threshold", C.c_float), ("filter_min_num_consistent", C.c_int32)]


class pm_cross_image(C.Structure):
    _fields_ = [("cam", pm_camera), ("width", C.c_int32), ("height", C.c_int32),
                ("depth", C.POINTER(C.c_float)), ("normal_codes", C.POINTER(C.c_uint32))]


class pm_cross_view(C.Structure):
    _fields_ = [("ref", C.c_int32), ("num_src", C.c_int32), ("src", C.POINTER(C.c_int32)),
                ("refined", C.POINTER(C.c_float))]


class pm_dedup_image(C.Structure):
    _fields_ = [("cam", pm_camera), ("width", C.c_int32), ("height", C.c_int32), ("depth", C.POINTER(C.c_float)),
                ("ref_flag", C.c_int32), ("reserved0", C.c_int32), ("visibility", C.POINTER(C.c_uint32)),
                ("dedup_depth", C.POINTER(C.c_float))]


class pm_dedup_view(C.Structure):
    _fields_ = [("ref", C.c_int32), ("num_src", C.c_int32), ("src", C.POINTER(C.c_int32)),
                ("delt", C.POINTER(C.c_float))]


class pm_fusion_options(C.Structure):
    _fields_ = [("step_size", C.c_int32), ("min_num_pixels", C.c_int32), ("max_num_pixels", C.c_int32),
                ("min_num_visible_images", C.c_int32), ("max_traversal_depth", C.c_int32), ("with_normal", C.c_int32),
                ("max_reproj_error", C.c_double), ("max_depth_error", C.c_double), ("max_normal_error", C.c_double)]


class pm_fusion_image(C.Structure):
    _fields_ = [("cam", pm_camera), ("width", C.c_int32), ("height", C.c_int32), ("depth", C.POINTER(C.c_float)),
                ("normal", C.POINTER(C.c_float)), ("rgb", C.POINTER(C.c_uint8)),
                ("overlapping", C.POINTER(C.c_int32)), ("num_overlapping", C.c_int32), ("is_rig", C.c_int32),
                ("used", C.c_int32), ("channels", C.c_int32)]


class pm_fusion_stats(C.Structure):
    _fields_ = [("walks", C.c_int64), ("kept", C.c_int64), ("rounds", C.c_int64), ("big_walks", C.c_int64),
                ("points", C.c_int64), ("total_ms", C.c_double), ("chain", C.c_int64), ("flow_ms", C.c_double)]


class pm_model(C.Structure):
    _fields_ = [("num_images", C.c_int32), ("R", C.POINTER(C.c_float)), ("T", C.POINTER(C.c_float)),
                ("is_rig", C.POINTER(C.c_uint8)), ("num_points", C.c_int64), ("xyz", C.POINTER(C.c_float)),
                ("track_offsets", C.POINTER(C.c_int64)), ("track_images", C.POINTER(C.c_int32))]


class pm_overlap_stats(C.Structure):
    _fields_ = [("pair_records", C.c_int64), ("valid_records", C.c_int64), ("image_pairs", C.c_int64),
                ("fallback_images", C.c_int32), ("reserved0", C.c_int32),
                ("emit_ms", C.c_float), ("sort_ms", C.c_float), ("reduce_ms", C.c_float), ("rank_ms", C.c_float),
                ("total_ms", C.c_float), ("reserved1", C.c_int32)]


_SIGNATURES = {
    "pm_options_init": (C.c_int, [C.POINTER(pm_options)]),
    "pm_options_check": (C.c_int, [C.POINTER(pm_options)]),
    "pm_last_error": (C.c_char_p, []),
    "pm_abi_version": (C.c_int32, []),
    "pm_engine_create": (C.c_int, [C.c_int32, C.POINTER(C.c_void_p)]),
    "pm_engine_destroy": (None, [C.c_void_p]),
    "pm_engine_set_counting": (C.c_int, [C.c_void_p, C.c_int32]),
    "pm_problem_create": (C.c_int, [C.c_void_p, C.POINTER(pm_options), C.POINTER(pm_view_desc),
                                    C.POINTER(C.c_void_p)]),
    "pm_problem_run": (C.c_int, [C.c_void_p]),
    "pm_problem_get_depth_map": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "pm_problem_get_normal_map": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "pm_problem_get_conf_map": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "pm_problem_get_cost_map": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "pm_problem_get_view_sel": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    "pm_problem_get_curvature_map": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint16)]),
    "pm_problem_get_stats": (C.c_int, [C.c_void_p, C.POINTER(pm_run_stats)]),
    "pm_problem_get_poses": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "pm_problem_destroy": (None, [C.c_void_p]),
    "pm_problem_init_from": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pm_problem_set_src_depth": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p]),
    "pm_run_view": (C.c_int, [C.c_void_p, C.POINTER(pm_options), C.POINTER(pm_view_desc),
                              C.POINTER(pm_outputs)]),
    "pm_batch_create": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(pm_options), C.POINTER(pm_view_desc),
                                  C.POINTER(C.c_void_p)]),
    "pm_batch_set_graph": (C.c_int, [C.c_void_p, C.c_int32]),
    "pm_batch_run": (C.c_int, [C.c_void_p]),
    "pm_batch_get_view": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_void_p)]),
    "pm_batch_get_stats": (C.c_int, [C.c_void_p, C.POINTER(pm_run_stats)]),
    "pm_batch_destroy": (None, [C.c_void_p]),
    "pm_batch_set_autotune": (C.c_int, [C.c_void_p, C.c_int32]),
    "pm_engine_set_tuning": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(pm_options), C.c_int32,
                                       C.POINTER(pm_tuning)]),
    "pm_engine_get_tuning": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(pm_options), C.c_int32,
                                       C.POINTER(pm_tuning)]),
    "pm_problem_autotune": (C.c_int, [C.c_void_p, C.POINTER(pm_tune_report)]),
    "pm_engine_save_tuning": (C.c_int, [C.c_void_p, C.c_char_p]),
    "pm_engine_load_tuning": (C.c_int, [C.c_void_p, C.c_char_p]),
    "pm_server_create": (C.c_int, [C.c_void_p, C.POINTER(pm_server_config), C.POINTER(C.c_void_p)]),
    "pm_server_submit": (C.c_int, [C.c_void_p, C.POINTER(pm_options), C.POINTER(pm_view_desc),
                                   C.POINTER(pm_outputs), C.POINTER(C.c_uint64)]),
    "pm_server_poll": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_int32)]),
    "pm_server_wait": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(pm_job_stats)]),
    "pm_server_get_stats": (C.c_int, [C.c_void_p, C.POINTER(pm_server_stats)]),
    "pm_server_destroy": (None, [C.c_void_p]),
    "pm_mat_write": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_void_p]),
    "pm_mat_read_header": (C.c_int, [C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                     C.POINTER(C.c_int32)]),
    "pm_mat_read": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_void_p]),
    "pm_normal_encode": (C.c_int, [C.c_int64, C.POINTER(C.c_float), C.POINTER(C.c_uint32)]),
    "pm_normal_decode": (C.c_int, [C.c_int64, C.POINTER(C.c_uint32), C.POINTER(C.c_float)]),
    "pm_problem_get_normal_codes": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    "pm_problem_write_maps": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p]),
    "pm_problem_run_filter": (C.c_int, [C.c_void_p, C.c_int32]),
    "pm_cross_options_init": (C.c_int, [C.POINTER(pm_cross_options)]),
    "pm_cross_filter": (C.c_int, [C.c_void_p, C.POINTER(pm_cross_options), C.c_int32, C.POINTER(pm_cross_image),
                                  C.c_int32, C.POINTER(pm_cross_view)]),
    "pm_cross_filter_problems": (C.c_int, [C.c_void_p, C.POINTER(pm_cross_options), C.c_int32,
                                           C.POINTER(C.c_void_p), C.c_int32, C.POINTER(pm_cross_view)]),
    "pm_fusion_options_init": (C.c_int, [C.POINTER(pm_fusion_options)]),
    "pm_fusion_run": (C.c_int, [C.c_void_p, C.POINTER(pm_fusion_options), C.c_int32, C.POINTER(pm_fusion_image),
                                C.POINTER(C.c_void_p)]),
    "pm_fusion_get_counts": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "pm_fusion_get_points": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p]),
    "pm_fusion_destroy": (None, [C.c_void_p]),
    "pm_fusion_get_stats": (C.c_int, [C.c_void_p, C.POINTER(pm_fusion_stats)]),
    "pm_model_depth_ranges_noalloc": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_int64,
                                                C.c_void_p, C.c_void_p]),
    "pm_cross_dedup": (C.c_int, [C.c_void_p, C.POINTER(pm_cross_options), C.c_int32, C.POINTER(pm_dedup_image),
                                 C.c_int32, C.POINTER(pm_dedup_view)]),
    "pm_model_max_overlapping_images": (C.c_int, [C.c_void_p, C.POINTER(pm_model), C.c_int32, C.c_double,
                                                  C.POINTER(C.c_uint8), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                                  C.POINTER(pm_overlap_stats)]),
    "pm_model_depth_ranges": (C.c_int, [C.c_void_p, C.POINTER(pm_model), C.POINTER(C.c_float)]),
}

# SENSEMAP_PM_LIB selects another build of the engine (tuning variants).
ENGINE_LIB = os.environ.get("SENSEMAP_PM_LIB") or os.path.join(LIB_DIR, "libsensemap_pm.so")
_engine = None


class PMError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


def load_engine() -> C.CDLL:
    """Load the in-tree HIP engine library; raise if it has not been built."""
    global _engine
    if _engine is None:
        if not os.path.exists(ENGINE_LIB):
            raise RuntimeError(
                f"{ENGINE_LIB} is missing: build it with __graft_entry__.build() "
                "(python -c 'import __graft_entry__ as g; g.build()')")
        lib = C.CDLL(ENGINE_LIB)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _engine = lib
    return _engine


def check(status: int) -> None:
    if status != PM_OK:
        raise PMError(status, load_engine().pm_last_error().decode(errors="replace"))


def default_options(**overrides) -> pm_options:
    o = pm_options()
    check(load_engine().pm_options_init(C.byref(o)))
    for k, v in overrides.items():
        if not hasattr(o, k):
            raise AttributeError(f"pm_options has no field {k}")
        setattr(o, k, v)
    return o
