"""ctypes binding of libsloam_gpu.so (the C ABI declared in include/sloam_gpu.h).

The library is required: importing this module raises if it has not been built
(python -m sloam_amd.build).  There is no CPU fallback.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsloam_gpu.so")

# mirrors of the C structs --------------------------------------------------


class CylFitCfg(C.Structure):
    _fields_ = [
        ("n_hyp", C.c_int32),
        ("thr", C.c_float),
        ("max_radius", C.c_float),
        ("refit_iters", C.c_int32),
        ("seed", C.c_uint64),
        ("stream_tag", C.c_int32),
        ("hyp_chunk", C.c_int32),
        ("tile_points", C.c_int32),
        ("refit_rounds", C.c_int32),
    ]


CYL_MODEL_STRIDE = 8


class GroundCfg(C.Structure):
    _fields_ = [
        ("radii_bins", C.c_int32),
        ("theta_bins", C.c_int32),
        ("min_dist", C.c_double),
        ("max_dist", C.c_double),
        ("retain_thresh", C.c_double),
        ("num_features", C.c_int32),
        ("origin", C.c_float * 3),
    ]


class PlaneFitCfg(C.Structure):
    _fields_ = [
        ("n_hyp", C.c_int32),
        ("thr", C.c_float),
        ("seed", C.c_uint64),
        ("stream_tag", C.c_int32),
        ("reserved", C.c_int32),
    ]


class PnTrainCfg(C.Structure):
    _fields_ = [
        ("lr", C.c_float),
        ("beta1", C.c_float),
        ("beta2", C.c_float),
        ("eps", C.c_float),
        ("w_axis", C.c_float),
        ("w_centre", C.c_float),
        ("w_radius", C.c_float),
        ("reserved", C.c_int32),
    ]


PN_PARAMS = 75399
PN_POINTS = 512
PN_PROPOSAL_STRIDE = 8


class TrellisCfg(C.Structure):
    _fields_ = [
        ("cluster_dist", C.c_float),
        ("max_dist_to_centroid", C.c_float),
        ("min_label_points", C.c_int32),
        ("min_row_points", C.c_int32),
        ("min_vertices", C.c_int32),
        ("max_vertices", C.c_int32),
        ("radius_scale", C.c_double),
    ]


class LineFitCfg(C.Structure):
    _fields_ = [
        ("threshold", C.c_double),
        ("probability", C.c_double),
        ("max_iterations", C.c_int32),
        ("optimize", C.c_int32),
        ("seed", C.c_uint32),
        ("reserved", C.c_int32),
    ]


class ProjCfg(C.Structure):
    _fields_ = [
        ("height", C.c_int32),
        ("width", C.c_int32),
        ("fov_up_deg", C.c_float),
        ("fov_down_deg", C.c_float),
        ("destagger", C.c_int32),
    ]


POSE_MODE_AUTO, POSE_MODE_BATCHED, POSE_MODE_FUSED = 0, 1, 2


class PoseCfg(C.Structure):
    _fields_ = [
        ("max_iters", C.c_int32),
        ("stop_on_convergence", C.c_int32),
        ("huber_delta", C.c_double),
        ("function_tolerance", C.c_double),
        ("gradient_tolerance", C.c_double),
        ("parameter_tolerance", C.c_double),
        ("initial_radius", C.c_double),
        ("min_relative_decrease", C.c_double),
        ("parametrization", C.c_int32),
        ("precision", C.c_int32),
    ]

class FrontendCfg(C.Structure):
    _fields_ = [
        ("proj", ProjCfg),
        ("trellis", TrellisCfg),
        ("ground", GroundCfg),
        ("use_graph", C.c_int32),
        ("host_results", C.c_int32),
        ("line", LineFitCfg),
        ("fit_lines", C.c_int32),
        ("reserved", C.c_int32),
    ]


class FrontendHost(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("trellis_counts", "tree_offsets", "vertex_int",
                                          "vertex_f64", "point_offsets", "points", "cell_info",
                                          "cell_model", "cell_features", "line_coeff",
                                          "line_stats")]


class FrontendOut(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("xyzi", "label", "mask", "ground", "n_ground", "tree")]
    _fields_ += [("cap_trees", C.c_int32), ("cap_vertices", C.c_int32), ("cap_points", C.c_int32)]
    _fields_ += [(n, C.c_void_p) for n in ("tree_offsets", "vertex_int", "vertex_f64",
                                           "point_offsets", "points", "trellis_counts")]
    _fields_ += [("n_cells", C.c_int32)]
    _fields_ += [(n, C.c_void_p) for n in ("cell_info", "cell_model", "cell_features",
                                           "kept_offsets", "kept_index", "line_coeff",
                                           "line_stats")]


class KfParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "max_lidar_dist", "default_tree_radius", "max_tree_radius", "max_axis_theta",
        "tree_match_thresh", "add_new_tree_thresh_dist", "min_tree_models", "min_ground_models",
        "min_planes")]
    _fields_ += [(n, C.c_int32) for n in ("features_per_tree", "num_ground_features", "two_step",
                                          "pose_precision")]


class KfInputs(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("trellis_counts", "tree_offsets", "vertex_int",
                                          "vertex_f64", "point_offsets", "points")]
    _fields_ += [("n_cells", C.c_int32), ("num_features", C.c_int32)]
    _fields_ += [(n, C.c_void_p) for n in ("cell_info", "cell_model", "cell_features",
                                           "tree_fits", "pose", "map", "n_map")]


class KfResultView(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("head", "pose", "delta", "lm_tree", "lm_id",
                                          "lm_match", "lm_model")]


_P = C.c_void_p
_I32 = C.c_int32
_I64 = C.c_int64

# name -> (restype, argtypes); every exported symbol of include/sloam_gpu.h
PROTOTYPES = {
    "sloam_ctx_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "sloam_cyl_plan_create_counted": (C.c_int, [_P, _I32, _I64, _P, C.POINTER(CylFitCfg),
                                                C.POINTER(_P)]),
    "sloam_cyl_plan_counts": (C.c_int, [_P, C.POINTER(_P)]),
    "sloam_pn_infer_counted": (C.c_int, [_P, _P, _P, _P, _I32, _P, _P]),
    "sloam_frontend_enqueue": (C.c_int, [_P]),
    "sloam_kf_tree_points": (C.c_int, [_P, C.POINTER(FrontendOut), _I32, _I64, _P, _P, _P]),
    "sloam_kf_estimator_create": (C.c_int, [_P, C.POINTER(KfParams), _I32, _I32, _I32,
                                            C.POINTER(_P)]),
    "sloam_kf_estimator_destroy": (None, [_P]),
    "sloam_kf_estimator_reset": (C.c_int, [_P]),
    "sloam_kf_estimator_exec": (C.c_int, [_P, C.POINTER(KfInputs)]),
    "sloam_kf_estimator_result": (C.c_int, [_P, C.POINTER(KfResultView)]),
    "sloam_ctx_destroy": (None, [_P]),
    "sloam_sync": (C.c_int, [_P]),
    "sloam_strerror": (C.c_char_p, [C.c_int]),
    "sloam_ctx_last_error": (C.c_char_p, [_P]),
    "sloam_ctx_stream": (_P, [_P]),
    "sloam_ctx_enable_timing": (C.c_int, [_P, C.c_int]),
    "sloam_ctx_kernel_time": (C.c_int, [_P, C.c_char_p, C.POINTER(C.c_double), C.POINTER(_I64)]),
    "sloam_ctx_reset_timing": (C.c_int, [_P]),
    "sloam_graph_begin": (C.c_int, [_P]),
    "sloam_graph_end": (C.c_int, [_P, C.POINTER(_P)]),
    "sloam_graph_launch": (C.c_int, [_P]),
    "sloam_graph_destroy": (None, [_P]),
    "sloam_cyl_plan_create": (C.c_int, [_P, _P, _I32, C.POINTER(CylFitCfg), C.POINTER(_P)]),
    "sloam_cyl_plan_reset": (C.c_int, [_P, _P, _I32]),
    "sloam_cyl_plan_destroy": (None, [_P]),
    "sloam_cyl_plan_exec": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "sloam_pose_plan_create": (C.c_int, [_P, _P, _I32, _I32, _I32, C.POINTER(PoseCfg),
                                         C.POINTER(_P)]),
    "sloam_pose_plan_destroy": (None, [_P]),
    "sloam_pose_plan_exec": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "sloam_pose_plan_reset": (C.c_int, [_P, _P, _I32, _I32, _I32, C.POINTER(PoseCfg)]),
    "sloam_pose_plan_set_mode": (C.c_int, [_P, _I32]),
    "sloam_pose_plan_uses_fused": (C.c_int, [_P]),
    "sloam_project_partition": (C.c_int, [_P, _P, _P, _I64, C.POINTER(ProjCfg), _P, _P, _P, _P,
                                          _P, _P, _P]),
    "sloam_project_partition_noalloc": (C.c_int, [_P, _P, _P, _I64, C.POINTER(ProjCfg), _P, _P,
                                                  _P, _P, _P, _P, _P]),
    "sloam_ground_planes": (C.c_int, [_P, _P, _I32, _P, C.POINTER(GroundCfg), _P, _P, _P, _P,
                                      _P]),
    "sloam_plane_plan_create": (C.c_int, [_P, _P, _I32, C.POINTER(PlaneFitCfg), C.POINTER(_P)]),
    "sloam_plane_plan_destroy": (None, [_P]),
    "sloam_fit_lines": (C.c_int, [_P, _P, _P, _I32, C.POINTER(LineFitCfg), _P, _P, _P]),
    "sloam_trellis": (C.c_int, [_P, _P, _I32, _I32, C.POINTER(TrellisCfg), _P, _I32, _I32, _I32,
                                _P, _P, _P, _P, _P, _P]),
    "sloam_memcpy_async": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "sloam_xyz_from_xyzi": (C.c_int, [_P, _P, C.c_int64, _P]),
    "sloam_frontend_create": (C.c_int, [_P, C.POINTER(FrontendCfg), C.POINTER(_P)]),
    "sloam_frontend_destroy": (None, [_P]),
    "sloam_frontend_buffers": (C.c_int, [_P, C.POINTER(FrontendOut)]),
    "sloam_frontend_exec": (C.c_int, [_P]),
    "sloam_frontend_host_view": (C.c_int, [_P, C.POINTER(FrontendHost)]),
    "sloam_cyl_plan_exec_ex": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "sloam_cyl_learned_exec": (C.c_int, [_P, _P, _P, _P, _P, C.c_float, _P, _P, _P, _P, _P, _P]),
    "sloam_cyl_fallback": (C.c_int, [_P, _P, _I32, _P, C.c_float, _P, _P]),
    "sloam_cyl_autotune": (C.c_int, [_P, _P, _I32, C.POINTER(CylFitCfg), _P, _P,
                                     C.POINTER(CylFitCfg)]),
    "sloam_pn_infer": (C.c_int, [_P, _P, _P, _P, _I32, _P]),
    "sloam_pn_loss_grad": (C.c_int, [_P, _P, _P, _P, _I32, _P, C.POINTER(PnTrainCfg), _P, _P]),
    "sloam_pn_train_step": (C.c_int, [_P, _P, _P, _P, _I32, C.POINTER(PnTrainCfg), _P, _P, _I32,
                                      _P, _P]),
    "sloam_plane_plan_exec": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, _P]),
}

_lib = None


def load():
    """Loads the library once and declares every prototype."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m sloam_amd.build` "
            "(the HIP path has no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class SloamError(RuntimeError):
    pass


def check(status, ctx=None):
    if status != 0:
        lib = load()
        msg = lib.sloam_strerror(status).decode()
        if ctx is not None:
            detail = lib.sloam_ctx_last_error(ctx)
            if detail:
                msg += f" ({detail.decode()})"
        raise SloamError(f"sloam status {status}: {msg}")
