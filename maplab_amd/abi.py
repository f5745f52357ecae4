"""ctypes view of the C ABI in include/maplab_ba.h and include/maplab_ba_synth.h.

The structures mirror the C declarations field for field. `load_library()`
loads the HIP library built in-tree (maplab_amd/libmaplab_ba.so) and raises
if it is missing: there is no CPU fallback for the solver.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmaplab_ba.so")
SYNTH_PATH = os.path.join(_HERE, "libmaplab_ba_synth.so")

c_i32 = ctypes.c_int32
c_f64 = ctypes.c_double
c_u64 = ctypes.c_uint64
c_vp = ctypes.c_void_p

MLBA_OK = 0
MLBA_EINVAL = -1
MLBA_ENOMEM = -2
MLBA_EDEVICE = -3
MLBA_ESTATE = -4

DIST_NONE, DIST_RADTAN, DIST_EQUIDISTANT, DIST_FISHEYE = 0, 1, 2, 3
CAMERA_PINHOLE, CAMERA_UNIFIED_PROJECTION = 0, 1
EVAL_WITH_XI = 2
LOCAL_KEYFRAME, LOCAL_MISSION, GLOBAL = 0, 1, 2
LOSS_MAPLAB_HUBER, LOSS_CAUCHY, LOSS_TRIVIAL = 0, 1, 2
SOLVER_SCHUR_PCG = 0
CONVERGENCE, NO_CONVERGENCE, FAILURE, USER_SUCCESS, USER_FAILURE = 0, 1, 2, 3, 4
EVAL_AMBIENT, EVAL_TANGENT = 0, 1

SYNTH_CIRCLE, SYNTH_RANDOM_WALK = 0, 1
ANCHOR_FIRST_OBSERVER, ANCHOR_GLOBAL = 0, 1

AMBIENT_BLOCK_SIZES = (3, 7, 7, 7, 7, 4, 3, 4, 4)
TANGENT_BLOCK_SIZES = (3, 6, 6, 6, 6, 3, 3, 4, 4)


class ProblemDesc(ctypes.Structure):
    _fields_ = [
        ("num_vertices", c_i32), ("num_landmarks", c_i32), ("num_observations", c_i32),
        ("num_cameras", c_i32), ("num_missions", c_i32), ("reserved0", c_i32),
        ("vertex_q_IM__M_p_MI", c_vp), ("baseframe_q_GM__G_p_GM", c_vp), ("landmark_p_B", c_vp),
        ("camera_q_CI__C_p_CI", c_vp), ("camera_intrinsics", c_vp), ("camera_distortion", c_vp),
        ("camera_distortion_type", c_vp), ("camera_resolution", c_vp), ("vertex_mission", c_vp),
        ("landmark_store_vertex", c_vp), ("obs_vertex", c_vp), ("obs_landmark", c_vp),
        ("obs_camera", c_vp), ("obs_keypoint", c_vp), ("obs_sigma", c_vp),
        ("vertex_constant", c_vp), ("landmark_constant", c_vp), ("camera_model", c_vp), ("camera_xi", c_vp),
    ]


class Options(ctypes.Structure):
    _fields_ = [
        ("max_num_iterations", c_i32), ("use_jacobi_scaling", c_i32),
        ("max_solver_time_in_seconds", c_f64), ("function_tolerance", c_f64),
        ("gradient_tolerance", c_f64), ("parameter_tolerance", c_f64),
        ("initial_trust_region_radius", c_f64), ("max_trust_region_radius", c_f64),
        ("min_trust_region_radius", c_f64), ("min_relative_decrease", c_f64),
        ("min_lm_diagonal", c_f64), ("max_lm_diagonal", c_f64),
        ("max_num_consecutive_invalid_steps", c_i32), ("loss_type", c_i32),
        ("loss_scale_px", c_f64), ("linear_solver_type", c_i32), ("pcg_max_iterations", c_i32),
        ("pcg_relative_tolerance", c_f64), ("reject_outliers_every_n_iters", c_i32),
        ("use_reprojection_error", c_i32), ("max_reprojection_error_px", c_f64),
        ("obs_tile", c_i32), ("schur_row_chunk", c_i32), ("pcg_coarse_nodes", c_i32), ("pcg_ritz_vectors", c_i32),
        ("jacobian_fp32", c_i32), ("pcg_preconditioner", c_i32),
    ]


class IterationSummary(ctypes.Structure):
    _fields_ = [
        ("iteration", c_i32), ("step_is_valid", c_i32), ("step_is_successful", c_i32),
        ("linear_solver_iterations", c_i32), ("cost", c_f64), ("cost_change", c_f64),
        ("gradient_max_norm", c_f64), ("step_norm", c_f64), ("relative_decrease", c_f64),
        ("trust_region_radius", c_f64), ("iteration_time_in_seconds", c_f64),
        ("cumulative_time_in_seconds", c_f64),
    ]


NUM_KERNELS = 12  # MLBA_NUM_KERNELS


class Summary(ctypes.Structure):
    _fields_ = [
        ("termination_type", c_i32), ("num_iterations", c_i32), ("num_successful_steps", c_i32),
        ("num_unsuccessful_steps", c_i32), ("num_residual_blocks", c_i32),
        ("num_removed_landmarks", c_i32), ("initial_cost", c_f64), ("final_cost", c_f64),
        ("fixed_cost", c_f64), ("lm_loop_seconds", c_f64), ("evaluate_seconds", c_f64),
        ("linear_solver_seconds", c_f64), ("total_linear_solver_iterations", ctypes.c_int64),
        ("iterations_capacity", c_i32), ("reserved0", c_i32),
        ("iterations", ctypes.POINTER(IterationSummary)),
        ("kernel_ms", c_f64 * NUM_KERNELS), ("kernel_launches", c_i32 * NUM_KERNELS),
    ]


KERNEL_LINEARIZE, KERNEL_COST, KERNEL_SCHUR_BLOCKS, KERNEL_PCG, KERNEL_BACK_SUBSTITUTE = 0, 1, 2, 3, 4
KERNEL_POSE_ROWS, KERNEL_LANDMARK_PREP, KERNEL_SCHUR_INCIDENCES, KERNEL_SCHUR_FACTORS = 5, 6, 7, 8


class SynthConfig(ctypes.Structure):
    _fields_ = [
        ("kind", c_i32), ("num_keyframes", c_i32), ("num_landmarks", c_i32), ("distortion_type", c_i32),
        ("distortion", c_f64 * 4), ("intrinsics", c_f64 * 4), ("width", c_i32), ("height", c_i32),
        ("p_C_I", c_f64 * 3), ("keypoint_sigma_px", c_f64), ("outlier_fraction", c_f64),
        ("rot_noise_rad", c_f64), ("pos_noise_m", c_f64), ("landmark_noise_m", c_f64),
        ("step_m", c_f64), ("yaw_drift_rad", c_f64), ("min_dist_m", c_f64), ("max_dist_m", c_f64),
        ("max_range_m", c_f64), ("min_depth_m", c_f64), ("box_half_m", c_f64),
        ("circle_radius_m", c_f64), ("min_obs_per_landmark", c_i32), ("max_obs_per_landmark", c_i32),
        ("detection_probability", c_f64), ("anchor_mode", c_i32), ("fix_scale_landmark", c_i32),
        ("seed", c_u64),
    ]


class KernelTimes(ctypes.Structure):
    _fields_ = [(n, c_f64) for n in (
        "linearize_ms", "cost_ms", "pose_rows_ms", "landmark_prep_ms", "schur_factors_ms", "schur_incidences_ms",
        "schur_blocks_ms",
        "block_inverse_ms", "pcg_solve_ms", "pcg_spmv_ms", "pcg_update_ms", "back_substitute_ms")] + [
        (n, ctypes.c_int64) for n in (
            "num_observations", "num_landmarks", "num_variable_vertices", "schur_blocks", "schur_upper_blocks",
            "schur_pairs", "incidences", "linearize_tiles", "pcg_iterations", "pcg_coarse_columns",
            "pcg_window_columns", "schur_factor_blocks", "schur_staged_blocks", "schur_visits", "schur_batches",
            "schur_segments", "schur_tile_pairs")]


class CalibrationOptions(ctypes.Structure):
    _fields_ = [("fix_intrinsics", c_i32), ("fix_distortion", c_i32), ("fix_extrinsics_rotation", c_i32),
                ("fix_extrinsics_translation", c_i32)]


class Schedule(ctypes.Structure):
    _fields_ = [("item_band_rows", c_i32), ("item_waves", c_i32), ("xcd_mapping", c_i32),
                ("schedules_tried", c_i32), ("default_ms", c_f64), ("tuned_ms", c_f64), ("pcg_partials", c_i32),
                ("reserved0", c_i32)]


class QualityOptions(ctypes.Structure):
    _fields_ = [("min_observers", c_i32), ("reserved0", c_i32), ("min_observation_angle_deg", c_f64),
                ("max_distance_from_closest_observer", c_f64), ("min_distance_from_closest_observer", c_f64),
                ("max_reprojection_error_px", c_f64)]


class ViProblemOptions(ctypes.Structure):
    _fields_ = [("add_visual_constraints", c_i32), ("fix_landmark_positions", c_i32), ("fix_vertices", c_i32),
                ("min_landmarks_per_frame", c_i32), ("enable_visual_outlier_rejection", c_i32),
                ("reserved0", c_i32), ("solver", Options), ("landmark_quality", QualityOptions),
                ("calibration", CalibrationOptions)]


class ProblemSelection(ctypes.Structure):
    _fields_ = [("num_observations", c_i32), ("num_vertices", c_i32), ("num_landmarks", c_i32),
                ("num_clusters", c_i32), ("obs_in_problem", c_vp), ("vertex_in_problem", c_vp),
                ("landmark_in_problem", c_vp), ("vertex_constant", c_vp), ("landmark_constant", c_vp),
                ("gauge_vertex", c_vp), ("gauge_landmark", c_vp)]


class InertialProblem(ctypes.Structure):
    _fields_ = [("num_vertices", c_i32), ("num_edges", c_i32), ("num_samples", c_i32), ("reserved0", c_i32),
                ("vertex_q_IM__M_p_MI", c_vp), ("vertex_gyro_bias", c_vp), ("vertex_velocity_M", c_vp),
                ("vertex_accel_bias", c_vp), ("edge_from", c_vp), ("edge_to", c_vp), ("edge_sample_begin", c_vp),
                ("imu", c_vp), ("timestamps_ns", c_vp), ("gyro_noise_sigma", c_f64), ("gyro_bias_sigma", c_f64),
                ("acc_noise_sigma", c_f64), ("acc_bias_sigma", c_f64), ("gravity_magnitude", c_f64)]


INERTIAL_BLOCK_SIZES = (7, 3, 3, 3, 7, 3, 3, 3)

QUALITY_UNKNOWN, QUALITY_BAD, QUALITY_GOOD, QUALITY_LOCALIZATION_SUMMARY = 0, 1, 2, 3


def _declare(lib, name, restype, argtypes):
    fn = getattr(lib, name)
    fn.restype = restype
    fn.argtypes = argtypes
    return fn


# Every symbol include/maplab_ba.h declares, with its ctypes signature.
LIB_SYMBOLS = {
    "mlba_abi_version": (c_i32, []),
    "mlba_options_init": (None, [ctypes.POINTER(Options)]),
    "mlba_create": (c_i32, [ctypes.POINTER(ProblemDesc), ctypes.POINTER(c_vp)]),
    "mlba_destroy": (c_i32, [c_vp]),
    "mlba_last_error": (ctypes.c_char_p, [c_vp]),
    "mlba_upload_state": (c_i32, [c_vp, c_vp, c_vp]),
    "mlba_download_state": (c_i32, [c_vp, c_vp, c_vp]),
    "mlba_set_landmark_active": (c_i32, [c_vp, c_vp]),
    "mlba_get_landmark_active": (c_i32, [c_vp, c_vp]),
    "mlba_evaluate": (c_i32, [c_vp, c_i32, ctypes.POINTER(Options), c_vp, c_vp, c_vp]),
    "mlba_evaluate_noalloc": (c_i32, [c_vp, c_i32, ctypes.POINTER(Options), c_vp, c_vp, c_vp]),
    "mlba_measure_copy": (c_i32, [ctypes.c_int64, c_i32, ctypes.POINTER(c_f64)]),
    "mlba_solve": (c_i32, [c_vp, ctypes.POINTER(Options), ctypes.POINTER(Summary)]),
    "mlba_solve_with_outlier_rejection": (c_i32, [c_vp, ctypes.POINTER(Options), ctypes.POINTER(Summary)]),
    "mlba_solve_many": (c_i32, [ctypes.POINTER(c_vp), ctypes.POINTER(ctypes.POINTER(Options)),
                                ctypes.POINTER(ctypes.POINTER(Summary)), ctypes.POINTER(c_i32), c_i32, c_i32,
                                c_i32]),
    "mlba_time_kernels": (c_i32, [c_vp, ctypes.POINTER(Options), c_i32, ctypes.POINTER(KernelTimes)]),
    "mlba_calibration_options_init": (None, [ctypes.POINTER(CalibrationOptions)]),
    "mlba_solve_calibration": (c_i32, [c_vp, ctypes.POINTER(Options), ctypes.POINTER(CalibrationOptions),
                                       ctypes.POINTER(Summary)]),
    "mlba_download_cameras": (c_i32, [c_vp, c_vp, c_vp, c_vp, c_vp]),
    "mlba_autotune": (c_i32, [c_vp, ctypes.POINTER(Options), c_i32, ctypes.POINTER(Schedule)]),
    "mlba_set_schedule": (c_i32, [c_vp, ctypes.POINTER(Schedule)]),
    "mlba_get_schedule": (c_i32, [c_vp, ctypes.POINTER(Schedule)]),
    "mlba_quality_options_init": (None, [ctypes.POINTER(QualityOptions)]),
    "mlba_landmark_quality": (c_i32, [c_vp, ctypes.POINTER(QualityOptions), c_vp]),
}

# Every symbol include/maplab_ba_optimizer.h declares (exported by the same library).
OPTIMIZER_SYMBOLS = {
    "mlba_vi_problem_options_init": (None, [ctypes.POINTER(ViProblemOptions)]),
    "mlba_select_problem": (c_i32, [ctypes.POINTER(ProblemDesc), ctypes.POINTER(ViProblemOptions), c_vp, c_vp,
                                    ctypes.POINTER(ProblemSelection)]),
    "mlba_vi_map_optimize": (c_i32, [ctypes.POINTER(ProblemDesc), ctypes.POINTER(ViProblemOptions), c_vp, c_vp,
                                     c_vp, c_vp, ctypes.POINTER(Summary)]),
    "mlba_vi_map_optimize_cameras": (c_i32, [ctypes.POINTER(ProblemDesc), ctypes.POINTER(ViProblemOptions), c_vp,
                                             c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, ctypes.POINTER(Summary)]),
    "mlba_optimizer_last_error": (ctypes.c_char_p, []),
}

# Every symbol include/maplab_ba_inertial.h declares (exported by the same library).
class InertialOptions(ctypes.Structure):
    _fields_ = [("fix_gyro_bias", c_i32), ("fix_accel_bias", c_i32), ("fix_velocity", c_i32), ("reserved0", c_i32)]


INERTIAL_SYMBOLS = {
    "mlba_inertial_evaluate": (c_i32, [ctypes.POINTER(InertialProblem), c_vp, c_vp]),
    "mlba_inertial_options_init": (None, [ctypes.POINTER(InertialOptions)]),
    "mlba_solve_inertial_states": (c_i32, [ctypes.POINTER(InertialProblem), ctypes.POINTER(Options),
                                           ctypes.POINTER(InertialOptions), c_vp, c_vp, c_vp,
                                           ctypes.POINTER(Summary)]),
    "mlba_inertial_last_error": (ctypes.c_char_p, []),
}

# Every symbol include/maplab_ba_learned.h declares (exported by the same library).
PRECOND_DEFLATED, PRECOND_BLOCK_JACOBI, PRECOND_LEARNED, PRECOND_LEARNED_DEFLATED = 0, 1, 2, 3
GNN_NODE_FEATURES, GNN_EDGE_FEATURES, GNN_NORM_FLOATS = 9, 3, 24


class PreconditionerStats(ctypes.Structure):
    _fields_ = [("solves", ctypes.c_int64), ("iterations", ctypes.c_int64), ("fallbacks", ctypes.c_int64),
                ("fallback_iterations", ctypes.c_int64)]


class GnnDataset(ctypes.Structure):
    _fields_ = [("nodes", c_i32), ("slots", c_i32), ("row_ptr", c_vp), ("col", c_vp), ("rev", c_vp),
                ("node_features", c_vp), ("edge_features", c_vp), ("qc", c_vp)]


LEARNED_SYMBOLS = {
    "mlba_gnn_num_params": (c_i32, []),
    "mlba_set_preconditioner_model": (c_i32, [c_vp, c_vp, c_vp]),
    "mlba_get_preconditioner_stats": (c_i32, [c_vp, ctypes.POINTER(PreconditionerStats)]),
    "mlba_reset_preconditioner_stats": (c_i32, [c_vp]),
    "mlba_rcs_size": (c_i32, [c_vp, ctypes.POINTER(c_i32), ctypes.POINTER(c_i32)]),
    "mlba_rcs_system": (c_i32, [c_vp, ctypes.POINTER(Options), c_f64, c_vp, c_vp, c_vp, c_vp]),
    "mlba_rcs_records": (c_i32, [c_vp, ctypes.POINTER(Options), c_f64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "mlba_rcs_pcg_iterations": (c_i32, [c_vp, ctypes.POINTER(Options), c_f64, c_i32, c_f64, c_i32,
                                        ctypes.POINTER(c_i32)]),
    "mlba_gnn_evaluate": (c_i32, [ctypes.POINTER(GnnDataset), c_vp, c_vp, c_vp, c_vp]),
    "mlba_gnn_train": (c_i32, [ctypes.POINTER(GnnDataset), c_i32, ctypes.c_float, c_vp, c_vp]),
    "mlba_gnn_last_error": (ctypes.c_char_p, []),
}

# Every symbol include/maplab_ba_vimap.h declares (exported by the same library).
VIMAP_SYMBOLS = {
    "mlba_vimap_load": (c_i32, [ctypes.c_char_p, ctypes.POINTER(c_vp)]),
    "mlba_vimap_free": (None, [c_vp]),
    "mlba_vimap_describe": (c_i32, [c_vp, ctypes.POINTER(ProblemDesc)]),
    "mlba_vimap_landmark_quality": (c_i32, [c_vp, c_vp]),
    "mlba_vimap_save": (c_i32, [c_vp, ctypes.c_char_p, c_vp, c_vp, c_vp]),
    "mlba_vimap_last_error": (ctypes.c_char_p, []),
}

SYNTH_SYMBOLS = {
    "mlba_synth_config_init": (c_i32, [ctypes.POINTER(SynthConfig), c_i32]),
    "mlba_synth_generate": (c_i32, [ctypes.POINTER(SynthConfig), ctypes.POINTER(c_vp)]),
    "mlba_synth_free": (None, [c_vp]),
    "mlba_synth_describe": (c_i32, [c_vp, ctypes.POINTER(ProblemDesc)]),
    "mlba_synth_true_vertices": (ctypes.POINTER(c_f64), [c_vp]),
    "mlba_synth_true_landmarks": (ctypes.POINTER(c_f64), [c_vp]),
}

_lib = None
_synth = None


def load_library(path: str = LIB_PATH):
    """Load libmaplab_ba.so; raises OSError when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise OSError(f"{path} is missing: build it with `make -C maplab_amd/csrc` "
                          "(the visual BA path has no CPU fallback)")
        lib = ctypes.CDLL(path)
        for name, (res, args) in {**LIB_SYMBOLS, **OPTIMIZER_SYMBOLS, **INERTIAL_SYMBOLS, **LEARNED_SYMBOLS,
                                  **VIMAP_SYMBOLS}.items():
            _declare(lib, name, res, args)
        _lib = lib
    return _lib


def load_synth(path: str = SYNTH_PATH):
    global _synth
    if _synth is None:
        if not os.path.exists(path):
            raise OSError(f"{path} is missing: build it with `make -C maplab_amd/csrc`")
        lib = ctypes.CDLL(path)
        for name, (res, args) in SYNTH_SYMBOLS.items():
            _declare(lib, name, res, args)
        _synth = lib
    return _synth


def default_quality_options() -> QualityOptions:
    """maplab's elq_* flag defaults (mlba_quality_options_init)."""
    o = QualityOptions()
    load_library().mlba_quality_options_init(ctypes.byref(o))
    return o


def default_options() -> Options:
    """maplab's flag defaults (solver-options.cc:3-52), filled by the library."""
    o = Options()
    load_library().mlba_options_init(ctypes.byref(o))
    return o


def check(rc: int, handle=None) -> None:
    if rc != MLBA_OK:
        msg = load_library().mlba_last_error(handle)
        raise RuntimeError(f"maplab_ba error {rc}: {msg.decode() if msg else ''}")
