/*
 * maplab_ba.h -- C ABI of the MI355X visual bundle-adjustment library
 * (libmaplab_ba.so).
 *
 * The library replaces the part of maplab's `optv` / `optvi` path that runs
 * inside `ceres::Solve`:
 *   - VisualReprojectionError::Evaluate for every keypoint observation
 *     (algorithms/ceres-error-terms/include/ceres-error-terms/
 *      visual-error-term-inl.h:13-379), pinhole camera with no / radtan /
 *     equidistant / fisheye distortion (aslam_cv2/aslam_cv_cameras/src/
 *     camera-pinhole.cc:65-164, distortion-radtan.cc:14-91,
 *     distortion-equidistant.cc:14-142, distortion-fisheye.cc:14-120);
 *   - the Levenberg-Marquardt trust-region loop that ceres::Solve runs with
 *     the options of algorithms/map-optimization/include/map-optimization/
 *     solver-options.h:26-74, with landmark (Schur) elimination and the
 *     reduced-camera-system solve;
 *   - the chunked outlier-rejection driver
 *     (algorithms/map-optimization/src/outlier-rejection-solver.cc:28-259).
 *
 * Conventions (identical to maplab's OptimizationStateBuffer,
 * algorithms/map-optimization/src/optimization-state-buffer.cc:153-306):
 *   vertex block    = [qx qy qz qw | px py pz], JPL q_IM and M_p_MI
 *   baseframe block = [q_GM (JPL) | G_p_GM]
 *   camera block    = [q_CI (JPL) | C_p_CI]
 *   landmark        = p_B, expressed in the frame of its store vertex.
 * Quaternions are unit, w >= 0.
 *
 * Ownership: the caller owns every host buffer passed in; the library copies
 * what it needs into device memory at mlba_create and owns device memory
 * until mlba_destroy. One handle may be used by one host thread at a time;
 * different handles are independent. Every entry point returns an MLBA_*
 * status code and never aborts; the text of the last failure is available
 * from mlba_last_error.
 */
#ifndef MAPLAB_BA_H_
#define MAPLAB_BA_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MLBA_ABI_VERSION 1

/* Status codes. */
#define MLBA_OK 0
#define MLBA_EINVAL -1   /* invalid argument (sigma <= 0, bad index, ...) */
#define MLBA_ENOMEM -2   /* device or host allocation failed */
#define MLBA_EDEVICE -3  /* HIP runtime error */
#define MLBA_ESTATE -4   /* state violates an invariant (non-unit quaternion) */

/* Distortion models (aslam::Distortion::Type subset). */
#define MLBA_DISTORTION_NONE 0
#define MLBA_DISTORTION_RADTAN 1
#define MLBA_DISTORTION_EQUIDISTANT 2
#define MLBA_DISTORTION_FISHEYE 3  /* FOV model, one parameter w (distortion[0]), |w| < 1e-16 or
                                      0.5 <= w <= 1.5 (distortion-fisheye.cc:145-156); its
                                      distortion Jacobian block has w in column 0, zeros after */

/* Camera models (aslam::Camera::Type subset used by the visual error term,
 * visual-error-term-factory-inl.h:33-95). */
#define MLBA_CAMERA_PINHOLE 0             /* camera-pinhole.cc:65-164 */
#define MLBA_CAMERA_UNIFIED_PROJECTION 1  /* camera-unified-projection.cc:74-196; mirror
                                             parameter xi from camera_xi, fu fv cu cv from
                                             camera_intrinsics */

/* ceres_error_terms::LandmarkErrorType (common.h:14). */
#define MLBA_LOCAL_KEYFRAME 0
#define MLBA_LOCAL_MISSION 1
#define MLBA_GLOBAL 2

/* Robust loss applied to the sigma-whitened residual. */
#define MLBA_LOSS_MAPLAB_HUBER 0 /* ceres::HuberLoss(delta*sigma), delta 3 (local)
                                    or 10 (global): optimization-terms-addition.cc:189-198 */
#define MLBA_LOSS_CAUCHY 1       /* ceres::CauchyLoss(scale_px / sigma) */
#define MLBA_LOSS_TRIVIAL 2      /* no loss */

/* Linear solver for the LM step. All solve (J'J + D'D) y = J'r. */
#define MLBA_SOLVER_SCHUR_PCG 0 /* landmark Schur elimination + deflated block-Jacobi
                                   PCG on the reduced camera system (device) */

/* ceres::TerminationType */
#define MLBA_CONVERGENCE 0
#define MLBA_NO_CONVERGENCE 1
#define MLBA_FAILURE 2
#define MLBA_USER_SUCCESS 3
#define MLBA_USER_FAILURE 4

/* Evaluation flags for mlba_evaluate. */
#define MLBA_EVAL_AMBIENT 0 /* Jacobians as VisualReprojectionError::Evaluate writes them */
#define MLBA_EVAL_TANGENT 1 /* quaternion columns lifted by the JPL local parameterization */
#define MLBA_EVAL_WITH_XI 2 /* or-ed in: `jacobians` has a 10th entry (nullable) receiving the
                               2 x 1 Jacobian of the unified-projection xi per observation
                               (zero for pinhole cameras); the reference's 5-wide intrinsics
                               block of such a camera is [xi | fu fv cu cv] */

/*
 * A map in structure-of-arrays form. Replaces the VIMap + OptimizationStateBuffer
 * pair that constructOptimizationProblem reads
 * (algorithms/map-optimization/src/vi-optimization-builder.cc:154-374).
 * Observation i is keypoint obs_keypoint[2i..2i+1] of camera obs_camera[i]
 * on vertex obs_vertex[i], associated to landmark obs_landmark[i]. The error
 * type of an observation follows optimization-terms-addition.cc:68-78:
 * observer == store vertex -> local keyframe; same mission -> local mission;
 * otherwise global.
 */
typedef struct mlba_problem_desc {
  int32_t num_vertices;
  int32_t num_landmarks;
  int32_t num_observations;
  int32_t num_cameras;
  int32_t num_missions;
  int32_t reserved0;
  const double* vertex_q_IM__M_p_MI;      /* 7 * num_vertices */
  const double* baseframe_q_GM__G_p_GM;   /* 7 * num_missions */
  const double* landmark_p_B;             /* 3 * num_landmarks */
  const double* camera_q_CI__C_p_CI;      /* 7 * num_cameras */
  const double* camera_intrinsics;        /* 4 * num_cameras: fu fv cu cv */
  const double* camera_distortion;        /* 4 * num_cameras */
  const int32_t* camera_distortion_type;  /* num_cameras, MLBA_DISTORTION_* */
  const int32_t* camera_resolution;       /* 2 * num_cameras: width height */
  const int32_t* vertex_mission;          /* num_vertices */
  const int32_t* landmark_store_vertex;   /* num_landmarks */
  const int32_t* obs_vertex;              /* num_observations */
  const int32_t* obs_landmark;            /* num_observations */
  const int32_t* obs_camera;              /* num_observations */
  const double* obs_keypoint;             /* 2 * num_observations, pixels */
  const double* obs_sigma;                /* num_observations, pixels, > 0 */
  const uint8_t* vertex_constant;         /* num_vertices, nullable (all variable) */
  const uint8_t* landmark_constant;       /* num_landmarks, nullable (all variable) */
  const int32_t* camera_model;            /* num_cameras, MLBA_CAMERA_*, nullable (all pinhole) */
  const double* camera_xi;                /* num_cameras, unified-projection xi >= 0, nullable */
} mlba_problem_desc;

/*
 * Solver options: the subset of ceres::Solver::Options that
 * initSolverOptionsFromFlags sets (solver-options.h:26-74) plus the outlier
 * rejection options (outlier-rejection-solver.h:15-27). mlba_options_init
 * fills maplab's flag defaults (solver-options.cc:3-52).
 */
typedef struct mlba_options {
  int32_t max_num_iterations;             /* --ba_num_iterations = 30 */
  int32_t use_jacobi_scaling;             /* --ba_use_jacobi_scaling = 1 */
  double max_solver_time_in_seconds;      /* --ba_max_time_seconds = 1e6 */
  double function_tolerance;              /* 1e-12 */
  double gradient_tolerance;              /* 1e-10 */
  double parameter_tolerance;             /* 1e-8 */
  double initial_trust_region_radius;     /* 1e5 */
  double max_trust_region_radius;         /* 1e20 */
  double min_trust_region_radius;         /* 1e-32 */
  double min_relative_decrease;           /* ceres default 1e-3 */
  double min_lm_diagonal;                 /* ceres default 1e-6 */
  double max_lm_diagonal;                 /* ceres default 1e32 */
  int32_t max_num_consecutive_invalid_steps; /* ceres default 5 */
  int32_t loss_type;                      /* MLBA_LOSS_* (default MAPLAB_HUBER) */
  double loss_scale_px;                   /* Cauchy scale in pixels (default 1) */
  int32_t linear_solver_type;             /* MLBA_SOLVER_* */
  int32_t pcg_max_iterations;             /* default 500 */
  double pcg_relative_tolerance;          /* ||r_k|| / ||b||, default 1e-10 */
  /* outlier rejection (outlier-rejection-solver.cc:9-18) */
  int32_t reject_outliers_every_n_iters;  /* 3 */
  int32_t use_reprojection_error;         /* 0 */
  double max_reprojection_error_px;       /* 50 */
  /* device launch tuning (0 = library default / autotuned) */
  int32_t obs_tile;                       /* reserved */
  int32_t schur_row_chunk;                /* reserved */
  int32_t pcg_coarse_nodes;               /* deflation hat nodes: 0 automatic, -1 none, else K (<= 15) */
  int32_t pcg_ritz_vectors;               /* recycled Ritz vectors added to the deflation space:
                                             0 automatic (16), -1 none, else r (<= 16) */
  int32_t jacobian_fp32;                  /* 1: the robust-corrected Jacobians are rounded to FP32 and
                                             stored so (88 instead of 160 bytes per observation);
                                             every product of them, S, the PCG, residuals, costs and
                                             LM scalars stay FP64, and S remains the Schur complement
                                             of J'J of that one rounded Jacobian (positive
                                             semi-definite). 0 (default): FP64 Jacobians */
  int32_t pcg_preconditioner;             /* MLBA_PRECOND_* of maplab_ba_learned.h: 0 deflated
                                             block-Jacobi (default), 1 block-Jacobi, 2 learned,
                                             3 learned with the hat deflation */
} mlba_options;

/* One row of the Ceres iteration log (ceres::IterationSummary subset). */
typedef struct mlba_iteration_summary {
  int32_t iteration;
  int32_t step_is_valid;
  int32_t step_is_successful;
  int32_t linear_solver_iterations;
  double cost;
  double cost_change;
  double gradient_max_norm;
  double step_norm;
  double relative_decrease;
  double trust_region_radius;
  double iteration_time_in_seconds;
  double cumulative_time_in_seconds;
} mlba_iteration_summary;

#define MLBA_NUM_KERNELS 12          /* slots of mlba_summary.kernel_ms / kernel_launches */

typedef struct mlba_summary {
  int32_t termination_type;               /* MLBA_CONVERGENCE ... */
  int32_t num_iterations;                 /* rows written (iteration 0 included) */
  int32_t num_successful_steps;
  int32_t num_unsuccessful_steps;
  int32_t num_residual_blocks;            /* active observations */
  int32_t num_removed_landmarks;          /* outlier rejection */
  double initial_cost;
  double final_cost;
  double fixed_cost;                      /* cost of residuals with no variable block */
  double lm_loop_seconds;                 /* wall time of the LM loop (no setup) */
  double evaluate_seconds;                /* residual+Jacobian+normal-equation passes */
  double linear_solver_seconds;
  int64_t total_linear_solver_iterations;
  int32_t iterations_capacity;            /* caller-set capacity of `iterations` */
  int32_t reserved0;
  mlba_iteration_summary* iterations;     /* caller-owned, nullable */
  /* Device time inside the LM loop, from HIP events on the library's stream:
   * summed milliseconds and launch count per MLBA_KERNEL_* id. */
  double kernel_ms[MLBA_NUM_KERNELS];
  int32_t kernel_launches[MLBA_NUM_KERNELS];
} mlba_summary;

/* Kernel ids of mlba_summary.kernel_ms. */
#define MLBA_KERNEL_LINEARIZE 0       /* residual + Jacobian + landmark reductions */
#define MLBA_KERNEL_COST 1            /* residual-only cost of the candidate */
#define MLBA_KERNEL_SCHUR_BLOCKS 2    /* reduced camera system blocks and rhs */
#define MLBA_KERNEL_PCG 3             /* whole PCG solve (all its launches) */
#define MLBA_KERNEL_BACK_SUBSTITUTE 4 /* landmark step + model cost change */
#define MLBA_KERNEL_POSE_ROWS 5       /* vertex gradient and diag(J'J) */
#define MLBA_KERNEL_LANDMARK_PREP 6   /* damped landmark pivots */
#define MLBA_KERNEL_SCHUR_INCIDENCES 7 /* per-iteration Schur factors of every row incidence */
#define MLBA_KERNEL_SCHUR_FACTORS 8   /* per-linearization Schur factors of every row incidence */

typedef struct mlba_handle mlba_handle;

int32_t mlba_abi_version(void);

/* Fill maplab's flag defaults (solver-options.cc:3-52). */
void mlba_options_init(mlba_options* options);

/* Copy the map to the device and build the observation / Schur topology.
 * Replaces constructOptimizationProblem + buildCeresProblemFromProblemInformation
 * (vi-optimization-builder.cc:154-374, problem-information.cc:22-97). */
int32_t mlba_create(const mlba_problem_desc* desc, mlba_handle** out_handle);
int32_t mlba_destroy(mlba_handle* handle);

/* Last error text for `handle` (or for the last failed mlba_create when NULL). */
const char* mlba_last_error(const mlba_handle* handle);

/* Replace / read back the optimized state (copyAllStatesBackToMap,
 * optimization-state-buffer.cc:42-151). Either pointer may be NULL. An
 * upload also clears what the linear solver carries from solve to solve
 * (recycled deflation vectors), so a solve from an uploaded state repeats
 * bit for bit. */
int32_t mlba_upload_state(mlba_handle* handle, const double* vertex_q_IM__M_p_MI,
                          const double* landmark_p_B);
int32_t mlba_download_state(const mlba_handle* handle, double* vertex_q_IM__M_p_MI,
                            double* landmark_p_B);

/* Deactivate / reactivate landmarks (ProblemInformation::deactivateCostFunction
 * for all observations of a landmark, outlier-rejection-solver.cc:168-177).
 * mask[l] != 0 means active. */
int32_t mlba_set_landmark_active(mlba_handle* handle, const uint8_t* mask);
int32_t mlba_get_landmark_active(const mlba_handle* handle, uint8_t* mask);

/*
 * Batched VisualReprojectionError::Evaluate over every observation, in the
 * caller's observation order. residuals: 2N (may be NULL). jacobians: array
 * of 9 nullable pointers in the reference block order
 * {p_B, anchor pose, landmark baseframe, observer baseframe, observer pose,
 *  q_CI, p_CI, intrinsics, distortion}; block k of observation i is written
 * row-major at jacobians[k] + i * 2 * size_k, with size_k = {3,7,7,7,7,4,3,4,4}
 * (ambient) or {3,6,6,6,6,3,3,4,4} (tangent). Blocks the error type does not
 * use are written as zeros. Failed projections give zero residual and
 * Jacobians (visual-error-term-inl.h:155-178). cost (nullable) receives
 * 0.5 * sum rho(|r|^2) over active observations with the options' loss.
 */
int32_t mlba_evaluate(mlba_handle* handle, int32_t flags, const mlba_options* options,
                      double* residuals, double* const* jacobians, double* cost);

/*
 * mlba_evaluate without any allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory during a call (tight serving
 * loops, harnesses that forbid hipMalloc). Same arguments, outputs and
 * results as mlba_evaluate (bit for bit). The residuals and Jacobian records
 * are staged in one fixed array of MLBA_NOALLOC_SCRATCH_BYTES per device,
 * held from the first call until the process ends; a call that would need
 * more (2N doubles for residuals plus, when any Jacobian is requested, N
 * records of 2 x (sum of the block sizes incl. the xi column) doubles: 84
 * tangent, 94 ambient) fails with MLBA_EINVAL and its reason in
 * mlba_last_error. Calls on one device share the array: the library
 * serializes them with a per-device lock held until the call's stream work
 * has finished, so overlapping calls from several threads wait for each
 * other. The handle's device is made current for the call and restored.
 */
#define MLBA_NOALLOC_SCRATCH_BYTES (64u << 20)
int32_t mlba_evaluate_noalloc(mlba_handle* handle, int32_t flags, const mlba_options* options,
                              double* residuals, double* const* jacobians, double* cost);

/* ceres::Solve on the visual problem (map_optimization::solve, solver.cc:14-44). */
int32_t mlba_solve(mlba_handle* handle, const mlba_options* options, mlba_summary* summary);

/*
 * Camera calibration refinement with the map held constant: the blocks the
 * visual error term takes for each camera (q_CI, C_p_CI, intrinsics with xi,
 * distortion; visual-error-term-inl.h:321-367) are optimized with the Ceres
 * LM rules of mlba_solve, on dense normal equations of at most 15 columns
 * per camera, while every vertex and landmark in the problem is constant
 * (MLBA_EINVAL otherwise; refining calibration jointly with the map is not
 * supported). The fix flags mirror --ba_fix_ncamera_intrinsics (here split
 * into intrinsics and distortion), --ba_fix_ncamera_extrinsics_rotation and
 * --ba_fix_ncamera_extrinsics_translation (vi-optimization-builder.cc:24-33,
 * optimization-terms-addition.cc:222-234). The refined cameras are used by
 * every later call on the handle and read back with mlba_download_cameras.
 */
typedef struct mlba_calibration_options {
  int32_t fix_intrinsics;             /* fu fv cu cv (and xi of a unified-projection camera) */
  int32_t fix_distortion;
  int32_t fix_extrinsics_rotation;    /* q_CI */
  int32_t fix_extrinsics_translation; /* C_p_CI */
} mlba_calibration_options;

void mlba_calibration_options_init(mlba_calibration_options* options); /* all fixed (maplab's defaults) */
int32_t mlba_solve_calibration(mlba_handle* handle, const mlba_options* options,
                               const mlba_calibration_options* calibration, mlba_summary* summary);
/* Current camera blocks: 7 (q_CI | C_p_CI), 4 intrinsics, 4 distortion and
 * xi per camera; every pointer nullable. */
int32_t mlba_download_cameras(const mlba_handle* handle, double* q_CI__C_p_CI, double* intrinsics,
                              double* distortion, double* xi);

/* Several independent problems solved concurrently: handle i runs
 * mlba_solve (or, with outlier_rejection, mlba_solve_with_outlier_rejection)
 * with options[i] and summaries[i] (nullable), each on its own stream; up to
 * `threads` host threads drive the handles (0: min(n, 8)), so the device
 * overlaps one problem's host decisions and small launches with another's
 * kernels. statuses[i] (nullable array) receives each status; the return
 * value is the first non-OK status in handle order (every handle is still
 * solved). The handles must be distinct. No counterpart in the reference
 * (maplab solves one map per VIMapOptimizer::optimize call); for callers
 * that optimize many maps, e.g. a batch of missions or sessions. */
int32_t mlba_solve_many(mlba_handle* const* handles, const mlba_options* const* options,
                        mlba_summary* const* summaries, int32_t* statuses, int32_t n, int32_t threads,
                        int32_t outlier_rejection);

/* map_optimization::solveWithOutlierRejection (outlier-rejection-solver.cc:206-259). */
int32_t mlba_solve_with_outlier_rejection(mlba_handle* handle, const mlba_options* options,
                                          mlba_summary* summary);

/* Landmark admission thresholds: maplab's elq_* flags
 * (map-structure/vi-map/src/landmark-quality-metrics.cc:11-33). */
typedef struct mlba_quality_options {
  int32_t min_observers;                     /* elq_min_observers (4) */
  int32_t reserved0;
  double min_observation_angle_deg;          /* elq_min_observation_angle_deg (5) */
  double max_distance_from_closest_observer; /* elq_max_distance_from_closest_observer (40 m) */
  double min_distance_from_closest_observer; /* elq_min_distance_from_closest_observer (0.05 m) */
  double max_reprojection_error_px;          /* elq_max_reprojection_error_px (-1: check off) */
} mlba_quality_options;

void mlba_quality_options_init(mlba_quality_options* options);

/* vi_map::isVisualLandmarkWellConstrained (landmark-quality-metrics.cc:103-171,
 * with common::getMaxDisparityRadAngleOfUnitVectorBundle, geometry.cc:30-44)
 * of every landmark at the current device state, over all of its
 * observations in the problem; good[l] = 1 when the landmark is well
 * constrained (the admission test of addLandmarkTermsForVertices,
 * optimization-terms-addition.cc:304-312). Evaluated on the device; the
 * active mask is not changed (pass the result to mlba_set_landmark_active). */
int32_t mlba_landmark_quality(mlba_handle* handle, const mlba_quality_options* options, uint8_t* good);

/*
 * Measurement hook (no counterpart in the reference): average device time of
 * each kernel of one LM iteration at the current state, timed with HIP
 * events on the library's stream over `repeats` launches, plus the sizes
 * that the algorithmic byte and flop counts of those kernels need.
 */
typedef struct mlba_kernel_times {
  double linearize_ms;        /* residual + Jacobian + landmark reductions */
  double cost_ms;             /* residual-only cost (candidate evaluation) */
  double pose_rows_ms;        /* vertex gradient and diag(J'J) */
  double landmark_prep_ms;    /* damped landmark pivots */
  double schur_factors_ms;    /* per-linearization Schur factors of every row incidence */
  double schur_incidences_ms; /* per-iteration Schur factors of every row incidence */
  double schur_blocks_ms;     /* reduced camera system blocks and rhs */
  double block_inverse_ms;    /* block-Jacobi preconditioner */
  double pcg_solve_ms;        /* one whole (deflated) PCG solve at the initial radius */
  double pcg_spmv_ms;         /* one PCG iteration's S p product */
  double pcg_update_ms;       /* one PCG iteration's vector update and reductions */
  double back_substitute_ms;  /* landmark step + model cost change */
  int64_t num_observations;
  int64_t num_landmarks;
  int64_t num_variable_vertices;
  int64_t schur_blocks;       /* 6x6 blocks of S */
  int64_t schur_upper_blocks; /* blocks of the upper triangle (diagonal included) */
  int64_t schur_pairs;        /* incidence pairs summed into the upper blocks */
  int64_t incidences;         /* vertex-observation and anchor incidences */
  int64_t linearize_tiles;
  int64_t pcg_iterations;     /* iterations of the timed PCG solve */
  int64_t pcg_coarse_columns; /* columns of the deflation space (6 x hat nodes) */
  int64_t pcg_window_columns; /* coarse columns a PCG row group couples to */
  /* The block kernel's work decomposition: (landmark, row) factor blocks,
   * factor blocks staged over all tile visits (each visit stages its
   * landmark's blocks in the tile's two bands), landmark tile visits,
   * LDS batches, work items (tile segments) and factor-block products. */
  int64_t schur_factor_blocks;
  int64_t schur_staged_blocks;
  int64_t schur_visits;
  int64_t schur_batches;
  int64_t schur_segments;
  int64_t schur_tile_pairs;
} mlba_kernel_times;

int32_t mlba_time_kernels(mlba_handle* handle, const mlba_options* options, int32_t repeats,
                          mlba_kernel_times* times);

/*
 * Measurement hook (no counterpart in the reference): a plain device copy of
 * `bytes` bytes (16-byte vector loads and stores, grid-stride), launched
 * `repeats` times back to back on its own stream and timed with HIP events
 * around the batch; ms_per_copy receives the average. The reference point of
 * the HBM bandwidth and of the rocprofv3 FETCH_SIZE / WRITE_SIZE counters
 * (a copy moves 2 x bytes). Allocates and frees its two buffers.
 */
int32_t mlba_measure_copy(int64_t bytes, int32_t repeats, double* ms_per_copy);

/*
 * Launch-schedule autotuner (no counterpart in the reference: Ceres has no
 * device schedules). The reduced-camera-system block kernel sums fixed
 * work items (segments of 16 x 16-block tiles of S, each summing its
 * landmarks' factor-block products in a fixed order) into fixed partial
 * slots; what can be chosen is only the order the items are dispatched in
 * (bands of item_band_rows tile rows swept tile column by tile column; 0 =
 * the items with the most work first), how many items a work-group runs
 * (item_waves: 1, 2, 4, 8) and whether work-groups are mapped
 * XCD-contiguously. The PCG sums the per-group
 * partials of each iteration (r'z, r'r and the coarse-space product of every
 * 42-row group) in group order, either in every SpMV work-group
 * (pcg_partials 1) or once, in the update work-group that finishes last
 * (2); 0 picks 2 when the partials exceed one 16 KB LDS chunk. None of these
 * changes a result bit: a solve after mlba_autotune or mlba_set_schedule
 * repeats the same iteration log and state.
 *
 * mlba_autotune times every schedule (item_band_rows in {0,1,2,4,8,16,32} x
 * item_waves x mapping, `repeats` launches each) on the Schur factors of
 * the current state and keeps the fastest for this handle (pcg_partials
 * is kept as set); it leaves the state unchanged. The schedule survives topology rebuilds (outlier
 * rejection, landmark masks).
 */
typedef struct mlba_schedule {
  int32_t item_band_rows;     /* 0 or 1..4096 */
  int32_t item_waves;         /* 1, 2, 4 or 8 */
  int32_t xcd_mapping;        /* 0 or 1 */
  int32_t schedules_tried;    /* mlba_autotune: schedules timed */
  double default_ms;          /* mlba_autotune: block kernel with the library default (0, 1, 0) */
  double tuned_ms;            /* mlba_autotune: block kernel with the chosen schedule */
  int32_t pcg_partials;       /* 0 automatic, 1 per SpMV work-group, 2 once per PCG update */
  int32_t reserved0;
} mlba_schedule;

int32_t mlba_autotune(mlba_handle* handle, const mlba_options* options, int32_t repeats,
                      mlba_schedule* chosen);
int32_t mlba_set_schedule(mlba_handle* handle, const mlba_schedule* schedule);
int32_t mlba_get_schedule(const mlba_handle* handle, mlba_schedule* schedule);

#ifdef __cplusplus
}
#endif
#endif /* MAPLAB_BA_H_ */
