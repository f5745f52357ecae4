/*
 * maplab_ba_optimizer.h -- the visual part of map_optimization::VIMapOptimizer
 * (algorithms/map-optimization/src/vi-map-optimizer.cc:34-80) over the device
 * BA library: problem construction (landmark admission, mission selection,
 * frame filter, gauge fixes; vi-optimization-builder.cc:154-374), the solve
 * (plain or with outlier rejection) and the copy back of the state.
 *
 * The map is the structure-of-arrays mlba_problem_desc of maplab_ba.h.
 * Vertices are taken in pose-graph order within each mission (the first
 * vertex of a mission is its lowest index) and a vertex's landmark store in
 * landmark index order.
 */
#ifndef MAPLAB_BA_OPTIMIZER_H_
#define MAPLAB_BA_OPTIMIZER_H_

#include <stdint.h>

#include "maplab_ba.h"

#ifdef __cplusplus
extern "C" {
#endif

/* vi_map::Landmark::Quality (map-structure/vi-map/include/vi-map/landmark.h). */
#define MLBA_QUALITY_UNKNOWN 0
#define MLBA_QUALITY_BAD 1
#define MLBA_QUALITY_GOOD 2
#define MLBA_QUALITY_LOCALIZATION_SUMMARY 3

/* ViProblemOptions, visual terms (vi-optimization-builder.cc:98-150). */
typedef struct mlba_vi_problem_options {
  int32_t add_visual_constraints;          /* --ba_include_visual (1) */
  int32_t fix_landmark_positions;          /* --ba_fix_landmark_positions (0) */
  int32_t fix_vertices;                    /* --ba_fix_vertices (0) */
  int32_t min_landmarks_per_frame;         /* --ba_min_landmarks_per_frame (0) */
  int32_t enable_visual_outlier_rejection; /* --ba_use_visual_outlier_rejection_solver (1) */
  int32_t reserved0;
  mlba_options solver;                     /* initSolverOptionsFromFlags + outlier rejection flags */
  mlba_quality_options landmark_quality;   /* --elq_* (landmark admission) */
  mlba_calibration_options calibration;    /* --ba_fix_ncamera_intrinsics (intrinsics and distortion),
                                              --ba_fix_ncamera_extrinsics_rotation / _translation
                                              (vi-optimization-builder.cc:24-33): all fixed by default */
} mlba_vi_problem_options;

void mlba_vi_problem_options_init(mlba_vi_problem_options* options);

/*
 * The visual problem constructOptimizationProblem builds from a map, given the
 * admission decision of every landmark (landmark_good, L): the observations
 * that enter (addLandmarkTermsForVertices, optimization-terms-addition.cc:
 * 243-331: observer and store missions selected, landmark well constrained,
 * frames with at least min_landmarks_per_frame well-constrained landmarks),
 * the vertices and landmarks in the problem, and the constant blocks: the
 * caller's constant flags, --ba_fix_vertices / --ba_fix_landmark_positions,
 * and per mission co-observation cluster (mission-clustering-coobservation.cc:
 * 84-104) the visual-only gauge of applyGaugeFixesForInitialVertices
 * (optimization-problem.cc:96-176): the pose of the cluster's first vertex in
 * the problem and the first landmark stored in that vertex. Host only.
 * Output arrays are caller-allocated.
 */
typedef struct mlba_problem_selection {
  int32_t num_observations;        /* observations in the problem */
  int32_t num_vertices;            /* keyframes_in_problem */
  int32_t num_landmarks;           /* landmarks_in_problem */
  int32_t num_clusters;            /* mission co-observation clusters */
  uint8_t* obs_in_problem;         /* num_observations of the map */
  uint8_t* vertex_in_problem;      /* num_vertices of the map */
  uint8_t* landmark_in_problem;    /* num_landmarks of the map */
  uint8_t* vertex_constant;        /* num_vertices of the map */
  uint8_t* landmark_constant;      /* num_landmarks of the map */
  int32_t* gauge_vertex;           /* per cluster (capacity num_missions), nullable */
  int32_t* gauge_landmark;         /* per cluster (capacity num_missions), nullable */
} mlba_problem_selection;

int32_t mlba_select_problem(const mlba_problem_desc* map, const mlba_vi_problem_options* options,
                            const uint8_t* mission_selected, const uint8_t* landmark_good,
                            mlba_problem_selection* selection);

/*
 * VIMapOptimizer::optimize for the visual terms. mission_selected (M,
 * nullable: all missions) is missions_to_optimize. landmark_quality (L,
 * nullable: all unknown) is each landmark's stored quality: unknown ones are
 * evaluated with isLandmarkWellConstrained on the device; on return,
 * landmarks rejected by the outlier-rejection solver are set to
 * MLBA_QUALITY_BAD (rejectOutliers, outlier-rejection-solver.cc:142-182).
 * vertices_out (7P) and landmarks_out (3L) receive the optimized state
 * (copyAllStatesBackToMap); entries not in the problem are copied unchanged.
 * summary (nullable) receives the Ceres-style summary of the solve.
 */
int32_t mlba_vi_map_optimize(const mlba_problem_desc* map, const mlba_vi_problem_options* options,
                             const uint8_t* mission_selected, uint8_t* landmark_quality, double* vertices_out,
                             double* landmarks_out, mlba_summary* summary);

/* The same, with the camera blocks: when options->calibration frees any of
 * them, the camera calibration is refined with mlba_solve_calibration
 * instead of the visual BA (every vertex and landmark in the selected
 * problem must then be constant, --ba_fix_vertices and
 * --ba_fix_landmark_positions; MLBA_EINVAL otherwise; the outlier-rejection
 * chunks do not apply). The camera outputs (7C q_CI | C_p_CI, 4C
 * intrinsics, 4C distortion, C xi; each nullable) receive the cameras after
 * the solve; vertices_out / landmarks_out receive the state as for
 * mlba_vi_map_optimize. No counterpart entry point in the reference: maplab
 * writes the refined NCamera back into the map's sensor manager. */
int32_t mlba_vi_map_optimize_cameras(const mlba_problem_desc* map, const mlba_vi_problem_options* options,
                                     const uint8_t* mission_selected, uint8_t* landmark_quality,
                                     double* vertices_out, double* landmarks_out, double* camera_q_CI__C_p_CI_out,
                                     double* camera_intrinsics_out, double* camera_distortion_out,
                                     double* camera_xi_out, mlba_summary* summary);

/* Last error text of the optimizer entry points of the calling thread. */
const char* mlba_optimizer_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* MAPLAB_BA_OPTIMIZER_H_ */
