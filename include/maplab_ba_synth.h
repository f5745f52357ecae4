/*
 * maplab_ba_synth.h -- seeded synthetic visual maps (libmaplab_ba_synth.so).
 *
 * Plays the role of maplab's synthetic map generator
 * (test/vi-map-generator-6dof/src/6dof-pose-graph-gen.cc:164-415) and of the
 * seeded corruption of test/vi-mapping-test-app/src/vi-mapping-test-app.cc:78-153,
 * for the benchmark workloads: a circle of inward-facing keyframes around a
 * landmark box, and long random-walk trajectories. The random stream is
 * xoshiro256** with Box-Muller normals, so one seed gives the same bytes on
 * every host. Host-only code; no device calls.
 */
#ifndef MAPLAB_BA_SYNTH_H_
#define MAPLAB_BA_SYNTH_H_

#include <stdint.h>

#include "maplab_ba.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MLBA_SYNTH_CIRCLE 0       /* keyframes on a circle facing its centre */
#define MLBA_SYNTH_RANDOM_WALK 1  /* forward-looking random walk, landmarks in a shell */

#define MLBA_ANCHOR_FIRST_OBSERVER 0 /* landmark stored in its first observer (maplab store vertex) */
#define MLBA_ANCHOR_GLOBAL 1         /* landmarks stored in a constant identity vertex of another mission */

typedef struct mlba_synth_config {
  int32_t kind;
  int32_t num_keyframes;
  int32_t num_landmarks;
  int32_t distortion_type;
  double distortion[4];
  double intrinsics[4];
  int32_t width;
  int32_t height;
  double p_C_I[3];               /* camera-from-IMU translation; rotation is identity */
  double keypoint_sigma_px;      /* keypoint noise std and the stored sigma */
  double outlier_fraction;       /* keypoints replaced by uniform image points */
  double rot_noise_rad;          /* initial-value corruption of vertex rotations */
  double pos_noise_m;            /* ... of vertex positions */
  double landmark_noise_m;       /* ... of landmark positions */
  double step_m;                 /* random walk step */
  double yaw_drift_rad;          /* random walk yaw change std per step */
  double min_dist_m;             /* landmark distance to trajectory, shell */
  double max_dist_m;
  double max_range_m;            /* visibility range */
  double min_depth_m;            /* observation kept when p_C.z >= this */
  double box_half_m;             /* circle: landmark box half size */
  double circle_radius_m;        /* circle: keyframe circle radius */
  int32_t min_obs_per_landmark;  /* landmarks with fewer observations are redrawn */
  int32_t max_obs_per_landmark;  /* observations beyond this are subsampled */
  double detection_probability;  /* a visible landmark is detected with this probability */
  int32_t anchor_mode;           /* MLBA_ANCHOR_* */
  int32_t fix_scale_landmark;    /* fix one landmark of the first vertex (visual-only gauge) */
  uint64_t seed;
} mlba_synth_config;

typedef struct mlba_synth_map mlba_synth_map;

/* Fill the configuration of a benchmark workload:
 * 1: 10 keyframes on a 1 m circle, 2000 landmarks in a 4 m box, radtan;
 * 2: 500-keyframe random walk, 200k landmarks, radtan;
 * 3: 2000-keyframe random walk, 1M landmarks, radtan.
 * Returns MLBA_EINVAL for an unknown id. */
int32_t mlba_synth_config_init(mlba_synth_config* config, int32_t workload);

int32_t mlba_synth_generate(const mlba_synth_config* config, mlba_synth_map** out_map);
void mlba_synth_free(mlba_synth_map* map);

/* Problem description pointing into the map's arrays (valid until free). The
 * state is the corrupted initial value; the first vertex (and, when asked,
 * one landmark stored in it) is constant. */
int32_t mlba_synth_describe(const mlba_synth_map* map, mlba_problem_desc* desc);

/* Ground-truth state (7 * num_vertices, 3 * num_landmarks). */
const double* mlba_synth_true_vertices(const mlba_synth_map* map);
const double* mlba_synth_true_landmarks(const mlba_synth_map* map);

#ifdef __cplusplus
}
#endif
#endif /* MAPLAB_BA_SYNTH_H_ */
