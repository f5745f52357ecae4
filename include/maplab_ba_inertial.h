/*
 * maplab_ba_inertial.h -- the inertial error term of optvi on the device:
 * ceres_error_terms::InertialErrorTerm::Evaluate
 * (algorithms/ceres-error-terms/src/inertial-error-term.cc:73-286) with the
 * RK4 IMU integrator imu_integrator::ImuIntegratorRK4
 * (algorithms/imu-integrator-rk4/include/imu-integrator/imu-integrator-inl.h),
 * evaluated for many edges (consecutive vertices with the IMU samples between
 * them, as addInertialTerms builds them, optimization-terms-addition.cc:361-389).
 */
#ifndef MAPLAB_BA_INERTIAL_H_
#define MAPLAB_BA_INERTIAL_H_

#include <stdint.h>

#include "maplab_ba.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mlba_inertial_problem {
  int32_t num_vertices;
  int32_t num_edges;
  int32_t num_samples;
  int32_t reserved0;
  const double* vertex_q_IM__M_p_MI;   /* 7 * num_vertices (JPL q_I_M | p_M_I) */
  const double* vertex_gyro_bias;      /* 3 * num_vertices */
  const double* vertex_velocity_M;     /* 3 * num_vertices */
  const double* vertex_accel_bias;     /* 3 * num_vertices */
  const int32_t* edge_from;            /* num_edges: begin vertex */
  const int32_t* edge_to;              /* num_edges: end vertex */
  const int32_t* edge_sample_begin;    /* num_edges + 1: samples of edge e are [begin[e], begin[e+1]) */
  const double* imu;                   /* 6 * num_samples: accelerometer xyz, gyroscope xyz */
  const int64_t* timestamps_ns;        /* num_samples, non-decreasing within an edge */
  double gyro_noise_sigma;             /* ImuIntegratorRK4 constructor arguments */
  double gyro_bias_sigma;
  double acc_noise_sigma;
  double acc_bias_sigma;
  double gravity_magnitude;
} mlba_inertial_problem;

/* Residuals (15 per edge) and the ambient Jacobians of the 8 parameter
 * blocks {pose_from (7), b_g_from, v_from, b_a_from, pose_to (7), b_g_to,
 * v_to, b_a_to} (row-major 15 x size per edge; jacobians and each entry
 * nullable). Returns MLBA_EINVAL for the reference's CHECK conditions
 * (sigmas <= 0, an edge without samples, unordered timestamps) and
 * MLBA_ESTATE when an accumulated covariance is not positive definite. */
int32_t mlba_inertial_evaluate(const mlba_inertial_problem* problem, double* residuals, double* const* jacobians);

const char* mlba_inertial_last_error(void);

/* optvi's speed and bias blocks with the keyframe poses held constant (as
 * with --ba_fix_vertices): the velocity and IMU biases of every vertex of
 * the edges are optimized over the inertial terms alone, with the Ceres LM
 * rules of mlba_solve on the block-tridiagonal normal equations of a
 * keyframe chain (every edge must join vertex i to vertex i + 1). The fix
 * flags are --ba_fix_gyro_bias, --ba_fix_accel_bias, --ba_fix_velocity
 * (vi-optimization-builder.cc:42-50, optimization-terms-addition.cc:
 * 468-485). The summary's costs are those of the inertial terms; the
 * optimized states go to the nullable outputs (3 P each). Refining poses
 * with the inertial terms (the full optvi problem) is not supported. */
typedef struct mlba_inertial_options {
  int32_t fix_gyro_bias;
  int32_t fix_accel_bias;
  int32_t fix_velocity;
  int32_t reserved0;
} mlba_inertial_options;

void mlba_inertial_options_init(mlba_inertial_options* options);  /* nothing fixed (maplab's defaults) */
int32_t mlba_solve_inertial_states(const mlba_inertial_problem* problem, const mlba_options* options,
                                   const mlba_inertial_options* inertial, double* gyro_bias_out,
                                   double* velocity_out, double* accel_bias_out, mlba_summary* summary);

#ifdef __cplusplus
}
#endif
#endif /* MAPLAB_BA_INERTIAL_H_ */
