/*
 * ba_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * FP64 CPU restatement of maplab's visual bundle-adjustment path, used as the
 * parity oracle for the HIP library and as the timed CPU baseline. Only the
 * tests, the smoke check and bench.py's cpu_baseline leg load it; nothing in
 * maplab_amd/ links, loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the maplab repository root). Parts that live in Ceres Solver (absent
 * from the reference tree; API bounds it to Ceres 1.13-2.1) are restated
 * from Ceres' published algorithm and marked "ceres:" in the comments.
 */
#ifndef MAPLAB_BA_ORACLE_H_
#define MAPLAB_BA_ORACLE_H_

#include <stdint.h>

#include "../include/maplab_ba.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Linear solvers of the oracle LM. */
#define ORACLE_SCHUR_DIRECT 0  /* landmark elimination + dense Cholesky of S */
#define ORACLE_DENSE_NORMAL 1  /* dense Cholesky of J'J + D'D (small problems) */
#define ORACLE_SCHUR_PCG 2     /* landmark elimination + block-Jacobi PCG on S */

/* PinholeCamera::project3Functional. Returns the ProjectionResult status:
 * 0 visible, 1 outside image box, 2 behind camera, 3 invalid.
 * jp (2x3), jint (2x4), jdist (2x4) are row-major and nullable. */
int32_t oracle_project(int32_t dist_type, const double* intrinsics, const double* dist,
                       int32_t width, int32_t height, const double* p_C, double* kp,
                       double* jp, double* jint, double* jdist);

/* The same for a camera model (MLBA_CAMERA_*; xi of a unified-projection
 * camera): UnifiedProjectionCamera::project3Functional
 * (camera-unified-projection.cc:74-196). jxi (2x1, nullable) = d kp / d xi. */
int32_t oracle_project_camera(int32_t model, double xi, int32_t dist_type, const double* intrinsics,
                              const double* dist, int32_t width, int32_t height, const double* p_C, double* kp,
                              double* jp, double* jint, double* jdist, double* jxi);

/* Distortion::distortUsingExternalCoefficients: in/out point, jd (2x2) nullable. */
void oracle_distort(int32_t dist_type, const double* dist, double* point, double* jd);

/* JplQuaternionParameterization::Plus and ::ComputeJacobian (4x3 row-major). */
void oracle_quat_plus(const double* q, const double* delta, double* out);
void oracle_quat_jacobian(const double* q, double* j43);
void oracle_rotation_jpl(const double* q, double* r33);

/* VisualReprojectionError::Evaluate on explicit parameter blocks.
 * params: 9 block pointers in the reference order; jac: 9 nullable pointers
 * (ambient, row-major). Returns 1 when the projection failed (r = J = 0). */
int32_t oracle_evaluate_visual(const double* meas, double sigma, int32_t error_type,
                               int32_t dist_type, int32_t width, int32_t height,
                               const double* const* params, double* residual,
                               double* const* jac);

/* ceres::LossFunction::Evaluate for the supported losses: rho[3]. */
void oracle_loss(int32_t loss_type, double a, double s, double* rho);

/* Batched evaluate over a problem, same contract as mlba_evaluate (flags
 * MLBA_EVAL_TANGENT and MLBA_EVAL_WITH_XI). */
int32_t oracle_problem_evaluate(const mlba_problem_desc* desc, const uint8_t* landmark_active,
                                int32_t flags, const mlba_options* options,
                                double* residuals, double* const* jacobians, double* cost);

/* One Levenberg-Marquardt solve on a copy of the problem state; the final
 * state is written to vertices_out (7P) and landmarks_out (3L) (nullable).
 * num_threads <= 0 selects max(1, hw/2) (solver-options.h:37-38). */
int32_t oracle_solve(const mlba_problem_desc* desc, const mlba_options* options,
                     int32_t linear_solver, int32_t num_threads,
                     const uint8_t* landmark_active, double* vertices_out,
                     double* landmarks_out, mlba_summary* summary);

/* Camera calibration refinement with every vertex and landmark constant:
 * fix = {intrinsics (with xi), distortion, extrinsics rotation, extrinsics
 * translation}; refined cameras into the nullable outputs (7C, 4C, 4C, C).
 * MLBA_EINVAL when a map block is variable or nothing is free. */
int32_t oracle_solve_calibration(const mlba_problem_desc* desc, const mlba_options* options, const int32_t* fix,
                                 const uint8_t* landmark_active, double* cam_q_out, double* intrinsics_out,
                                 double* distortion_out, double* xi_out, mlba_summary* summary);

/* solveWithOutlierRejection; landmark_active_out (L, nullable) receives the
 * final active mask. */
int32_t oracle_solve_with_outlier_rejection(const mlba_problem_desc* desc,
                                            const mlba_options* options,
                                            int32_t linear_solver, int32_t num_threads,
                                            double* vertices_out, double* landmarks_out,
                                            uint8_t* landmark_active_out,
                                            mlba_summary* summary);

/* vi_map::isVisualLandmarkWellConstrained (landmark-quality-metrics.cc:103-171)
 * of every landmark over all of its observations in desc; good (L) receives
 * 1 for well-constrained landmarks. */
int32_t oracle_landmark_quality(const mlba_problem_desc* desc, const mlba_quality_options* options,
                                uint8_t* good);

/* InertialErrorTerm::Evaluate (inertial-error-term.cc:73-286) for one edge:
 * n IMU samples imu (n x 6: accelerometer, gyroscope) at t_ns (n), noise =
 * {gyro noise, gyro bias, acc noise, acc bias sigmas, gravity}, params =
 * {pose_from (7), b_g_from, v_from, b_a_from, pose_to (7), b_g_to, v_to,
 * b_a_to}. residual (15, nullable), jacobians[8] (nullable each, row-major
 * 15 x {7,3,3,3,7,3,3,3}, ambient as the reference), and optionally the
 * accumulated transition (15x15), covariance (15x15) and end state (16). */
int32_t oracle_inertial_evaluate(int32_t n, const double* imu, const int64_t* t_ns, const double* noise,
                                 const double* const* params, double* residual, double* const* jacobians,
                                 double* phi_out, double* Q_out, double* end_state_out);

/* Speed and bias states of the edges' vertices with the poses constant:
 * dense LM over the free [b_g | v_M | b_a] columns; fix = {gyro bias, accel
 * bias, velocity}; noise as oracle_inertial_evaluate. Outputs nullable. */
int32_t oracle_solve_inertial_states(int32_t P, int32_t E, const double* vertices, const double* gyro_bias,
                                     const double* velocity, const double* accel_bias, const int32_t* edge_from,
                                     const int32_t* edge_to, const int32_t* sample_begin, const double* imu,
                                     const int64_t* t_ns, const double* noise, const int32_t* fix,
                                     const mlba_options* options, double* gyro_bias_out, double* velocity_out,
                                     double* accel_bias_out, mlba_summary* summary);

/* The damped, Jacobi-scaled reduced camera system at the given state:
 * S (n x n dense, n = 6 * number of variable vertices that have residuals,
 * ordered by vertex index) and rhs (n), for LM radius `mu`. Used to check the
 * device Schur build. Returns n, or a negative error. */
int32_t oracle_reduced_camera_system(const mlba_problem_desc* desc, const mlba_options* options,
                                     double mu, int32_t capacity, double* S, double* rhs,
                                     int32_t* vertex_of_block);

#ifdef __cplusplus
}
#endif
#endif /* MAPLAB_BA_ORACLE_H_ */
