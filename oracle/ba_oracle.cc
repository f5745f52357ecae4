// ba_oracle.cc -- TEST INFRASTRUCTURE ONLY.
//
// FP64 CPU restatement of maplab's visual bundle-adjustment hot path. It is
// the parity oracle for the HIP library and the CPU baseline of bench.py.
// Nothing in the product links or calls it.
//
// Reference parts restated (paths relative to the maplab repository root):
//   VisualReprojectionError::Evaluate
//       algorithms/ceres-error-terms/include/ceres-error-terms/visual-error-term-inl.h:13-379
//   PinholeCamera::project3Functional / evaluateProjectionResult / isKeypointVisible
//       aslam_cv2/aslam_cv_cameras/src/camera-pinhole.cc:65-164,
//       include/aslam/cameras/camera-pinhole-inl.h:48-65, camera-inl.h:17-25
//   RadTanDistortion  aslam_cv2/aslam_cv_cameras/src/distortion-radtan.cc:14-91
//   EquidistantDistortion aslam_cv2/aslam_cv_cameras/src/distortion-equidistant.cc:14-142
//   FisheyeDistortion aslam_cv2/aslam_cv_cameras/src/distortion-fisheye.cc:14-120
//   JPL quaternion math common/maplab-common/include/maplab-common/quaternion-math-inl.h:7-27,113-154
//   JplQuaternionParameterization
//       algorithms/ceres-error-terms/src/parameterization/quaternion-param-jpl.cc:8-76
//   skew  common/maplab-common/include/maplab-common/geometry-inl.h:12-21
//   Huber wrapping  algorithms/map-optimization/src/optimization-terms-addition.cc:189-198
//   error-type choice / dummies  optimization-terms-addition.cc:68-78,
//       algorithms/ceres-error-terms/include/ceres-error-terms/landmark-common-inl.h:9-51
//   outlier rejection  algorithms/map-optimization/src/outlier-rejection-solver.cc:28-259
// Ceres Solver (absent from the tree; bounded to 1.13-2.1 by API use) parts,
// restated from its published algorithm and marked "ceres:":
//   HuberLoss / CauchyLoss, the robust-loss Corrector, Jacobi column scaling,
//   LevenbergMarquardtStrategy, TrustRegionMinimizer, Schur elimination.

#include "ba_oracle.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------- helpers

template <typename F>
void parallel_for(int n, int num_threads, F&& f) {
  if (num_threads <= 1 || n < 64) {
    f(0, n, 0);
    return;
  }
  std::vector<std::thread> threads;
  const int chunk = (n + num_threads - 1) / num_threads;
  for (int t = 0; t < num_threads; ++t) {
    const int b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    threads.emplace_back([&f, b, e, t] { f(b, e, t); });
  }
  for (auto& th : threads) th.join();
}

inline void matmul(const double* a, const double* b, double* c, int n, int k, int m) {
  // c (n x m) = a (n x k) * b (k x m), row-major.
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < m; ++j) {
      double s = 0.0;
      for (int t = 0; t < k; ++t) s += a[i * k + t] * b[t * m + j];
      c[i * m + j] = s;
    }
}
inline void transpose33(const double* a, double* at) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) at[j * 3 + i] = a[i * 3 + j];
}
// geometry-inl.h:12-21
inline void skew(const double* v, double* s) {
  s[0] = 0.0;   s[1] = -v[2]; s[2] = v[1];
  s[3] = v[2];  s[4] = 0.0;   s[5] = -v[0];
  s[6] = -v[1]; s[7] = v[0];  s[8] = 0.0;
}

// quaternion-math-inl.h:7-27 (toRotationMatrixJPL), coefficients [x y z w].
void rotJPL(const double* q, double* R) {
  const double x = q[0], y = q[1], z = q[2], w = q[3];
  R[0] = 1.0 - 2.0 * (y * y + z * z);
  R[1] = 2.0 * (x * y + z * w);
  R[2] = 2.0 * (x * z - y * w);
  R[3] = 2.0 * (x * y - z * w);
  R[4] = 1.0 - 2.0 * (x * x + z * z);
  R[5] = 2.0 * (y * z + x * w);
  R[6] = 2.0 * (x * z + y * w);
  R[7] = 2.0 * (y * z - x * w);
  R[8] = 1.0 - 2.0 * (x * x + y * y);
}

// quaternion-param-jpl.cc:8-32 (get_dQuaternionJpl_dTheta), 4x3 row-major.
void quatJacobian(const double* q, double* J) {
  J[0] = q[3];  J[1] = -q[2]; J[2] = q[1];
  J[3] = q[2];  J[4] = q[3];  J[5] = -q[0];
  J[6] = -q[1]; J[7] = q[0];  J[8] = q[3];
  J[9] = -q[0]; J[10] = -q[1]; J[11] = -q[2];
  for (int i = 0; i < 12; ++i) J[i] *= 0.5;
}

// quaternion-param-jpl.cc:35-68 (Plus) with quaternion-math-inl.h:113-154
// (signed / positive JPL product).
void quatPlus(const double* q, const double* d, double* out) {
  const double sq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  const double kThr = (0.262 * 2) * (0.262 * 2);
  double qd[4];
  if (sq < kThr) {
    qd[0] = 0.5 * d[0]; qd[1] = 0.5 * d[1]; qd[2] = 0.5 * d[2]; qd[3] = 1.0;
    const double n = std::sqrt(qd[0] * qd[0] + qd[1] * qd[1] + qd[2] * qd[2] + 1.0);
    for (int i = 0; i < 4; ++i) qd[i] /= n;
  } else {
    const double h = 0.5 * std::sqrt(sq);
    const double s = std::sin(h);
    for (int i = 0; i < 3; ++i) qd[i] = 0.5 * d[i] / h * s;
    qd[3] = std::cos(h);
  }
  // L(q1) = [[w I - [v]x, v], [-v^T, w]], product = L(qd) * q.
  const double ax = qd[0], ay = qd[1], az = qd[2], aw = qd[3];
  double p[4];
  p[0] = aw * q[0] + az * q[1] - ay * q[2] + ax * q[3];
  p[1] = -az * q[0] + aw * q[1] + ax * q[2] + ay * q[3];
  p[2] = ay * q[0] - ax * q[1] + aw * q[2] + az * q[3];
  p[3] = -ax * q[0] - ay * q[1] - az * q[2] + aw * q[3];
  const double n = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2] + p[3] * p[3]);
  for (int i = 0; i < 4; ++i) p[i] /= n;
  if (p[3] < 0.0)
    for (int i = 0; i < 4; ++i) p[i] = -p[i];
  for (int i = 0; i < 4; ++i) out[i] = p[i];
}

// ---------------------------------------------------------------- camera

// distortion-radtan.cc:14-65
void distortRadtan(const double* k, double* pt, double* J) {
  double& x = pt[0];
  double& y = pt[1];
  const double k1 = k[0], k2 = k[1], p1 = k[2], p2 = k[3];
  const double mx2 = x * x, my2 = y * y, mxy = x * y;
  const double rho2 = mx2 + my2;
  const double rad = k1 * rho2 + k2 * rho2 * rho2;
  if (J) {
    J[0] = 1.0 + rad + 2.0 * k1 * mx2 + 4.0 * k2 * rho2 * mx2 + 2.0 * p1 * y + 6.0 * p2 * x;
    J[1] = 2.0 * k1 * mxy + 4.0 * k2 * rho2 * mxy + 2.0 * p1 * x + 2.0 * p2 * y;
    J[2] = J[1];
    J[3] = 1.0 + rad + 2.0 * k1 * my2 + 4.0 * k2 * rho2 * my2 + 2.0 * p2 * x + 6.0 * p1 * y;
  }
  x += x * rad + 2.0 * p1 * mxy + p2 * (rho2 + 2.0 * mx2);
  y += y * rad + 2.0 * p2 * mxy + p1 * (rho2 + 2.0 * my2);
}
// distortion-radtan.cc:67-91
void radtanParamJac(const double* pt, double* J) {
  const double y0 = pt[0], y1 = pt[1];
  const double r2 = y0 * y0 + y1 * y1, r4 = r2 * r2;
  J[0] = y0 * r2; J[1] = y0 * r4; J[2] = 2.0 * y0 * y1; J[3] = r2 + 2.0 * y0 * y0;
  J[4] = y1 * r2; J[5] = y1 * r4; J[6] = r2 + 2.0 * y1 * y1; J[7] = 2.0 * y0 * y1;
}
// distortion-equidistant.cc:14-107. The 2x2 Jacobian is the derivative of
// (x, y) * theta_d(r) / r; the reference writes it as MATLAB output, this is
// the same derivative collected by hand: s I + s'(r)/r [x y]^T [x y].
void distortEquidistant(const double* k, double* pt, double* J) {
  double& x = pt[0];
  double& y = pt[1];
  const double x2 = x * x, y2 = y * y;
  const double r = std::sqrt(x2 + y2);
  if (r < 1e-10) {
    if (J) J[0] = J[1] = J[2] = J[3] = 0.0;
    return;
  }
  const double th = std::atan(r);
  const double t2 = th * th, t4 = t2 * t2, t6 = t2 * t4, t8 = t4 * t4;
  const double poly = 1.0 + k[0] * t2 + k[1] * t4 + k[2] * t6 + k[3] * t8;
  const double thd = th * poly;
  if (J) {
    // d(theta_d)/d(theta) = poly + theta * poly'(theta)
    const double dthd = poly + 2.0 * k[0] * t2 + 4.0 * k[1] * t4 + 6.0 * k[2] * t6 + 8.0 * k[3] * t8;
    const double r2 = x2 + y2;
    const double s = thd / r;
    const double ds_dr = dthd / (1.0 + r2) / r - thd / r2;
    const double c = ds_dr / r;
    J[0] = s + c * x2;
    J[1] = c * x * y;
    J[2] = c * x * y;
    J[3] = s + c * y2;
  }
  const double scaling = (r > 1e-8) ? thd / r : 1.0;
  x *= scaling;
  y *= scaling;
}
// distortion-equidistant.cc:109-142
void equidistantParamJac(const double* pt, double* J) {
  const double x = pt[0], y = pt[1];
  const double r = std::sqrt(x * x + y * y);
  const double th = std::atan(r);
  if (r < 1e-10) {
    for (int i = 0; i < 8; ++i) J[i] = 0.0;
    return;
  }
  const double t3 = th * th * th, t5 = t3 * th * th, t7 = t5 * th * th, t9 = t7 * th * th;
  J[0] = x * t3 / r; J[1] = x * t5 / r; J[2] = x * t7 / r; J[3] = x * t9 / r;
  J[4] = y * t3 / r; J[5] = y * t5 / r; J[6] = y * t7 / r; J[7] = y * t9 / r;
}

// distortion-fisheye.cc:14-82 (FOV model, one parameter w). The 2x2
// Jacobian is written entry by entry as the reference does.
void distortFisheye(const double* k, double* pt, double* J) {
  const double w = k[0];
  const double u = pt[0], v = pt[1];
  const double r_u = std::sqrt(u * u + v * v);
  const double r_u3 = r_u * r_u * r_u;
  const double tw = std::tan(w / 2.0);
  const double tw2 = tw * tw;
  const double at = std::atan(2.0 * tw * r_u);
  double r_rd;
  if (w * w < 1e-5) r_rd = 1.0;
  else if (r_u * r_u < 1e-5) r_rd = 2.0 * tw / w;
  else r_rd = at / (r_u * w);
  if (J) {
    if (w * w < 1e-5) {
      J[0] = 1.0; J[1] = 0.0; J[2] = 0.0; J[3] = 1.0;
    } else if (r_u * r_u < 1e-5) {
      const double f = 2.0 * tw / w;
      J[0] = f; J[1] = 0.0; J[2] = 0.0; J[3] = f;
    } else {
      const double n2 = u * u + v * v;
      const double den = w * n2 * (4.0 * tw2 * n2 + 1.0);
      J[0] = at / (w * r_u) - (u * u * at) / (w * r_u3) + (2.0 * u * u * tw) / den;
      J[1] = (2.0 * u * v * tw) / den - (u * v * at) / (w * r_u3);
      J[2] = J[1];
      J[3] = at / (w * r_u) - (v * v * at) / (w * r_u3) + (2.0 * v * v * tw) / den;
    }
  }
  pt[0] = u * r_rd;
  pt[1] = v * r_rd;
}
// distortion-fisheye.cc:84-120, one column (w) written into a 2x4 block.
// In the limit r_u^2 < 1e-5 the reference sets both rows to the constant
// (w - sin w) / (w^2 cos^2(w / 2)); that is kept.
void fisheyeParamJac(const double* k, const double* pt, double* J) {
  for (int i = 0; i < 8; ++i) J[i] = 0.0;
  const double w = k[0];
  const double tw = std::tan(w / 2.0), tw2 = tw * tw;
  const double u = pt[0], v = pt[1];
  const double r_u = std::sqrt(u * u + v * v);
  const double at = std::atan(2.0 * tw * r_u);
  if (w * w < 1e-5) return;
  if (r_u * r_u < 1e-5) {
    J[0] = J[4] = (w - std::sin(w)) / (w * w * std::cos(w / 2) * std::cos(w / 2));
    return;
  }
  J[0] = (2.0 * u * (tw2 / 2.0 + 0.5)) / (w * (4.0 * tw2 * r_u * r_u + 1.0)) - (u * at) / (w * w * r_u);
  J[4] = (2.0 * v * (tw2 / 2.0 + 0.5)) / (w * (4.0 * tw2 * r_u * r_u + 1.0)) - (v * at) / (w * w * r_u);
}

void distort(int type, const double* k, double* pt, double* J) {
  if (type == MLBA_DISTORTION_RADTAN) {
    distortRadtan(k, pt, J);
  } else if (type == MLBA_DISTORTION_EQUIDISTANT) {
    distortEquidistant(k, pt, J);
  } else if (type == MLBA_DISTORTION_FISHEYE) {
    distortFisheye(k, pt, J);
  } else if (J) {  // NullDistortion: identity
    J[0] = 1.0; J[1] = 0.0; J[2] = 0.0; J[3] = 1.0;
  }
}

enum { kVisible = 0, kOutside = 1, kBehind = 2, kInvalid = 3 };

// Camera model of one camera (pinhole, or unified projection with its xi).
struct CamModel {
  int model = MLBA_CAMERA_PINHOLE;
  double xi = 0.0;
};
CamModel camModel(const mlba_problem_desc* d, int c) {
  CamModel m;
  if (d->camera_model) m.model = d->camera_model[c];
  if (m.model == MLBA_CAMERA_UNIFIED_PROJECTION && d->camera_xi) m.xi = d->camera_xi[c];
  return m;
}

// camera-unified-projection.cc:74-196 (+ evaluateProjectionResult). The
// intrinsics Jacobian is written for (fu, fv, cu, cv) in Jint (2x4) and for
// xi in Jxi (2x1).
int projectUnified(double xi, int dist_type, const double* intr, const double* dist, int width, int height,
                   const double* p, double* kp, double* Jp, double* Jint, double* Jdist, double* Jxi) {
  const double fu = intr[0], fv = intr[1], cu = intr[2], cv = intr[3];
  const double x = p[0], y = p[1], z = p[2];
  const double d = std::sqrt(x * x + y * y + z * z);
  const double rz = 1.0 / (z + xi * d);
  const double fov_parameter = (xi <= 1.0) ? xi : (1 / xi);
  if (!(z > -(fov_parameter * d))) {
    if (Jint) for (int i = 0; i < 8; ++i) Jint[i] = 0.0;
    if (Jdist) for (int i = 0; i < 8; ++i) Jdist[i] = 0.0;
    if (Jxi) Jxi[0] = Jxi[1] = 0.0;
    if (Jp) for (int i = 0; i < 6; ++i) Jp[i] = 0.0;
    kp[0] = kp[1] = 0.0;
    return kInvalid;
  }
  kp[0] = x * rz;
  kp[1] = y * rz;
  double Jd[4] = {1.0, 0.0, 0.0, 1.0};
  if (Jdist) {
    if (dist_type == MLBA_DISTORTION_RADTAN) radtanParamJac(kp, Jdist);
    else if (dist_type == MLBA_DISTORTION_EQUIDISTANT) equidistantParamJac(kp, Jdist);
    else if (dist_type == MLBA_DISTORTION_FISHEYE) fisheyeParamJac(dist, kp, Jdist);
    else for (int i = 0; i < 8; ++i) Jdist[i] = 0.0;
    for (int i = 0; i < 4; ++i) { Jdist[i] *= fu; Jdist[4 + i] *= fv; }
  }
  distort(dist_type, dist, kp, (Jp || Jint || Jxi) ? Jd : nullptr);
  if (Jp) {
    // J (2x3, row-major) as the reference fills it, then through distortion.
    double J[6];
    double rz2 = rz * rz / d;
    J[0] = rz2 * (d * z + xi * (y * y + z * z));
    J[3] = -rz2 * xi * x * y;
    J[1] = J[3];
    J[4] = rz2 * (d * z + xi * (x * x + z * z));
    rz2 = rz2 * (-xi * z - d);
    J[2] = x * rz2;
    J[5] = y * rz2;
    for (int c = 0; c < 3; ++c) {
      Jp[c] = fu * (J[c] * Jd[0] + J[3 + c] * Jd[1]);
      Jp[3 + c] = fv * (J[c] * Jd[2] + J[3 + c] * Jd[3]);
    }
  }
  if (Jxi) {
    const double jx = -x * rz * d * rz, jy = -y * rz * d * rz;
    Jxi[0] = fu * (Jd[0] * jx + Jd[1] * jy);
    Jxi[1] = fv * (Jd[2] * jx + Jd[3] * jy);
  }
  if (Jint) {
    Jint[0] = kp[0]; Jint[1] = 0.0; Jint[2] = 1.0; Jint[3] = 0.0;
    Jint[4] = 0.0; Jint[5] = kp[1]; Jint[6] = 0.0; Jint[7] = 1.0;
  }
  kp[0] = fu * kp[0] + cu;
  kp[1] = fv * kp[1] + cv;
  const bool visible = kp[0] >= 0.0 && kp[1] >= 0.0 && kp[0] < static_cast<double>(width) &&
                       kp[1] < static_cast<double>(height);
  const double d2 = x * x + y * y + z * z, min_depth2 = 1e-10 * 1e-10;  // kMinimumDepth
  if (visible && d2 > min_depth2) return kVisible;
  if (!visible && d2 > min_depth2) return kOutside;
  return kInvalid;
}

// camera-pinhole.cc:65-164 + camera-pinhole-inl.h:48-65 + camera-inl.h:17-25;
// a unified-projection camera goes to projectUnified.
int project3Functional(int dist_type, const double* intr, const double* dist, int width,
                       int height, const double* p, double* kp, double* Jp, double* Jint,
                       double* Jdist, const CamModel& cm = CamModel(), double* Jxi = nullptr) {
  if (cm.model == MLBA_CAMERA_UNIFIED_PROJECTION)
    return projectUnified(cm.xi, dist_type, intr, dist, width, height, p, kp, Jp, Jint, Jdist, Jxi);
  if (Jxi) Jxi[0] = Jxi[1] = 0.0;
  const double fu = intr[0], fv = intr[1], cu = intr[2], cv = intr[3];
  const double x = p[0], y = p[1], z = p[2];
  const double rz = 1.0 / z;
  kp[0] = x * rz;
  kp[1] = y * rz;
  double Jd[4] = {1.0, 0.0, 0.0, 1.0};
  if (Jdist) {
    if (dist_type == MLBA_DISTORTION_RADTAN) radtanParamJac(kp, Jdist);
    else if (dist_type == MLBA_DISTORTION_EQUIDISTANT) equidistantParamJac(kp, Jdist);
    else if (dist_type == MLBA_DISTORTION_FISHEYE) fisheyeParamJac(dist, kp, Jdist);
    else for (int i = 0; i < 8; ++i) Jdist[i] = 0.0;
    for (int i = 0; i < 4; ++i) { Jdist[i] *= fu; Jdist[4 + i] *= fv; }
  }
  distort(dist_type, dist, kp, Jp ? Jd : nullptr);
  if (Jp) {
    const double rz2 = rz * rz;
    Jp[0] = fu * Jd[0] * rz;
    Jp[1] = fu * Jd[1] * rz;
    Jp[2] = -fu * (x * Jd[0] + y * Jd[1]) * rz2;
    Jp[3] = fv * Jd[2] * rz;
    Jp[4] = fv * Jd[3] * rz;
    Jp[5] = -fv * (x * Jd[2] + y * Jd[3]) * rz2;
  }
  if (Jint) {
    Jint[0] = kp[0]; Jint[1] = 0.0; Jint[2] = 1.0; Jint[3] = 0.0;
    Jint[4] = 0.0; Jint[5] = kp[1]; Jint[6] = 0.0; Jint[7] = 1.0;
  }
  kp[0] = fu * kp[0] + cu;
  kp[1] = fv * kp[1] + cv;
  const bool visible = kp[0] >= 0.0 && kp[1] >= 0.0 && kp[0] < static_cast<double>(width) &&
                       kp[1] < static_cast<double>(height);
  const double kMinimumDepth = 1e-10;  // camera-pinhole.cc:234
  if (visible && z > kMinimumDepth) return kVisible;
  if (!visible && z > kMinimumDepth) return kOutside;
  if (z < 0.0) return kBehind;
  return kInvalid;
}

// ------------------------------------------------------- visual error term

// visual-error-term-inl.h:13-379. params in the reference block order;
// jac entries nullable (ambient, row-major). Returns true on projection failure.
bool evaluateVisual(const double* meas, double sigma_inv, int type, int dist_type, int width,
                    int height, const double* const* params, double* res, double* const* jac,
                    const CamModel& cm = CamModel(), double* jac_xi = nullptr) {
  const double* p_B = params[0];
  const double* q_B_LM = params[1];
  const double* p_LM_B = params[1] + 4;
  const double* q_G_LM = params[2];
  const double* p_G_LM = params[2] + 4;
  const double* q_G_M = params[3];
  const double* p_G_M = params[3] + 4;
  const double* q_I_M = params[4];
  const double* p_M_I = params[4] + 4;
  const double* q_C_I = params[5];
  const double* p_C_I = params[6];
  const double* intr = params[7];
  const double* dist = params[8];

  double R_C_I[9], R_B_LM[9], R_LM_B[9], R_G_LM[9], R_G_M[9], R_M_G[9], R_I_M[9];
  rotJPL(q_C_I, R_C_I);
  double p_M[3] = {0, 0, 0}, p_G[3] = {0, 0, 0}, p_I[3] = {0, 0, 0};
  if (type == MLBA_LOCAL_KEYFRAME) {
    for (int i = 0; i < 3; ++i) p_I[i] = p_B[i];
  } else {
    rotJPL(q_B_LM, R_B_LM);
    transpose33(R_B_LM, R_LM_B);
    rotJPL(q_I_M, R_I_M);
    double t[3];
    if (type == MLBA_LOCAL_MISSION) {
      matmul(R_LM_B, p_B, p_M, 3, 3, 1);
      for (int i = 0; i < 3; ++i) { p_M[i] += p_LM_B[i]; t[i] = p_M[i] - p_M_I[i]; }
      matmul(R_I_M, t, p_I, 3, 3, 1);
    } else {
      rotJPL(q_G_LM, R_G_LM);
      rotJPL(q_G_M, R_G_M);
      transpose33(R_G_M, R_M_G);
      double p_LM[3];
      matmul(R_LM_B, p_B, p_LM, 3, 3, 1);
      for (int i = 0; i < 3; ++i) p_LM[i] += p_LM_B[i];
      matmul(R_G_LM, p_LM, p_G, 3, 3, 1);
      for (int i = 0; i < 3; ++i) { p_G[i] += p_G_LM[i]; t[i] = p_G[i] - p_G_M[i]; }
      matmul(R_M_G, t, p_M, 3, 3, 1);
      for (int i = 0; i < 3; ++i) t[i] = p_M[i] - p_M_I[i];
      matmul(R_I_M, t, p_I, 3, 3, 1);
    }
  }
  double p_C[3];
  matmul(R_C_I, p_I, p_C, 3, 3, 1);
  for (int i = 0; i < 3; ++i) p_C[i] += p_C_I[i];

  const bool want_j = jac != nullptr;
  double kp[2], Jkp[6], Jint[8], Jdist[8], Jxi[2];
  const int status = project3Functional(dist_type, intr, dist, width, height, p_C, kp,
                                        want_j ? Jkp : nullptr,
                                        (want_j && jac[7]) ? Jint : nullptr,
                                        (want_j && jac[8] && dist_type != MLBA_DISTORTION_NONE)
                                            ? Jdist : nullptr,
                                        cm, jac_xi ? Jxi : nullptr);
  const bool failed = status == kBehind || status == kInvalid || p_C[2] < 0.05 ||
                      (kp[0] * kp[0] + kp[1] * kp[1]) > 1.0e5 * 1.0e5;

  if (want_j) {
    double J_pC_pCI[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    double tmp[3], J_pC_qCI[9];
    matmul(R_C_I, p_I, tmp, 3, 3, 1);
    skew(tmp, J_pC_qCI);
    double J_pMI[9] = {0}, J_pGM[9] = {0}, J_pGLM[9] = {0}, J_pLMB[9] = {0}, J_pB[9] = {0};
    double J_qBLM[9] = {0}, J_qGLM[9] = {0}, J_qGM[9] = {0}, J_qIM[9] = {0};
    double sk[9], A[9];
    if (type == MLBA_GLOBAL) {
      matmul(R_C_I, R_I_M, A, 3, 3, 3);
      for (int i = 0; i < 9; ++i) J_pMI[i] = -A[i];
      matmul(J_pMI, R_M_G, J_pGM, 3, 3, 3);
      for (int i = 0; i < 9; ++i) J_pGLM[i] = -J_pGM[i];
      matmul(J_pGLM, R_G_LM, J_pLMB, 3, 3, 3);
      matmul(J_pLMB, R_LM_B, J_pB, 3, 3, 3);
      skew(p_B, sk);
      matmul(J_pB, sk, J_qBLM, 3, 3, 3);
      for (int i = 0; i < 9; ++i) J_qBLM[i] = -J_qBLM[i];
      skew(p_G, sk);
      matmul(J_pGLM, sk, J_qGLM, 3, 3, 3);
      matmul(J_pGM, sk, J_qGM, 3, 3, 3);
      skew(p_I, sk);
      matmul(R_C_I, sk, J_qIM, 3, 3, 3);
    } else if (type == MLBA_LOCAL_MISSION) {
      matmul(R_C_I, R_I_M, A, 3, 3, 3);
      for (int i = 0; i < 9; ++i) { J_pMI[i] = -A[i]; J_pLMB[i] = A[i]; }
      skew(p_I, sk);
      matmul(R_C_I, sk, J_qIM, 3, 3, 3);
      matmul(J_pLMB, R_LM_B, J_pB, 3, 3, 3);
      skew(p_B, sk);
      matmul(J_pB, sk, J_qBLM, 3, 3, 3);
      for (int i = 0; i < 9; ++i) J_qBLM[i] = -J_qBLM[i];
    } else {
      for (int i = 0; i < 9; ++i) J_pB[i] = R_C_I[i];
    }
    auto posBlock = [&](const double* Jx, double* out, int stride, int col0) {
      double t6[6];
      matmul(Jkp, Jx, t6, 2, 3, 3);
      for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 3; ++c) out[r * stride + col0 + c] = failed ? 0.0 : t6[r * 3 + c] * sigma_inv;
    };
    auto rotBlock = [&](const double* Jx, const double* q, double* out, int stride) {
      double Jq[12], JqT[12], t6[6], t8[8];
      quatJacobian(q, Jq);
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 3; ++j) JqT[j * 4 + i] = Jq[i * 3 + j];
      matmul(Jkp, Jx, t6, 2, 3, 3);
      matmul(t6, JqT, t8, 2, 3, 4);
      for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 4; ++c) out[r * stride + c] = failed ? 0.0 : t8[r * 4 + c] * 4.0 * sigma_inv;
    };
    if (jac[0]) posBlock(J_pB, jac[0], 3, 0);
    if (jac[1]) { rotBlock(J_qBLM, q_B_LM, jac[1], 7); posBlock(J_pLMB, jac[1], 7, 4); }
    if (jac[2]) { rotBlock(J_qGLM, q_G_LM, jac[2], 7); posBlock(J_pGLM, jac[2], 7, 4); }
    if (jac[3]) { rotBlock(J_qGM, q_G_M, jac[3], 7); posBlock(J_pGM, jac[3], 7, 4); }
    if (jac[4]) { rotBlock(J_qIM, q_I_M, jac[4], 7); posBlock(J_pMI, jac[4], 7, 4); }
    if (jac[5]) rotBlock(J_pC_qCI, q_C_I, jac[5], 4);
    if (jac[6]) posBlock(J_pC_pCI, jac[6], 3, 0);
    if (jac[7])
      for (int i = 0; i < 8; ++i) jac[7][i] = failed ? 0.0 : Jint[i] * sigma_inv;
    if (jac[8]) {
      for (int i = 0; i < 8; ++i)
        jac[8][i] = (failed || dist_type == MLBA_DISTORTION_NONE) ? 0.0 : Jdist[i] * sigma_inv;
    }
  }
  if (jac_xi) {
    jac_xi[0] = failed ? 0.0 : Jxi[0] * sigma_inv;
    jac_xi[1] = failed ? 0.0 : Jxi[1] * sigma_inv;
  }
  if (!failed) {
    res[0] = (kp[0] - meas[0]) * sigma_inv;
    res[1] = (kp[1] - meas[1]) * sigma_inv;
  } else {
    res[0] = res[1] = 0.0;
  }
  return failed;
}

// ------------------------------------------------------------ robust loss

// ceres: HuberLoss::Evaluate, CauchyLoss::Evaluate, TrivialLoss.
void lossEvaluate(int type, double a, double s, double* rho) {
  if (type == MLBA_LOSS_MAPLAB_HUBER) {
    const double b = a * a;
    if (s > b) {
      const double r = std::sqrt(s);
      rho[0] = 2.0 * a * r - b;
      rho[1] = std::max(std::numeric_limits<double>::min(), a / r);
      rho[2] = -rho[1] / (2.0 * s);
    } else {
      rho[0] = s; rho[1] = 1.0; rho[2] = 0.0;
    }
  } else if (type == MLBA_LOSS_CAUCHY) {
    const double b = a * a, c = 1.0 / b;
    const double sum = 1.0 + s * c, inv = 1.0 / sum;
    rho[0] = b * std::log(sum);
    rho[1] = std::max(std::numeric_limits<double>::min(), inv);
    rho[2] = -c * (inv * inv);
  } else {
    rho[0] = s; rho[1] = 1.0; rho[2] = 0.0;
  }
}

// Loss scale `a` of one observation (optimization-terms-addition.cc:189-198:
// HuberLoss(delta * sigma), delta = 3, or 10 for the global error type).
double lossScale(const mlba_options& o, int type, double sigma) {
  if (o.loss_type == MLBA_LOSS_MAPLAB_HUBER) return (type == MLBA_GLOBAL ? 10.0 : 3.0) * sigma;
  if (o.loss_type == MLBA_LOSS_CAUCHY) return o.loss_scale_px / sigma;
  return 1.0;
}

// ceres: Corrector (robust-loss correction of residual and Jacobian).
struct Corrector {
  double sqrt_rho1, residual_scaling, alpha_sq_norm;
  Corrector(double sq_norm, const double* rho) {
    sqrt_rho1 = std::sqrt(rho[1]);
    if (sq_norm == 0.0 || rho[2] <= 0.0) {
      residual_scaling = sqrt_rho1;
      alpha_sq_norm = 0.0;
      return;
    }
    const double D = 1.0 + 2.0 * sq_norm * rho[2] / rho[1];
    const double alpha = 1.0 - std::sqrt(D);
    residual_scaling = sqrt_rho1 / (1.0 - alpha);
    alpha_sq_norm = alpha / sq_norm;
  }
  // Uses the uncorrected residual r (ceres corrects J before r).
  void correctJacobian(const double* r, double* J, int cols) const {
    if (alpha_sq_norm == 0.0) {
      for (int i = 0; i < 2 * cols; ++i) J[i] *= sqrt_rho1;
      return;
    }
    for (int c = 0; c < cols; ++c) {
      const double rtj = J[c] * r[0] + J[cols + c] * r[1];
      for (int k = 0; k < 2; ++k) J[k * cols + c] = sqrt_rho1 * (J[k * cols + c] - alpha_sq_norm * r[k] * rtj);
    }
  }
};

// ------------------------------------------------------------ the problem

struct Problem {
  const mlba_problem_desc* d;
  int P, L, N;
  std::vector<double> vertices, landmarks;  // mutable state
  std::vector<int> obs_type;                // LandmarkErrorType per observation
  std::vector<uint8_t> lm_active;
  // Variable-block indexing: -1 when the block is constant or not in the problem.
  std::vector<int> pose_var, lm_var;
  std::vector<int> var_pose_list, var_lm_list;
  std::vector<int> active_obs;              // observations in the problem

  explicit Problem(const mlba_problem_desc* desc, const uint8_t* active) : d(desc) {
    P = desc->num_vertices; L = desc->num_landmarks; N = desc->num_observations;
    vertices.assign(desc->vertex_q_IM__M_p_MI, desc->vertex_q_IM__M_p_MI + 7 * P);
    landmarks.assign(desc->landmark_p_B, desc->landmark_p_B + 3 * L);
    obs_type.resize(N);
    for (int i = 0; i < N; ++i) {
      const int v = desc->obs_vertex[i];
      const int s = desc->landmark_store_vertex[desc->obs_landmark[i]];
      if (v == s) obs_type[i] = MLBA_LOCAL_KEYFRAME;
      else if (desc->vertex_mission[v] == desc->vertex_mission[s]) obs_type[i] = MLBA_LOCAL_MISSION;
      else obs_type[i] = MLBA_GLOBAL;
    }
    lm_active.assign(L, 1);
    if (active) for (int l = 0; l < L; ++l) lm_active[l] = active[l] ? 1 : 0;
    rebuildIndex();
  }
  bool poseConst(int v) const { return d->vertex_constant && d->vertex_constant[v]; }
  bool lmConst(int l) const { return d->landmark_constant && d->landmark_constant[l]; }

  // Parameter blocks exist in the (Ceres) problem only when some active
  // residual uses them; constant blocks are not optimized.
  void rebuildIndex() {
    active_obs.clear();
    pose_var.assign(P, -1);
    lm_var.assign(L, -1);
    std::vector<uint8_t> pose_used(P, 0), lm_used(L, 0);
    for (int i = 0; i < N; ++i) {
      const int l = d->obs_landmark[i];
      if (!lm_active[l]) continue;
      active_obs.push_back(i);
      lm_used[l] = 1;
      if (obs_type[i] != MLBA_LOCAL_KEYFRAME) {
        pose_used[d->obs_vertex[i]] = 1;
        pose_used[d->landmark_store_vertex[l]] = 1;
      }
    }
    var_pose_list.clear();
    var_lm_list.clear();
    for (int v = 0; v < P; ++v)
      if (pose_used[v] && !poseConst(v)) { pose_var[v] = (int)var_pose_list.size(); var_pose_list.push_back(v); }
    for (int l = 0; l < L; ++l)
      if (lm_used[l] && !lmConst(l)) { lm_var[l] = (int)var_lm_list.size(); var_lm_list.push_back(l); }
  }
  int numTangent() const { return 6 * (int)var_pose_list.size() + 3 * (int)var_lm_list.size(); }

  // Gather the 9 reference parameter blocks of observation i. Dummies for
  // unused blocks are never read by the error term
  // (landmark-common-inl.h:9-51).
  void params(int i, const double* verts, const double* lms, const double** p) const {
    const int l = d->obs_landmark[i];
    const int v = d->obs_vertex[i];
    const int s = d->landmark_store_vertex[l];
    const int c = d->obs_camera[i];
    p[0] = lms + 3 * l;
    p[1] = verts + 7 * s;
    p[2] = d->baseframe_q_GM__G_p_GM + 7 * d->vertex_mission[s];
    p[3] = d->baseframe_q_GM__G_p_GM + 7 * d->vertex_mission[v];
    p[4] = verts + 7 * v;
    p[5] = d->camera_q_CI__C_p_CI + 7 * c;
    p[6] = d->camera_q_CI__C_p_CI + 7 * c + 4;
    p[7] = d->camera_intrinsics + 4 * c;
    p[8] = d->camera_distortion + 4 * c;
  }
};

// Per residual data of one linearization (J scaled and robust-corrected).
struct Linearization {
  // For each active observation k: residual (2), J wrt landmark (2x3),
  // observer pose (2x6), anchor pose (2x6), and which are variable.
  std::vector<double> r, Jl, Jo, Ja;
  std::vector<int> il, io, ia;  // variable indices (-1 if not variable)
  double cost = 0.0, fixed_cost = 0.0;
  std::vector<double> gradient;  // unscaled tangent gradient J^T r
};

// Lift an ambient pose Jacobian (2x7) to the tangent space (2x6) with the
// JplPoseParameterization = Product(JPL quaternion, Identity(3))
// (pose-param-jpl.h:11-18).
void liftPose(const double* Jamb, const double* q, double* Jt) {
  double Jq[12];
  quatJacobian(q, Jq);
  for (int r = 0; r < 2; ++r) {
    for (int c = 0; c < 3; ++c) {
      double s = 0.0;
      for (int k = 0; k < 4; ++k) s += Jamb[r * 7 + k] * Jq[k * 3 + c];
      Jt[r * 6 + c] = s;
    }
    for (int c = 0; c < 3; ++c) Jt[r * 6 + 3 + c] = Jamb[r * 7 + 4 + c];
  }
}

// ceres: Evaluator -- residuals, robust correction and tangent Jacobians of
// every active residual block, the cost 0.5 * sum rho, and the gradient.
void linearize(const Problem& pb, const mlba_options& o, int nthreads, bool with_jacobian,
               Linearization* lin) {
  const int K = (int)pb.active_obs.size();
  const mlba_problem_desc* d = pb.d;
  if (with_jacobian) {
    lin->r.assign(2 * K, 0.0);
    lin->Jl.assign(6 * K, 0.0);
    lin->Jo.assign(12 * K, 0.0);
    lin->Ja.assign(12 * K, 0.0);
    lin->il.assign(K, -1); lin->io.assign(K, -1); lin->ia.assign(K, -1);
  }
  std::vector<double> part_cost(std::max(1, nthreads), 0.0), part_fixed(std::max(1, nthreads), 0.0);
  parallel_for(K, nthreads, [&](int b, int e, int t) {
    double c_sum = 0.0, f_sum = 0.0;
    for (int k = b; k < e; ++k) {
      const int i = pb.active_obs[k];
      const double* params[9];
      pb.params(i, pb.vertices.data(), pb.landmarks.data(), params);
      const int l = d->obs_landmark[i];
      const int v = d->obs_vertex[i];
      const int s = d->landmark_store_vertex[l];
      const int c = d->obs_camera[i];
      const int type = pb.obs_type[i];
      const int vl = pb.lm_var[l];
      const int vo = type != MLBA_LOCAL_KEYFRAME ? pb.pose_var[v] : -1;
      const int va = type != MLBA_LOCAL_KEYFRAME ? pb.pose_var[s] : -1;
      double r[2], J0[6], J1[14], J4[14];
      double* jac[9] = {nullptr};
      if (with_jacobian) {
        jac[0] = vl >= 0 ? J0 : nullptr;
        jac[1] = va >= 0 ? J1 : nullptr;
        jac[4] = vo >= 0 ? J4 : nullptr;
      }
      const double sigma = d->obs_sigma[i];
      evaluateVisual(d->obs_keypoint + 2 * i, 1.0 / sigma, type, d->camera_distortion_type[c],
                     d->camera_resolution[2 * c], d->camera_resolution[2 * c + 1], params, r,
                     with_jacobian ? jac : nullptr, camModel(d, c));
      const double sq = r[0] * r[0] + r[1] * r[1];
      double rho[3];
      lossEvaluate(o.loss_type, lossScale(o, type, sigma), sq, rho);
      const bool fixed = vl < 0 && vo < 0 && va < 0;
      if (fixed) f_sum += 0.5 * rho[0]; else c_sum += 0.5 * rho[0];
      if (!with_jacobian) continue;
      Corrector corr(sq, rho);
      if (vl >= 0) { corr.correctJacobian(r, J0, 3); std::memcpy(&lin->Jl[6 * k], J0, sizeof(J0)); }
      if (vo >= 0) {
        double Jt[12];
        liftPose(J4, params[4], Jt);
        corr.correctJacobian(r, Jt, 6);
        std::memcpy(&lin->Jo[12 * k], Jt, sizeof(Jt));
      }
      if (va >= 0) {
        double Jt[12];
        liftPose(J1, params[1], Jt);
        corr.correctJacobian(r, Jt, 6);
        std::memcpy(&lin->Ja[12 * k], Jt, sizeof(Jt));
      }
      lin->r[2 * k] = r[0] * corr.residual_scaling;
      lin->r[2 * k + 1] = r[1] * corr.residual_scaling;
      lin->il[k] = vl; lin->io[k] = vo; lin->ia[k] = va;
    }
    part_cost[t] = c_sum;
    part_fixed[t] = f_sum;
  });
  lin->cost = 0.0;
  lin->fixed_cost = 0.0;
  for (size_t t = 0; t < part_cost.size(); ++t) { lin->cost += part_cost[t]; lin->fixed_cost += part_fixed[t]; }
  if (!with_jacobian) return;
  // gradient = J^T r (tangent, unscaled), accumulated in observation order.
  const int np = (int)pb.var_pose_list.size();
  lin->gradient.assign(pb.numTangent(), 0.0);
  for (int k = 0; k < K; ++k) {
    const double* r = &lin->r[2 * k];
    if (lin->il[k] >= 0)
      for (int c = 0; c < 3; ++c)
        lin->gradient[6 * np + 3 * lin->il[k] + c] += lin->Jl[6 * k + c] * r[0] + lin->Jl[6 * k + 3 + c] * r[1];
    if (lin->io[k] >= 0)
      for (int c = 0; c < 6; ++c)
        lin->gradient[6 * lin->io[k] + c] += lin->Jo[12 * k + c] * r[0] + lin->Jo[12 * k + 6 + c] * r[1];
    if (lin->ia[k] >= 0)
      for (int c = 0; c < 6; ++c)
        lin->gradient[6 * lin->ia[k] + c] += lin->Ja[12 * k + c] * r[0] + lin->Ja[12 * k + 6 + c] * r[1];
  }
}

// Tangent index layout: [6 per variable pose | 3 per variable landmark].
// ceres: Jacobi scaling -- column c scaled by 1 / (1 + ||J_c||).
void columnSquaredNorms(const Problem& pb, const Linearization& lin, std::vector<double>* out) {
  const int np = (int)pb.var_pose_list.size();
  out->assign(pb.numTangent(), 0.0);
  const int K = (int)pb.active_obs.size();
  for (int k = 0; k < K; ++k) {
    if (lin.il[k] >= 0)
      for (int c = 0; c < 3; ++c) {
        const double a = lin.Jl[6 * k + c], b = lin.Jl[6 * k + 3 + c];
        (*out)[6 * np + 3 * lin.il[k] + c] += a * a + b * b;
      }
    if (lin.io[k] >= 0)
      for (int c = 0; c < 6; ++c) {
        const double a = lin.Jo[12 * k + c], b = lin.Jo[12 * k + 6 + c];
        (*out)[6 * lin.io[k] + c] += a * a + b * b;
      }
    if (lin.ia[k] >= 0)
      for (int c = 0; c < 6; ++c) {
        const double a = lin.Ja[12 * k + c], b = lin.Ja[12 * k + 6 + c];
        (*out)[6 * lin.ia[k] + c] += a * a + b * b;
      }
  }
}

void scaleColumns(const Problem& pb, const std::vector<double>& s, Linearization* lin) {
  const int np = (int)pb.var_pose_list.size();
  const int K = (int)pb.active_obs.size();
  for (int k = 0; k < K; ++k) {
    if (lin->il[k] >= 0)
      for (int c = 0; c < 3; ++c) {
        const double f = s[6 * np + 3 * lin->il[k] + c];
        lin->Jl[6 * k + c] *= f; lin->Jl[6 * k + 3 + c] *= f;
      }
    if (lin->io[k] >= 0)
      for (int c = 0; c < 6; ++c) {
        const double f = s[6 * lin->io[k] + c];
        lin->Jo[12 * k + c] *= f; lin->Jo[12 * k + 6 + c] *= f;
      }
    if (lin->ia[k] >= 0)
      for (int c = 0; c < 6; ++c) {
        const double f = s[6 * lin->ia[k] + c];
        lin->Ja[12 * k + c] *= f; lin->Ja[12 * k + 6 + c] *= f;
      }
  }
}

// ------------------------------------------------------------ dense algebra

// Dot product of two panel rows with eight independent partial sums (the
// chains vectorize and overlap; the sum order is fixed).
static inline double panel_dot(const double* a, const double* b, int w) {
  double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int t = 0;
  for (; t + 8 <= w; t += 8)
    for (int u = 0; u < 8; ++u) s[u] += a[t + u] * b[t + u];
  for (; t < w; ++t) s[0] += a[t] * b[t];
  return ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
}

// Blocked right-looking Cholesky of an n x n row-major SPD matrix (lower
// factor in place). Returns false when not positive definite.
bool cholesky(std::vector<double>& A, int n, int nthreads) {
  const int B = 64;
  for (int k0 = 0; k0 < n; k0 += B) {
    const int k1 = std::min(n, k0 + B);
    // Factor the diagonal block.
    for (int j = k0; j < k1; ++j) {
      double s = A[(size_t)j * n + j];
      for (int t = k0; t < j; ++t) s -= A[(size_t)j * n + t] * A[(size_t)j * n + t];
      if (!(s > 0.0)) return false;
      const double ljj = std::sqrt(s);
      A[(size_t)j * n + j] = ljj;
      for (int i = j + 1; i < k1; ++i) {
        double v = A[(size_t)i * n + j];
        for (int t = k0; t < j; ++t) v -= A[(size_t)i * n + t] * A[(size_t)j * n + t];
        A[(size_t)i * n + j] = v / ljj;
      }
    }
    // Panel below: rows k1..n, columns k0..k1.
    parallel_for(n - k1, nthreads, [&](int b, int e, int) {
      for (int i = k1 + b; i < k1 + e; ++i)
        for (int j = k0; j < k1; ++j) {
          double v = A[(size_t)i * n + j];
          for (int t = k0; t < j; ++t) v -= A[(size_t)i * n + t] * A[(size_t)j * n + t];
          A[(size_t)i * n + j] = v / A[(size_t)j * n + j];
        }
    });
    // Trailing update of the lower triangle: A_ij -= L_ik L_jk^T, four rows
    // i at a time (each panel row j is read once per quad). Quads are taken
    // from both ends in turn so that every thread's share of the triangle
    // is about the same.
    const int w = k1 - k0, nq = (n - k1 + 3) / 4;
    parallel_for(nq, nthreads, [&](int b, int e, int) {
      for (int idx = b; idx < e; ++idx) {
        const int q = (idx & 1) ? nq - 1 - idx / 2 : idx / 2;
        const int i0 = k1 + 4 * q, ni = std::min(4, n - i0);
        for (int j = k1; j < i0 + ni; ++j) {
          const double* lj = &A[(size_t)j * n + k0];
          for (int r = std::max(0, j - i0); r < ni; ++r) {
            const int i = i0 + r;
            A[(size_t)i * n + j] -= panel_dot(&A[(size_t)i * n + k0], lj, w);
          }
        }
      }
    });
  }
  return true;
}

void choleskySolve(const std::vector<double>& L, int n, std::vector<double>& x) {
  for (int i = 0; i < n; ++i) {
    double s = x[i];
    for (int t = 0; t < i; ++t) s -= L[(size_t)i * n + t] * x[t];
    x[i] = s / L[(size_t)i * n + i];
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = x[i];
    for (int t = i + 1; t < n; ++t) s -= L[(size_t)t * n + i] * x[t];
    x[i] = s / L[(size_t)i * n + i];
  }
}

bool invertSpd3(const double* A, double* inv) {
  // Cholesky of a 3x3 SPD matrix, then the inverse.
  const double l00 = std::sqrt(A[0]);
  if (!(l00 > 0)) return false;
  const double l10 = A[3] / l00, l20 = A[6] / l00;
  const double d11 = A[4] - l10 * l10;
  if (!(d11 > 0)) return false;
  const double l11 = std::sqrt(d11);
  const double l21 = (A[7] - l20 * l10) / l11;
  const double d22 = A[8] - l20 * l20 - l21 * l21;
  if (!(d22 > 0)) return false;
  const double l22 = std::sqrt(d22);
  // inverse of L
  const double i00 = 1 / l00, i11 = 1 / l11, i22 = 1 / l22;
  const double i10 = -l10 * i00 * i11;
  const double i21 = -l21 * i11 * i22;
  const double i20 = -(l20 * i00 + l21 * i10) * i22;
  // A^-1 = L^-T L^-1
  const double Li[9] = {i00, 0, 0, i10, i11, 0, i20, i21, i22};
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += Li[k * 3 + r] * Li[k * 3 + c];
      inv[r * 3 + c] = s;
    }
  return true;
}

// ------------------------------------------------------------ linear solve

// Per-landmark elimination data of one LM step.
struct Elimination {
  std::vector<double> Vinv;                  // 9 per variable landmark
  std::vector<std::vector<int>> inc_pose;    // variable poses touching the landmark
  std::vector<std::vector<double>> inc_W;    // 18 per incidence (3x6)
  std::vector<double> gl;                    // 3 per variable landmark
};

// Solve (J^T J + D^T D) y = J^T r for the scaled, corrected J held in lin.
// D is given per tangent column (already sqrt(diag / mu)). Returns the number
// of iterations of the inner solver (1 for direct), or -1 on failure.
int solveNormalEquations(const Problem& pb, const Linearization& lin, const std::vector<double>& D,
                         int solver, int nthreads, double pcg_tol, int pcg_max,
                         std::vector<double>* y, std::vector<double>* S_out = nullptr,
                         std::vector<double>* rhs_out = nullptr) {
  const int np = (int)pb.var_pose_list.size();
  const int nl = (int)pb.var_lm_list.size();
  const int n = 6 * np + 3 * nl;
  const int K = (int)pb.active_obs.size();
  y->assign(n, 0.0);
  // Jacobian-transpose-residual (rhs of the normal equations).
  std::vector<double> g(n, 0.0);
  for (int k = 0; k < K; ++k) {
    const double* r = &lin.r[2 * k];
    if (lin.il[k] >= 0)
      for (int c = 0; c < 3; ++c) g[6 * np + 3 * lin.il[k] + c] += lin.Jl[6 * k + c] * r[0] + lin.Jl[6 * k + 3 + c] * r[1];
    if (lin.io[k] >= 0)
      for (int c = 0; c < 6; ++c) g[6 * lin.io[k] + c] += lin.Jo[12 * k + c] * r[0] + lin.Jo[12 * k + 6 + c] * r[1];
    if (lin.ia[k] >= 0)
      for (int c = 0; c < 6; ++c) g[6 * lin.ia[k] + c] += lin.Ja[12 * k + c] * r[0] + lin.Ja[12 * k + 6 + c] * r[1];
  }

  if (solver == ORACLE_DENSE_NORMAL) {
    std::vector<double> A((size_t)n * n, 0.0);
    for (int k = 0; k < K; ++k) {
      int idx[15];
      double col[2][15];
      int m = 0;
      auto add = [&](int base, const double* J, int w) {
        for (int c = 0; c < w; ++c) { idx[m] = base + c; col[0][m] = J[c]; col[1][m] = J[w + c]; ++m; }
      };
      if (lin.il[k] >= 0) add(6 * np + 3 * lin.il[k], &lin.Jl[6 * k], 3);
      if (lin.io[k] >= 0) add(6 * lin.io[k], &lin.Jo[12 * k], 6);
      if (lin.ia[k] >= 0) add(6 * lin.ia[k], &lin.Ja[12 * k], 6);
      for (int a = 0; a < m; ++a)
        for (int b = 0; b < m; ++b)
          A[(size_t)idx[a] * n + idx[b]] += col[0][a] * col[0][b] + col[1][a] * col[1][b];
    }
    for (int i = 0; i < n; ++i) A[(size_t)i * n + i] += D[i] * D[i];
    if (!cholesky(A, n, nthreads)) return -1;
    *y = g;
    choleskySolve(A, n, *y);
    return 1;
  }

  // ceres: Schur elimination of the landmark (e-)blocks.
  // Group active observations by variable landmark.
  std::vector<std::vector<int>> lm_obs(nl);
  for (int k = 0; k < K; ++k)
    if (lin.il[k] >= 0) lm_obs[lin.il[k]].push_back(k);
  Elimination el;
  el.Vinv.assign(9 * nl, 0.0);
  el.inc_pose.assign(nl, {});
  el.inc_W.assign(nl, {});
  el.gl.assign(3 * nl, 0.0);
  bool ok = true;
  parallel_for(nl, nthreads, [&](int b, int e, int) {
    for (int j = b; j < e; ++j) {
      double V[9] = {0};
      std::vector<int>& poses = el.inc_pose[j];
      std::vector<double>& W = el.inc_W[j];
      auto slot = [&](int p) {
        for (size_t s = 0; s < poses.size(); ++s) if (poses[s] == p) return (int)s;
        poses.push_back(p);
        W.resize(W.size() + 18, 0.0);
        return (int)poses.size() - 1;
      };
      for (int k : lm_obs[j]) {
        const double* Jl = &lin.Jl[6 * k];
        for (int a = 0; a < 3; ++a)
          for (int c = 0; c < 3; ++c) V[a * 3 + c] += Jl[a] * Jl[c] + Jl[3 + a] * Jl[3 + c];
        for (int a = 0; a < 3; ++a) el.gl[3 * j + a] += Jl[a] * lin.r[2 * k] + Jl[3 + a] * lin.r[2 * k + 1];
        auto addW = [&](int p, const double* Jp) {
          double* Wp = &W[18 * slot(p)];
          for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 6; ++c) Wp[a * 6 + c] += Jl[a] * Jp[c] + Jl[3 + a] * Jp[6 + c];
        };
        if (lin.io[k] >= 0) addW(lin.io[k], &lin.Jo[12 * k]);
        if (lin.ia[k] >= 0) addW(lin.ia[k], &lin.Ja[12 * k]);
      }
      for (int a = 0; a < 3; ++a) V[a * 4] += D[6 * np + 3 * j + a] * D[6 * np + 3 * j + a];
      if (!invertSpd3(V, &el.Vinv[9 * j])) ok = false;
    }
  });
  if (!ok) return -1;

  // Reduced camera system S = F'F + D_f^2 - sum_l W_l^T V_l^-1 W_l, dense.
  const int nf = 6 * np;
  std::vector<double> S((size_t)nf * nf, 0.0), rhs(g.begin(), g.begin() + nf);
  // Direct pose-pose terms F'F, row-partitioned by pose to stay deterministic.
  std::vector<std::vector<int>> pose_obs(np), pose_lms(np);
  for (int k = 0; k < K; ++k) {
    if (lin.io[k] >= 0) pose_obs[lin.io[k]].push_back(k);
    if (lin.ia[k] >= 0) pose_obs[lin.ia[k]].push_back(k);
  }
  for (int j = 0; j < nl; ++j)
    for (int p : el.inc_pose[j]) pose_lms[p].push_back(j);
  parallel_for(np, nthreads, [&](int b, int e, int) {
    for (int p = b; p < e; ++p) {
      double* Srow = &S[(size_t)6 * p * nf];
      for (int k : pose_obs[p]) {
        const double* Jp = lin.io[k] == p ? &lin.Jo[12 * k] : &lin.Ja[12 * k];
        auto addF = [&](int q, const double* Jq) {
          for (int a = 0; a < 6; ++a)
            for (int c = 0; c < 6; ++c)
              Srow[(size_t)a * nf + 6 * q + c] += Jp[a] * Jq[c] + Jp[6 + a] * Jq[6 + c];
        };
        if (lin.io[k] >= 0) addF(lin.io[k], &lin.Jo[12 * k]);
        if (lin.ia[k] >= 0) addF(lin.ia[k], &lin.Ja[12 * k]);
      }
      for (int a = 0; a < 6; ++a) Srow[(size_t)a * nf + 6 * p + a] += D[6 * p + a] * D[6 * p + a];
      for (int j : pose_lms[p]) {
        const std::vector<int>& poses = el.inc_pose[j];
        const double* Vi = &el.Vinv[9 * j];
        int sp = 0;
        while (poses[sp] != p) ++sp;
        const double* Wp = &el.inc_W[j][18 * sp];
        double A[18];  // W_p^T V^-1 (6x3)
        for (int a = 0; a < 6; ++a)
          for (int c = 0; c < 3; ++c) {
            double s = 0;
            for (int t = 0; t < 3; ++t) s += Wp[t * 6 + a] * Vi[t * 3 + c];
            A[a * 3 + c] = s;
          }
        for (int a = 0; a < 6; ++a) {
          double s = 0;
          for (int t = 0; t < 3; ++t) s += A[a * 3 + t] * el.gl[3 * j + t];
          rhs[6 * p + a] -= s;
        }
        for (size_t q = 0; q < poses.size(); ++q) {
          const double* Wq = &el.inc_W[j][18 * q];
          for (int a = 0; a < 6; ++a)
            for (int c = 0; c < 6; ++c) {
              double s = 0;
              for (int t = 0; t < 3; ++t) s += A[a * 3 + t] * Wq[t * 6 + c];
              Srow[(size_t)a * nf + 6 * poses[q] + c] -= s;
            }
        }
      }
    }
  });
  if (S_out) *S_out = S;
  if (rhs_out) *rhs_out = rhs;

  std::vector<double> yp(nf, 0.0);
  int iters = 1;
  if (nf > 0) {
    if (solver == ORACLE_SCHUR_PCG) {
      // ceres: ConjugateGradientsSolver with the SCHUR_JACOBI preconditioner
      // (inverse of the 6x6 diagonal blocks of S); stops at
      // ||r_k|| <= tol * ||b||.
      std::vector<double> Minv((size_t)np * 36);
      for (int p = 0; p < np; ++p) {
        std::vector<double> blk(36);
        for (int a = 0; a < 6; ++a)
          for (int c = 0; c < 6; ++c) blk[a * 6 + c] = S[(size_t)(6 * p + a) * nf + 6 * p + c];
        if (!cholesky(blk, 6, 1)) return -1;
        for (int c = 0; c < 6; ++c) {
          std::vector<double> e(6, 0.0);
          e[c] = 1.0;
          choleskySolve(blk, 6, e);
          for (int a = 0; a < 6; ++a) Minv[(size_t)p * 36 + a * 6 + c] = e[a];
        }
      }
      auto matvec = [&](const std::vector<double>& x, std::vector<double>& out) {
        parallel_for(nf, nthreads, [&](int b, int e, int) {
          for (int i = b; i < e; ++i) {
            double s = 0;
            const double* row = &S[(size_t)i * nf];
            for (int t = 0; t < nf; ++t) s += row[t] * x[t];
            out[i] = s;
          }
        });
      };
      auto precond = [&](const std::vector<double>& r, std::vector<double>& z) {
        for (int p = 0; p < np; ++p)
          for (int a = 0; a < 6; ++a) {
            double s = 0;
            for (int c = 0; c < 6; ++c) s += Minv[(size_t)p * 36 + a * 6 + c] * r[6 * p + c];
            z[6 * p + a] = s;
          }
      };
      std::vector<double> r(rhs), z(nf), pdir(nf), q(nf);
      double bnorm = 0;
      for (double v : rhs) bnorm += v * v;
      bnorm = std::sqrt(bnorm);
      iters = 0;
      if (bnorm > 0) {
        precond(r, z);
        pdir = z;
        double rz = 0;
        for (int i = 0; i < nf; ++i) rz += r[i] * z[i];
        for (iters = 1; iters <= pcg_max; ++iters) {
          matvec(pdir, q);
          double pq = 0;
          for (int i = 0; i < nf; ++i) pq += pdir[i] * q[i];
          const double alpha = rz / pq;
          double rn = 0;
          for (int i = 0; i < nf; ++i) { yp[i] += alpha * pdir[i]; r[i] -= alpha * q[i]; rn += r[i] * r[i]; }
          if (std::sqrt(rn) <= pcg_tol * bnorm) break;
          precond(r, z);
          double rz_new = 0;
          for (int i = 0; i < nf; ++i) rz_new += r[i] * z[i];
          const double beta = rz_new / rz;
          rz = rz_new;
          for (int i = 0; i < nf; ++i) pdir[i] = z[i] + beta * pdir[i];
        }
        iters = std::min(iters, pcg_max);
      }
    } else {
      if (!cholesky(S, nf, nthreads)) return -1;
      yp = rhs;
      choleskySolve(S, nf, yp);
    }
  }
  for (int i = 0; i < nf; ++i) (*y)[i] = yp[i];
  // Back substitution: y_l = V^-1 (g_l - sum_p W_lp y_p).
  for (int j = 0; j < nl; ++j) {
    double t[3] = {el.gl[3 * j], el.gl[3 * j + 1], el.gl[3 * j + 2]};
    const std::vector<int>& poses = el.inc_pose[j];
    for (size_t q = 0; q < poses.size(); ++q) {
      const double* Wq = &el.inc_W[j][18 * q];
      for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 6; ++c) t[a] -= Wq[a * 6 + c] * yp[6 * poses[q] + c];
    }
    for (int a = 0; a < 3; ++a) {
      double s = 0;
      for (int c = 0; c < 3; ++c) s += el.Vinv[9 * j + a * 3 + c] * t[c];
      (*y)[nf + 3 * j + a] = s;
    }
  }
  return iters;
}

// ------------------------------------------------------------ LM minimizer

double vecNorm(const std::vector<double>& v) {
  double s = 0;
  for (double x : v) s += x * x;
  return std::sqrt(s);
}

// ceres: Evaluator::Plus over all variable blocks (JplPoseParameterization
// for vertices, identity for landmarks).
void plusState(const Problem& pb, const std::vector<double>& delta, const std::vector<double>& vin,
               const std::vector<double>& lin_, std::vector<double>* vout, std::vector<double>* lout) {
  *vout = vin;
  *lout = lin_;
  const int np = (int)pb.var_pose_list.size();
  for (int j = 0; j < np; ++j) {
    const int v = pb.var_pose_list[j];
    quatPlus(&vin[7 * v], &delta[6 * j], &(*vout)[7 * v]);
    for (int c = 0; c < 3; ++c) (*vout)[7 * v + 4 + c] = vin[7 * v + 4 + c] + delta[6 * j + 3 + c];
  }
  for (size_t j = 0; j < pb.var_lm_list.size(); ++j) {
    const int l = pb.var_lm_list[j];
    for (int c = 0; c < 3; ++c) (*lout)[3 * l + c] = lin_[3 * l + c] + delta[6 * np + 3 * j + c];
  }
}

double stateNorm(const Problem& pb, const std::vector<double>& v, const std::vector<double>& l) {
  double s = 0;
  for (int p : pb.var_pose_list) for (int c = 0; c < 7; ++c) s += v[7 * p + c] * v[7 * p + c];
  for (int m : pb.var_lm_list) for (int c = 0; c < 3; ++c) s += l[3 * m + c] * l[3 * m + c];
  return std::sqrt(s);
}

double stepAmbientNorm(const Problem& pb, const std::vector<double>& v0, const std::vector<double>& l0,
                       const std::vector<double>& v1, const std::vector<double>& l1) {
  double s = 0;
  for (int p : pb.var_pose_list) for (int c = 0; c < 7; ++c) { const double t = v1[7 * p + c] - v0[7 * p + c]; s += t * t; }
  for (int m : pb.var_lm_list) for (int c = 0; c < 3; ++c) { const double t = l1[3 * m + c] - l0[3 * m + c]; s += t * t; }
  return std::sqrt(s);
}

// ceres: gradient_max_norm = max |x - Plus(x, -g)|.
double projectedGradientMaxNorm(const Problem& pb, const std::vector<double>& g) {
  std::vector<double> neg(g.size());
  for (size_t i = 0; i < g.size(); ++i) neg[i] = -g[i];
  std::vector<double> v1, l1;
  plusState(pb, neg, pb.vertices, pb.landmarks, &v1, &l1);
  double m = 0;
  for (int p : pb.var_pose_list) for (int c = 0; c < 7; ++c) m = std::max(m, std::fabs(pb.vertices[7 * p + c] - v1[7 * p + c]));
  for (int q : pb.var_lm_list) for (int c = 0; c < 3; ++c) m = std::max(m, std::fabs(pb.landmarks[3 * q + c] - l1[3 * q + c]));
  return m;
}

struct LmState {
  double radius;
  double decrease_factor = 2.0;
};

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void pushIteration(mlba_summary* sm, const mlba_iteration_summary& it) {
  if (sm && sm->iterations && sm->num_iterations < sm->iterations_capacity) sm->iterations[sm->num_iterations] = it;
  if (sm) sm->num_iterations++;
}

// ceres: TrustRegionMinimizer::Minimize with LevenbergMarquardtStrategy
// (monotonic steps, no inner iterations).
int minimize(Problem& pb, const mlba_options& o, int solver, int nthreads, double initial_radius,
             mlba_summary* sm, double* last_radius) {
  const double t_start = now();
  LmState lm;
  lm.radius = initial_radius;
  const int np = (int)pb.var_pose_list.size();
  Linearization lin;
  linearize(pb, o, nthreads, true, &lin);
  std::vector<double> scale(pb.numTangent(), 1.0);
  if (o.use_jacobi_scaling) {
    columnSquaredNorms(pb, lin, &scale);
    for (double& s : scale) s = 1.0 / (1.0 + std::sqrt(s));
    scaleColumns(pb, scale, &lin);
  }
  double x_cost = lin.cost;
  const double fixed_cost = lin.fixed_cost;
  if (sm) { sm->initial_cost = x_cost + fixed_cost; sm->fixed_cost = fixed_cost; }
  double x_norm = stateNorm(pb, pb.vertices, pb.landmarks);
  mlba_iteration_summary it{};
  it.iteration = 0;
  it.step_is_valid = 1;
  it.step_is_successful = 1;
  it.cost = x_cost + fixed_cost;
  it.gradient_max_norm = projectedGradientMaxNorm(pb, lin.gradient);
  it.trust_region_radius = lm.radius;
  it.cumulative_time_in_seconds = now() - t_start;
  it.iteration_time_in_seconds = it.cumulative_time_in_seconds;
  pushIteration(sm, it);
  int termination = MLBA_NO_CONVERGENCE;
  if (it.gradient_max_norm <= o.gradient_tolerance) { *last_radius = lm.radius; return MLBA_CONVERGENCE; }
  if (o.max_num_iterations <= 0) { *last_radius = lm.radius; return MLBA_NO_CONVERGENCE; }

  std::vector<double> diag;
  bool reuse_diagonal = false;
  int consecutive_invalid = 0;
  int iteration = 0;
  double prev_grad_max = it.gradient_max_norm;
  while (true) {
    const double t_iter = now();
    ++iteration;
    mlba_iteration_summary cur{};
    cur.iteration = iteration;
    // ceres: LevenbergMarquardtStrategy::ComputeStep
    if (!reuse_diagonal) {
      columnSquaredNorms(pb, lin, &diag);  // of the scaled Jacobian
      for (double& v : diag) v = std::min(std::max(v, o.min_lm_diagonal), o.max_lm_diagonal);
    }
    std::vector<double> D(diag.size());
    for (size_t i = 0; i < diag.size(); ++i) D[i] = std::sqrt(diag[i] / lm.radius);
    std::vector<double> y;
    const double t_ls = now();
    const int ls_iters = solveNormalEquations(pb, lin, D, solver, nthreads, o.pcg_relative_tolerance,
                                              o.pcg_max_iterations, &y);
    if (sm) sm->linear_solver_seconds += now() - t_ls;
    reuse_diagonal = true;
    bool step_valid = ls_iters >= 0;
    for (double v : y) if (!std::isfinite(v)) step_valid = false;
    cur.linear_solver_iterations = std::max(ls_iters, 0);
    if (sm && ls_iters > 0) sm->total_linear_solver_iterations += ls_iters;
    // step = -y; model_cost_change = -(J step)^T (r + J step / 2)
    double model_cost_change = 0.0;
    if (step_valid) {
      const int K = (int)pb.active_obs.size();
      double mcc = 0.0;
      for (int k = 0; k < K; ++k) {
        double js[2] = {0, 0};
        for (int rr = 0; rr < 2; ++rr) {
          if (lin.il[k] >= 0) for (int c = 0; c < 3; ++c) js[rr] -= lin.Jl[6 * k + 3 * rr + c] * y[6 * np + 3 * lin.il[k] + c];
          if (lin.io[k] >= 0) for (int c = 0; c < 6; ++c) js[rr] -= lin.Jo[12 * k + 6 * rr + c] * y[6 * lin.io[k] + c];
          if (lin.ia[k] >= 0) for (int c = 0; c < 6; ++c) js[rr] -= lin.Ja[12 * k + 6 * rr + c] * y[6 * lin.ia[k] + c];
        }
        mcc -= js[0] * (lin.r[2 * k] + 0.5 * js[0]) + js[1] * (lin.r[2 * k + 1] + 0.5 * js[1]);
      }
      model_cost_change = mcc;
      step_valid = model_cost_change > 0.0;
    }
    cur.step_is_valid = step_valid;
    if (!step_valid) {
      // ceres: HandleInvalidStep
      if (++consecutive_invalid >= o.max_num_consecutive_invalid_steps) {
        termination = MLBA_FAILURE;
        break;
      }
      lm.radius /= lm.decrease_factor;
      lm.decrease_factor *= 2.0;
      cur.cost = x_cost + fixed_cost;
      cur.gradient_max_norm = prev_grad_max;
      cur.trust_region_radius = lm.radius;
      cur.iteration_time_in_seconds = now() - t_iter;
      cur.cumulative_time_in_seconds = now() - t_start;
      pushIteration(sm, cur);
      if (sm) sm->num_unsuccessful_steps++;
    } else {
      consecutive_invalid = 0;
      // delta = step * scale (undo Jacobi scaling); step = -y
      std::vector<double> delta(y.size());
      for (size_t i = 0; i < y.size(); ++i) delta[i] = -y[i] * scale[i];
      std::vector<double> cv, cl;
      plusState(pb, delta, pb.vertices, pb.landmarks, &cv, &cl);
      std::swap(pb.vertices, cv);
      std::swap(pb.landmarks, cl);
      Linearization cand;
      const double t_ev = now();
      linearize(pb, o, nthreads, false, &cand);
      if (sm) sm->evaluate_seconds += now() - t_ev;
      std::swap(pb.vertices, cv);  // restore x
      std::swap(pb.landmarks, cl);
      const double candidate_cost = cand.cost;
      cur.step_norm = stepAmbientNorm(pb, pb.vertices, pb.landmarks, cv, cl);
      // ceres: ParameterToleranceReached
      const double step_size_norm = vecNorm(delta);
      if (step_size_norm <= o.parameter_tolerance * (x_norm + o.parameter_tolerance)) {
        // ceres returns here without recording the iteration.
        termination = MLBA_CONVERGENCE;
        break;
      }
      cur.cost_change = x_cost - candidate_cost;
      // ceres: FunctionToleranceReached
      if (std::fabs(cur.cost_change) <= o.function_tolerance * x_cost) {
        termination = MLBA_CONVERGENCE;
        break;
      }
      cur.relative_decrease = (x_cost - candidate_cost) / model_cost_change;
      if (cur.relative_decrease > o.min_relative_decrease) {
        // ceres: HandleSuccessfulStep
        std::swap(pb.vertices, cv);
        std::swap(pb.landmarks, cl);
        x_norm = stateNorm(pb, pb.vertices, pb.landmarks);
        const double t_ev2 = now();
        linearize(pb, o, nthreads, true, &lin);
        if (sm) sm->evaluate_seconds += now() - t_ev2;
        if (o.use_jacobi_scaling) scaleColumns(pb, scale, &lin);
        x_cost = lin.cost;
        cur.step_is_successful = 1;
        cur.cost = x_cost + fixed_cost;
        cur.gradient_max_norm = projectedGradientMaxNorm(pb, lin.gradient);
        prev_grad_max = cur.gradient_max_norm;
        const double q = 2.0 * cur.relative_decrease - 1.0;
        lm.radius = lm.radius / std::max(1.0 / 3.0, 1.0 - q * q * q);
        lm.radius = std::min(o.max_trust_region_radius, lm.radius);
        lm.decrease_factor = 2.0;
        reuse_diagonal = false;
        if (sm) sm->num_successful_steps++;
      } else {
        // ceres: HandleUnsuccessfulStep
        lm.radius /= lm.decrease_factor;
        lm.decrease_factor *= 2.0;
        cur.cost = candidate_cost + fixed_cost;
        cur.gradient_max_norm = prev_grad_max;
        if (sm) sm->num_unsuccessful_steps++;
      }
      cur.trust_region_radius = lm.radius;
      cur.iteration_time_in_seconds = now() - t_iter;
      cur.cumulative_time_in_seconds = now() - t_start;
      pushIteration(sm, cur);
    }
    // ceres: FinalizeIterationAndCheckIfMinimizerCanContinue
    if (now() - t_start > o.max_solver_time_in_seconds) { termination = MLBA_NO_CONVERGENCE; break; }
    if (iteration >= o.max_num_iterations) { termination = MLBA_NO_CONVERGENCE; break; }
    if (cur.step_is_successful && cur.gradient_max_norm <= o.gradient_tolerance) { termination = MLBA_CONVERGENCE; break; }
    if (lm.radius <= o.min_trust_region_radius) { termination = MLBA_CONVERGENCE; break; }
  }
  if (sm) {
    sm->final_cost = x_cost + fixed_cost;
    sm->lm_loop_seconds += now() - t_start;
  }
  *last_radius = lm.radius;
  return termination;
}

bool validateDesc(const mlba_problem_desc* d) {
  if (!d || d->num_vertices < 0 || d->num_landmarks < 0 || d->num_observations < 0) return false;
  for (int i = 0; i < d->num_observations; ++i) {
    if (d->obs_landmark[i] < 0 || d->obs_landmark[i] >= d->num_landmarks) return false;
    if (d->obs_vertex[i] < 0 || d->obs_vertex[i] >= d->num_vertices) return false;
    if (d->obs_camera[i] < 0 || d->obs_camera[i] >= d->num_cameras) return false;
    if (!(d->obs_sigma[i] > 0.0)) return false;
  }
  return true;
}

// outlier-rejection-solver.cc:28-103 (findOutlierLandmarks): a landmark is an
// outlier when any of its observations has p_C.z <= 0, or (optionally) a
// squared pixel reprojection error above the threshold.
std::vector<int> findOutlierLandmarks(const Problem& pb, const mlba_options& o) {
  const mlba_problem_desc* d = pb.d;
  std::vector<uint8_t> bad(pb.L, 0);
  const double max_sq = o.max_reprojection_error_px * o.max_reprojection_error_px;
  for (int i = 0; i < pb.N; ++i) {
    const int l = d->obs_landmark[i];
    if (!pb.lm_active[l] || bad[l]) continue;
    const double* params[9];
    pb.params(i, pb.vertices.data(), pb.landmarks.data(), params);
    // p_C from the same frame chain as the error term.
    double r[2];
    const int c = d->obs_camera[i];
    // Reuse evaluateVisual's chain through a projection with a huge image so
    // that only p_C matters: recompute p_C explicitly instead.
    double R_C_I[9], R_B_LM[9], R_LM_B[9], R_I_M[9], p_I[3], p_C[3];
    rotJPL(params[5], R_C_I);
    const int type = pb.obs_type[i];
    if (type == MLBA_LOCAL_KEYFRAME) {
      for (int a = 0; a < 3; ++a) p_I[a] = params[0][a];
    } else {
      rotJPL(params[1], R_B_LM);
      transpose33(R_B_LM, R_LM_B);
      rotJPL(params[4], R_I_M);
      double p_M[3], t[3];
      matmul(R_LM_B, params[0], p_M, 3, 3, 1);
      for (int a = 0; a < 3; ++a) p_M[a] += params[1][4 + a];
      if (type == MLBA_GLOBAL) {
        double R_G_LM[9], R_G_M[9], R_M_G[9], p_G[3];
        rotJPL(params[2], R_G_LM);
        rotJPL(params[3], R_G_M);
        transpose33(R_G_M, R_M_G);
        matmul(R_G_LM, p_M, p_G, 3, 3, 1);
        for (int a = 0; a < 3; ++a) t[a] = p_G[a] + params[2][4 + a] - params[3][4 + a];
        matmul(R_M_G, t, p_M, 3, 3, 1);
      }
      for (int a = 0; a < 3; ++a) t[a] = p_M[a] - params[4][4 + a];
      matmul(R_I_M, t, p_I, 3, 3, 1);
    }
    matmul(R_C_I, p_I, p_C, 3, 3, 1);
    for (int a = 0; a < 3; ++a) p_C[a] += params[6][a];
    (void)r;
    if (p_C[2] <= 0.0) { bad[l] = 1; continue; }
    if (o.use_reprojection_error) {
      // landmark-quality-metrics.cc:57-73 (computeSquaredReprojectionError)
      double kp[2];
      const int st = project3Functional(d->camera_distortion_type[c], params[7], params[8],
                                        d->camera_resolution[2 * c], d->camera_resolution[2 * c + 1],
                                        p_C, kp, nullptr, nullptr, nullptr, camModel(d, c));
      double e2 = std::numeric_limits<double>::max();
      if (st == kVisible || st == kOutside) {
        const double du = kp[0] - d->obs_keypoint[2 * i], dv = kp[1] - d->obs_keypoint[2 * i + 1];
        e2 = du * du + dv * dv;
      }
      if (e2 > max_sq) bad[l] = 1;
    }
  }
  std::vector<int> out;
  for (int l = 0; l < pb.L; ++l) if (bad[l]) out.push_back(l);
  return out;
}

// ------------------------------------------------------------ landmark admission

// Rigid transform T (R, p): x -> R x + p (kindr / minkindr Transformation).
struct Tf {
  double R[9];
  double p[3];
};

Tf tfCompose(const Tf& a, const Tf& b) {  // a * b
  Tf c;
  matmul(a.R, b.R, c.R, 3, 3, 3);
  matmul(a.R, b.p, c.p, 3, 3, 1);
  for (int k = 0; k < 3; ++k) c.p[k] += a.p[k];
  return c;
}

Tf tfInverse(const Tf& a) {
  Tf c;
  transpose33(a.R, c.R);
  matmul(c.R, a.p, c.p, 3, 3, 1);
  for (int k = 0; k < 3; ++k) c.p[k] = -c.p[k];
  return c;
}

void tfApply(const Tf& a, const double* x, double* y) {
  matmul(a.R, x, y, 3, 3, 1);
  for (int k = 0; k < 3; ++k) y[k] += a.p[k];
}

// T_M_I of a vertex block [q_IM | M_p_MI] (rotJPL(q) = R_I_M), T_G_M of a
// baseframe block (rotJPL(q) = R_G_M), T_C_I of a camera block (rotJPL(q) = R_C_I)
// (optimization-state-buffer.cc:161-263).
Tf vertexT_M_I(const double* b) {
  Tf t;
  double R_I_M[9];
  rotJPL(b, R_I_M);
  transpose33(R_I_M, t.R);
  for (int k = 0; k < 3; ++k) t.p[k] = b[4 + k];
  return t;
}

Tf blockT(const double* b) {
  Tf t;
  rotJPL(b, t.R);
  for (int k = 0; k < 3; ++k) t.p[k] = b[4 + k];
  return t;
}

int resolveThreads(int n) {
  if (n > 0) return n;
  const unsigned hw = std::thread::hardware_concurrency();
  return std::max(1u, hw / 2u);
}

}  // namespace

// ====================================================================== C ABI

extern "C" {

int32_t oracle_project(int32_t dist_type, const double* intrinsics, const double* dist, int32_t width,
                       int32_t height, const double* p_C, double* kp, double* jp, double* jint,
                       double* jdist) {
  return project3Functional(dist_type, intrinsics, dist, width, height, p_C, kp, jp, jint, jdist);
}

int32_t oracle_project_camera(int32_t model, double xi, int32_t dist_type, const double* intrinsics,
                              const double* dist, int32_t width, int32_t height, const double* p_C, double* kp,
                              double* jp, double* jint, double* jdist, double* jxi) {
  CamModel cm;
  cm.model = model;
  cm.xi = xi;
  return project3Functional(dist_type, intrinsics, dist, width, height, p_C, kp, jp, jint, jdist, cm, jxi);
}

void oracle_distort(int32_t dist_type, const double* dist, double* point, double* jd) {
  distort(dist_type, dist, point, jd);
}

void oracle_quat_plus(const double* q, const double* delta, double* out) { quatPlus(q, delta, out); }
void oracle_quat_jacobian(const double* q, double* j43) { quatJacobian(q, j43); }
void oracle_rotation_jpl(const double* q, double* r33) { rotJPL(q, r33); }

int32_t oracle_evaluate_visual(const double* meas, double sigma, int32_t error_type, int32_t dist_type,
                               int32_t width, int32_t height, const double* const* params,
                               double* residual, double* const* jac) {
  return evaluateVisual(meas, 1.0 / sigma, error_type, dist_type, width, height, params, residual, jac) ? 1 : 0;
}

void oracle_loss(int32_t loss_type, double a, double s, double* rho) { lossEvaluate(loss_type, a, s, rho); }

int32_t oracle_problem_evaluate(const mlba_problem_desc* desc, const uint8_t* landmark_active, int32_t flags,
                                const mlba_options* options, double* residuals, double* const* jacobians,
                                double* cost) {
  if (!validateDesc(desc)) return MLBA_EINVAL;
  mlba_options o;
  if (options) o = *options; else { std::memset(&o, 0, sizeof(o)); o.loss_type = MLBA_LOSS_MAPLAB_HUBER; }
  Problem pb(desc, landmark_active);
  static const int amb[9] = {3, 7, 7, 7, 7, 4, 3, 4, 4};
  static const int tan[9] = {3, 6, 6, 6, 6, 3, 3, 4, 4};
  const int* sz = (flags & MLBA_EVAL_TANGENT) ? tan : amb;
  const bool with_xi = (flags & MLBA_EVAL_WITH_XI) != 0;
  double total = 0.0;
  for (int i = 0; i < pb.N; ++i) {
    const double* params[9];
    pb.params(i, pb.vertices.data(), pb.landmarks.data(), params);
    const int type = pb.obs_type[i];
    const int c = desc->obs_camera[i];
    double J[9][14];
    double* jp[9];
    for (int b = 0; b < 9; ++b) jp[b] = J[b];
    if (type == MLBA_LOCAL_KEYFRAME) jp[1] = jp[2] = jp[3] = jp[4] = nullptr;
    if (type == MLBA_LOCAL_MISSION) jp[2] = jp[3] = nullptr;
    double r[2], jxi[2];
    evaluateVisual(desc->obs_keypoint + 2 * i, 1.0 / desc->obs_sigma[i], type, desc->camera_distortion_type[c],
                   desc->camera_resolution[2 * c], desc->camera_resolution[2 * c + 1], params, r,
                   jacobians ? jp : nullptr, camModel(desc, c), (jacobians && with_xi) ? jxi : nullptr);
    if (jacobians && with_xi && jacobians[9]) { jacobians[9][2 * i] = jxi[0]; jacobians[9][2 * i + 1] = jxi[1]; }
    if (residuals) { residuals[2 * i] = r[0]; residuals[2 * i + 1] = r[1]; }
    const int l = desc->obs_landmark[i];
    if (pb.lm_active[l]) {
      double rho[3];
      lossEvaluate(o.loss_type, lossScale(o, type, desc->obs_sigma[i]), r[0] * r[0] + r[1] * r[1], rho);
      total += 0.5 * rho[0];
    }
    if (!jacobians) continue;
    for (int b = 0; b < 9; ++b) {
      if (!jacobians[b]) continue;
      double* out = jacobians[b] + (size_t)i * 2 * sz[b];
      if (!jp[b]) { for (int t = 0; t < 2 * sz[b]; ++t) out[t] = 0.0; continue; }
      if ((flags & MLBA_EVAL_TANGENT) && (b >= 1 && b <= 4)) {
        liftPose(J[b], params[b], out);
      } else if ((flags & MLBA_EVAL_TANGENT) && b == 5) {
        double Jq[12];
        quatJacobian(params[5], Jq);
        for (int rr = 0; rr < 2; ++rr)
          for (int cc = 0; cc < 3; ++cc) {
            double s = 0;
            for (int k = 0; k < 4; ++k) s += J[5][rr * 4 + k] * Jq[k * 3 + cc];
            out[rr * 3 + cc] = s;
          }
      } else {
        for (int t = 0; t < 2 * sz[b]; ++t) out[t] = J[b][t];
      }
    }
  }
  if (cost) *cost = total;
  return MLBA_OK;
}

int32_t oracle_solve(const mlba_problem_desc* desc, const mlba_options* options, int32_t linear_solver,
                     int32_t num_threads, const uint8_t* landmark_active, double* vertices_out,
                     double* landmarks_out, mlba_summary* summary) {
  if (!validateDesc(desc) || !options) return MLBA_EINVAL;
  Problem pb(desc, landmark_active);
  const int nthreads = resolveThreads(num_threads);
  if (summary) {
    summary->num_iterations = 0;
    summary->num_successful_steps = summary->num_unsuccessful_steps = 0;
    summary->lm_loop_seconds = summary->evaluate_seconds = summary->linear_solver_seconds = 0;
    summary->total_linear_solver_iterations = 0;
    summary->num_residual_blocks = (int)pb.active_obs.size();
    summary->num_removed_landmarks = 0;
  }
  double radius;
  const int term = minimize(pb, *options, linear_solver, nthreads, options->initial_trust_region_radius,
                            summary, &radius);
  if (summary) summary->termination_type = term;
  if (vertices_out) std::memcpy(vertices_out, pb.vertices.data(), sizeof(double) * 7 * pb.P);
  if (landmarks_out) std::memcpy(landmarks_out, pb.landmarks.data(), sizeof(double) * 3 * pb.L);
  return MLBA_OK;
}

// Calibration refinement with the map held constant: the camera blocks
// q_CI (JplQuaternionParameterization), C_p_CI, intrinsics (with xi for a
// unified-projection camera) and distortion of visual-error-term-inl.h:321-367
// are the only variables (fix flags: optimization-terms-addition.cc:222-234);
// ceres: LM on dense normal equations (DENSE_QR / DENSE_SCHUR give the same
// step here), with the rules of minimize() above. fix = {intrinsics,
// distortion, extrinsics rotation, extrinsics translation}.
int32_t oracle_solve_calibration(const mlba_problem_desc* desc, const mlba_options* options, const int32_t* fix,
                                 const uint8_t* landmark_active, double* cam_q_out, double* intrinsics_out,
                                 double* distortion_out, double* xi_out, mlba_summary* summary) {
  if (!validateDesc(desc) || !options || !fix) return MLBA_EINVAL;
  const mlba_options& o = *options;
  Problem pb(desc, landmark_active);
  const int C = desc->num_cameras;
  std::vector<double> cq(desc->camera_q_CI__C_p_CI, desc->camera_q_CI__C_p_CI + 7 * C);
  std::vector<double> ci(desc->camera_intrinsics, desc->camera_intrinsics + 4 * C);
  std::vector<double> cd(desc->camera_distortion, desc->camera_distortion + 4 * C);
  std::vector<double> cx(C, 0.0);
  for (int c = 0; c < C; ++c) cx[c] = camModel(desc, c).xi;
  // Every active residual's map blocks must be constant.
  std::vector<uint8_t> used(C, 0);
  for (int i : pb.active_obs) {
    const int l = desc->obs_landmark[i], v = desc->obs_vertex[i], st = desc->landmark_store_vertex[l];
    if (!pb.lmConst(l) || !pb.poseConst(v) || (pb.obs_type[i] != MLBA_LOCAL_KEYFRAME && !pb.poseConst(st)))
      return MLBA_EINVAL;
    used[desc->obs_camera[i]] = 1;
  }
  // Columns: per camera [rot 3 | trans 3 | fu fv cu cv | distortion | xi].
  std::vector<std::array<int, 15>> col(C);
  int n = 0;
  for (int c = 0; c < C; ++c) {
    col[c].fill(-1);
    if (!used[c]) continue;
    const int dt = desc->camera_distortion_type[c];
    const int nd = (dt == MLBA_DISTORTION_RADTAN || dt == MLBA_DISTORTION_EQUIDISTANT) ? 4
                   : (dt == MLBA_DISTORTION_FISHEYE ? 1 : 0);
    if (!fix[2]) for (int k = 0; k < 3; ++k) col[c][k] = n++;
    if (!fix[3]) for (int k = 3; k < 6; ++k) col[c][k] = n++;
    if (!fix[0]) for (int k = 6; k < 10; ++k) col[c][k] = n++;
    if (!fix[1]) for (int k = 0; k < nd; ++k) col[c][10 + k] = n++;
    if (!fix[0] && camModel(desc, c).model == MLBA_CAMERA_UNIFIED_PROJECTION) col[c][14] = n++;
  }
  if (n == 0) return MLBA_EINVAL;
  auto camFree = [&](int c) {
    for (int k = 0; k < 15; ++k) if (col[c][k] >= 0) return true;
    return false;
  };
  // Cost, fixed cost and (with A) J'J and J'r of the free columns.
  auto linearize = [&](const std::vector<double>& q, const std::vector<double>& in, const std::vector<double>& di,
                       const std::vector<double>& xi, double* cost, double* fixed, std::vector<double>* A,
                       std::vector<double>* g) {
    *cost = *fixed = 0.0;
    if (A) { A->assign((size_t)n * n, 0.0); g->assign(n, 0.0); }
    for (int i : pb.active_obs) {
      const int c = desc->obs_camera[i];
      const double* params[9];
      pb.params(i, pb.vertices.data(), pb.landmarks.data(), params);
      params[5] = &q[7 * c];
      params[6] = &q[7 * c + 4];
      params[7] = &in[4 * c];
      params[8] = &di[4 * c];
      CamModel cm = camModel(desc, c);
      cm.xi = xi[c];
      double J5[8], J6[6], J7[8], J8[8], Jx[2], r[2];
      double* jac[9] = {nullptr, nullptr, nullptr, nullptr, nullptr, J5, J6, J7, J8};
      const int type = pb.obs_type[i];
      const double sigma = desc->obs_sigma[i];
      evaluateVisual(desc->obs_keypoint + 2 * i, 1.0 / sigma, type, desc->camera_distortion_type[c],
                     desc->camera_resolution[2 * c], desc->camera_resolution[2 * c + 1], params, r,
                     A ? jac : nullptr, cm, A ? Jx : nullptr);
      const double sq = r[0] * r[0] + r[1] * r[1];
      double rho[3];
      lossEvaluate(o.loss_type, lossScale(o, type, sigma), sq, rho);
      (camFree(c) ? *cost : *fixed) += 0.5 * rho[0];
      if (!A || !camFree(c)) continue;
      // Tangent columns: q_CI through the JPL parameterization Jacobian.
      double Jq[12], Jt[30];
      quatJacobian(params[5], Jq);
      for (int k = 0; k < 2; ++k) {
        for (int a = 0; a < 3; ++a) {
          double v = 0.0;
          for (int m = 0; m < 4; ++m) v += J5[4 * k + m] * Jq[3 * m + a];
          Jt[15 * k + a] = v;
          Jt[15 * k + 3 + a] = J6[3 * k + a];
        }
        for (int a = 0; a < 4; ++a) { Jt[15 * k + 6 + a] = J7[4 * k + a]; Jt[15 * k + 10 + a] = J8[4 * k + a]; }
        Jt[15 * k + 14] = Jx[k];
      }
      Corrector corr(sq, rho);
      corr.correctJacobian(r, Jt, 15);
      const double rr[2] = {r[0] * corr.residual_scaling, r[1] * corr.residual_scaling};
      for (int a = 0; a < 15; ++a) {
        const int ga = col[c][a];
        if (ga < 0) continue;
        (*g)[ga] += Jt[a] * rr[0] + Jt[15 + a] * rr[1];
        for (int b = 0; b < 15; ++b) {
          const int gb = col[c][b];
          if (gb >= 0) (*A)[(size_t)ga * n + gb] += Jt[a] * Jt[b] + Jt[15 + a] * Jt[15 + b];
        }
      }
    }
  };
  auto gradMax = [&](const std::vector<double>& q, const std::vector<double>& g) {
    double m = 0.0;
    for (int c = 0; c < C; ++c)
      for (int k = 0; k < 15; ++k) {
        const int j = col[c][k];
        if (j < 0) continue;
        if (k == 0) {
          const double d[3] = {-g[j], -g[col[c][1]], -g[col[c][2]]};
          double qq[4];
          quatPlus(&q[7 * c], d, qq);
          for (int a = 0; a < 4; ++a) m = std::max(m, std::fabs(q[7 * c + a] - qq[a]));
        } else if (k >= 3) {
          m = std::max(m, std::fabs(g[j]));
        }
      }
    return m;
  };
  if (summary) {
    summary->num_iterations = 0;
    summary->num_successful_steps = summary->num_unsuccessful_steps = 0;
    summary->total_linear_solver_iterations = 0;
    summary->num_residual_blocks = (int)pb.active_obs.size();
    summary->num_removed_landmarks = 0;
  }
  auto push = [&](const mlba_iteration_summary& it) {
    if (!summary) return;
    if (summary->iterations && summary->num_iterations < summary->iterations_capacity)
      summary->iterations[summary->num_iterations] = it;
    summary->num_iterations++;
  };
  std::vector<double> A, g;
  double cost, fixed;
  linearize(cq, ci, cd, cx, &cost, &fixed, &A, &g);
  double gmax = gradMax(cq, g);
  if (summary) { summary->initial_cost = cost + fixed; summary->fixed_cost = fixed; }
  mlba_iteration_summary it0{};
  it0.step_is_valid = it0.step_is_successful = 1;
  it0.cost = cost + fixed;
  it0.gradient_max_norm = gmax;
  double radius = o.initial_trust_region_radius, decrease = 2.0;
  it0.trust_region_radius = radius;
  push(it0);
  int term = gmax <= o.gradient_tolerance ? MLBA_CONVERGENCE : MLBA_NO_CONVERGENCE;
  int invalid = 0, iteration = 0;
  while (term == MLBA_NO_CONVERGENCE && o.max_num_iterations > 0) {
    ++iteration;
    mlba_iteration_summary cur{};
    cur.iteration = iteration;
    // ceres: Jacobi scaling 1 / (1 + |column|), LM diagonal clamp(diag) / radius.
    std::vector<double> sc(n), M((size_t)n * n), b(n), w(n);
    for (int j = 0; j < n; ++j) sc[j] = o.use_jacobi_scaling ? 1.0 / (1.0 + std::sqrt(A[(size_t)j * n + j])) : 1.0;
    for (int a = 0; a < n; ++a) {
      b[a] = sc[a] * g[a];
      for (int c2 = 0; c2 < n; ++c2) M[(size_t)a * n + c2] = sc[a] * A[(size_t)a * n + c2] * sc[c2];
    }
    for (int j = 0; j < n; ++j)
      M[(size_t)j * n + j] += std::min(std::max(M[(size_t)j * n + j], o.min_lm_diagonal), o.max_lm_diagonal) / radius;
    // Dense LDL'-free Cholesky solve.
    bool ok = true;
    std::vector<double> L(M);
    for (int j = 0; j < n && ok; ++j) {
      double dj = L[(size_t)j * n + j];
      for (int k = 0; k < j; ++k) dj -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
      if (!(dj > 0.0)) { ok = false; break; }
      const double lj = std::sqrt(dj);
      L[(size_t)j * n + j] = lj;
      for (int i2 = j + 1; i2 < n; ++i2) {
        double v = L[(size_t)i2 * n + j];
        for (int k = 0; k < j; ++k) v -= L[(size_t)i2 * n + k] * L[(size_t)j * n + k];
        L[(size_t)i2 * n + j] = v / lj;
      }
    }
    double mcc = 0.0;
    if (ok) {
      for (int i2 = 0; i2 < n; ++i2) {
        double v = b[i2];
        for (int k = 0; k < i2; ++k) v -= L[(size_t)i2 * n + k] * w[k];
        w[i2] = v / L[(size_t)i2 * n + i2];
      }
      for (int i2 = n - 1; i2 >= 0; --i2) {
        double v = w[i2];
        for (int k = i2 + 1; k < n; ++k) v -= L[(size_t)k * n + i2] * w[k];
        w[i2] = v / L[(size_t)i2 * n + i2];
      }
      for (int j = 0; j < n; ++j) w[j] *= sc[j];
      // ceres: model cost change of the step -w, -(J d)'(r + J d / 2).
      double gw = 0.0, wAw = 0.0;
      for (int a = 0; a < n; ++a) {
        gw += g[a] * w[a];
        for (int c2 = 0; c2 < n; ++c2) wAw += w[a] * A[(size_t)a * n + c2] * w[c2];
      }
      mcc = gw - 0.5 * wAw;
      ok = std::isfinite(mcc) && mcc > 0.0;
    }
    cur.step_is_valid = ok;
    cur.linear_solver_iterations = 1;
    if (!ok) {
      if (++invalid >= o.max_num_consecutive_invalid_steps) { term = MLBA_FAILURE; break; }
      radius /= decrease;
      decrease *= 2.0;
      cur.cost = cost + fixed;
      cur.gradient_max_norm = gmax;
      cur.trust_region_radius = radius;
      if (summary) summary->num_unsuccessful_steps++;
    } else {
      invalid = 0;
      // Candidate x [+] (-w) and the ambient norm of the variable blocks.
      std::vector<double> q2(cq), in2(ci), di2(cd), xi2(cx);
      double xx = 0.0, ss = 0.0;
      for (int j = 0; j < n; ++j) ss += w[j] * w[j];
      for (int c = 0; c < C; ++c) {
        const auto& cc = col[c];
        if (cc[0] >= 0) {
          const double d[3] = {-w[cc[0]], -w[cc[1]], -w[cc[2]]};
          quatPlus(&cq[7 * c], d, &q2[7 * c]);
          for (int a = 0; a < 4; ++a) xx += cq[7 * c + a] * cq[7 * c + a];
        }
        if (cc[3] >= 0)
          for (int a = 0; a < 3; ++a) { q2[7 * c + 4 + a] -= w[cc[3 + a]]; xx += cq[7 * c + 4 + a] * cq[7 * c + 4 + a]; }
        if (cc[6] >= 0)
          for (int a = 0; a < 4; ++a) { in2[4 * c + a] -= w[cc[6 + a]]; xx += ci[4 * c + a] * ci[4 * c + a]; }
        for (int a = 0; a < 4; ++a)
          if (cc[10 + a] >= 0) { di2[4 * c + a] -= w[cc[10 + a]]; xx += cd[4 * c + a] * cd[4 * c + a]; }
        if (cc[14] >= 0) { xi2[c] -= w[cc[14]]; xx += cx[c] * cx[c]; }
      }
      const double step_norm = std::sqrt(ss), x_norm = std::sqrt(xx);
      cur.step_norm = step_norm;
      if (step_norm <= o.parameter_tolerance * (x_norm + o.parameter_tolerance)) { term = MLBA_CONVERGENCE; break; }
      double ccost, cfixed;
      linearize(q2, in2, di2, xi2, &ccost, &cfixed, nullptr, nullptr);
      cur.cost_change = cost - ccost;
      if (std::fabs(cur.cost_change) <= o.function_tolerance * cost) { term = MLBA_CONVERGENCE; break; }
      cur.relative_decrease = (cost - ccost) / mcc;
      if (cur.relative_decrease > o.min_relative_decrease) {
        cq = q2; ci = in2; cd = di2; cx = xi2;
        linearize(cq, ci, cd, cx, &cost, &fixed, &A, &g);
        gmax = gradMax(cq, g);
        cur.step_is_successful = 1;
        cur.cost = cost + fixed;
        cur.gradient_max_norm = gmax;
        const double qq = 2.0 * cur.relative_decrease - 1.0;
        radius = std::min(o.max_trust_region_radius, radius / std::max(1.0 / 3.0, 1.0 - qq * qq * qq));
        decrease = 2.0;
        if (summary) summary->num_successful_steps++;
      } else {
        radius /= decrease;
        decrease *= 2.0;
        cur.cost = ccost + fixed;
        cur.gradient_max_norm = gmax;
        if (summary) summary->num_unsuccessful_steps++;
      }
      cur.trust_region_radius = radius;
    }
    push(cur);
    if (iteration >= o.max_num_iterations) break;
    if (cur.step_is_successful && cur.gradient_max_norm <= o.gradient_tolerance) { term = MLBA_CONVERGENCE; break; }
    if (radius <= o.min_trust_region_radius) { term = MLBA_CONVERGENCE; break; }
  }
  if (summary) {
    summary->termination_type = term;
    summary->final_cost = cost + fixed;
  }
  if (cam_q_out) std::memcpy(cam_q_out, cq.data(), sizeof(double) * 7 * C);
  if (intrinsics_out) std::memcpy(intrinsics_out, ci.data(), sizeof(double) * 4 * C);
  if (distortion_out) std::memcpy(distortion_out, cd.data(), sizeof(double) * 4 * C);
  if (xi_out) std::memcpy(xi_out, cx.data(), sizeof(double) * C);
  return MLBA_OK;
}

int32_t oracle_solve_with_outlier_rejection(const mlba_problem_desc* desc, const mlba_options* options,
                                            int32_t linear_solver, int32_t num_threads, double* vertices_out,
                                            double* landmarks_out, uint8_t* landmark_active_out,
                                            mlba_summary* summary) {
  if (!validateDesc(desc) || !options) return MLBA_EINVAL;
  if (options->reject_outliers_every_n_iters <= 0 || options->max_num_iterations <= 0) {
    if (summary) summary->termination_type = MLBA_FAILURE;
    return MLBA_OK;  // outlier-rejection-solver.cc:214-218
  }
  Problem pb(desc, nullptr);
  const int nthreads = resolveThreads(num_threads);
  if (summary) {
    summary->num_iterations = 0;
    summary->num_successful_steps = summary->num_unsuccessful_steps = 0;
    summary->lm_loop_seconds = summary->evaluate_seconds = summary->linear_solver_seconds = 0;
    summary->total_linear_solver_iterations = 0;
    summary->num_removed_landmarks = 0;
  }
  int remaining = options->max_num_iterations;
  double radius = options->initial_trust_region_radius;
  int term = MLBA_NO_CONVERGENCE;
  double budget = options->max_solver_time_in_seconds;
  bool first = true;
  double initial_cost = 0.0;
  while (remaining > 0 && budget > 0) {
    const double t0 = now();
    const int step_iters = std::min(remaining, options->reject_outliers_every_n_iters);
    mlba_options local = *options;
    local.max_num_iterations = step_iters;
    term = minimize(pb, local, linear_solver, nthreads, radius, summary, &radius);
    if (first && summary) initial_cost = summary->initial_cost;
    first = false;
    // outlier-rejection-solver.cc:142-182
    const std::vector<int> outliers = findOutlierLandmarks(pb, *options);
    for (int l : outliers) pb.lm_active[l] = 0;
    if (summary) summary->num_removed_landmarks += (int)outliers.size();
    pb.rebuildIndex();
    if (term != MLBA_NO_CONVERGENCE) break;
    budget -= now() - t0;
    remaining -= step_iters;
  }
  if (summary) {
    summary->termination_type = term;
    summary->initial_cost = initial_cost;
    summary->num_residual_blocks = (int)pb.active_obs.size();
  }
  if (vertices_out) std::memcpy(vertices_out, pb.vertices.data(), sizeof(double) * 7 * pb.P);
  if (landmarks_out) std::memcpy(landmarks_out, pb.landmarks.data(), sizeof(double) * 3 * pb.L);
  if (landmark_active_out) std::memcpy(landmark_active_out, pb.lm_active.data(), pb.L);
  return MLBA_OK;
}

int32_t oracle_landmark_quality(const mlba_problem_desc* d, const mlba_quality_options* o, uint8_t* good) {
  if (!d || !o || !good) return MLBA_EINVAL;
  const int L = d->num_landmarks, N = d->num_observations;
  std::vector<std::vector<int>> obs(L);
  for (int i = 0; i < N; ++i) obs[d->obs_landmark[i]].push_back(i);
  const double max_e2 = o->max_reprojection_error_px * o->max_reprojection_error_px;
  for (int l = 0; l < L; ++l) {
    good[l] = 0;
    // landmark-quality-metrics.cc:105-108
    if ((int)obs[l].size() < o->min_observers) continue;
    // map.getLandmark_G_p_fi: T_G_M (store mission) * T_M_I (store vertex) * p_B.
    const int a = d->landmark_store_vertex[l];
    const Tf T_G_Ia = tfCompose(blockT(d->baseframe_q_GM__G_p_GM + 7 * d->vertex_mission[a]),
                                vertexT_M_I(d->vertex_q_IM__M_p_MI + 7 * a));
    double p_G_fi[3];
    tfApply(T_G_Ia, d->landmark_p_B + 3 * l, p_G_fi);
    std::vector<std::array<double, 3>> rays;
    double signed_min = std::numeric_limits<double>::max();
    bool reproj_bad = false;
    for (int i : obs[l]) {
      const int v = d->obs_vertex[i], c = d->obs_camera[i];
      // T_G_C = T_G_I * T_C_B^-1 (:118-121); p_C_fi = T_C_B * T_G_I^-1 * p_G_fi (:123-124).
      const Tf T_G_I = tfCompose(blockT(d->baseframe_q_GM__G_p_GM + 7 * d->vertex_mission[v]),
                                 vertexT_M_I(d->vertex_q_IM__M_p_MI + 7 * v));
      const Tf T_C_B = blockT(d->camera_q_CI__C_p_CI + 7 * c);
      const Tf T_G_C = tfCompose(T_G_I, tfInverse(T_C_B));
      double p_C[3];
      tfApply(tfCompose(T_C_B, tfInverse(T_G_I)), p_G_fi, p_C);
      double ray[3];
      for (int k = 0; k < 3; ++k) ray[k] = T_G_C.p[k] - p_G_fi[k];
      const double dist = std::sqrt(ray[0] * ray[0] + ray[1] * ray[1] + ray[2] * ray[2]);
      signed_min = std::min(signed_min, dist * (p_C[2] < 0.0 ? -1.0 : 1.0));
      if (dist > 0) rays.push_back({ray[0] / dist, ray[1] / dist, ray[2] / dist});
      if (o->max_reprojection_error_px > 0.0) {
        // computeSquaredReprojectionError (:57-73)
        double kp[2];
        const int st = project3Functional(d->camera_distortion_type[c], d->camera_intrinsics + 4 * c,
                                          d->camera_distortion + 4 * c, d->camera_resolution[2 * c],
                                          d->camera_resolution[2 * c + 1], p_C, kp, nullptr, nullptr, nullptr,
                                          camModel(d, c));
        double e2 = std::numeric_limits<double>::max();
        if (st == kVisible || st == kOutside) {
          const double du = kp[0] - d->obs_keypoint[2 * i], dv = kp[1] - d->obs_keypoint[2 * i + 1];
          e2 = du * du + dv * dv;
        }
        if (e2 > max_e2) reproj_bad = true;
      }
    }
    if (reproj_bad) continue;
    if (signed_min > o->max_distance_from_closest_observer || signed_min < o->min_distance_from_closest_observer)
      continue;
    // common::getMaxDisparityRadAngleOfUnitVectorBundle (geometry.cc:30-44)
    double angle = 0.0;
    if (rays.size() >= 2) {
      double min_cos = 1.0;
      for (size_t x = 0; x < rays.size(); ++x)
        for (size_t y = x + 1; y < rays.size(); ++y)
          min_cos = std::min(min_cos, std::abs(rays[x][0] * rays[y][0] + rays[x][1] * rays[y][1] +
                                               rays[x][2] * rays[y][2]));
      angle = std::acos(min_cos);
    }
    if (angle * (180.0 / M_PI) < o->min_observation_angle_deg) continue;
    good[l] = 1;
  }
  return MLBA_OK;
}

int32_t oracle_reduced_camera_system(const mlba_problem_desc* desc, const mlba_options* options, double mu,
                                     int32_t capacity, double* S, double* rhs, int32_t* vertex_of_block) {
  if (!validateDesc(desc) || !options) return MLBA_EINVAL;
  Problem pb(desc, nullptr);
  Linearization lin;
  linearize(pb, *options, 1, true, &lin);
  std::vector<double> scale(pb.numTangent(), 1.0);
  if (options->use_jacobi_scaling) {
    columnSquaredNorms(pb, lin, &scale);
    for (double& s : scale) s = 1.0 / (1.0 + std::sqrt(s));
    scaleColumns(pb, scale, &lin);
  }
  std::vector<double> diag;
  columnSquaredNorms(pb, lin, &diag);
  std::vector<double> D(diag.size());
  for (size_t i = 0; i < diag.size(); ++i)
    D[i] = std::sqrt(std::min(std::max(diag[i], options->min_lm_diagonal), options->max_lm_diagonal) / mu);
  std::vector<double> y, Sd, rd;
  const int nf = 6 * (int)pb.var_pose_list.size();
  if (nf * nf > capacity) return MLBA_EINVAL;
  if (solveNormalEquations(pb, lin, D, ORACLE_SCHUR_DIRECT, 1, 0.0, 0, &y, &Sd, &rd) < 0) return MLBA_ESTATE;
  std::memcpy(S, Sd.data(), sizeof(double) * nf * nf);
  std::memcpy(rhs, rd.data(), sizeof(double) * nf);
  for (size_t j = 0; j < pb.var_pose_list.size(); ++j) vertex_of_block[j] = pb.var_pose_list[j];
  return nf;
}

}  // extern "C"
