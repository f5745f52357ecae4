// ba_math.h -- fixed-size FP64 geometry for the visual reprojection error,
// shared by the gfx950 kernels and the host generator.
//
// Conventions follow maplab:
//  * quaternions are stored [x y z w] (JPL), w >= 0; the rotation matrix of
//    the coefficients is toRotationMatrixJPL
//    (common/maplab-common/include/maplab-common/quaternion-math-inl.h:7-27);
//  * JPL Plus (algorithms/ceres-error-terms/src/parameterization/
//    quaternion-param-jpl.cc:35-68) left-multiplies the small rotation, so
//    its tangent derivative of R(q) p is [R(q) p]x (see below);
//  * pinhole projection with radtan / equidistant / fisheye (FOV) distortion and the
//    projection status of aslam_cv2/aslam_cv_cameras/src/camera-pinhole.cc:65-164.
//
// Unlike the reference, Jacobians are produced directly in the tangent space
// that the solver works in: for q' = Plus(q, d), R(q') ~ (I - [d]x) R(q) with
// |q|^2 factors that the reference gets through 4 Jq^T Jq = |q|^2 I.
#ifndef MAPLAB_AMD_BA_MATH_H_
#define MAPLAB_AMD_BA_MATH_H_

#include <math.h>
#include <stdint.h>

#ifndef MLBA_HD
#if defined(__HIPCC__)
#define MLBA_HD __host__ __device__ inline
#else
#define MLBA_HD inline
#endif
#endif

namespace mlba {

enum : int32_t { kDistNone = 0, kDistRadtan = 1, kDistEquidistant = 2, kDistFisheye = 3 };
enum : int32_t { kCameraPinhole = 0, kCameraUnifiedProjection = 1 };
enum : int32_t { kTypeKeyframe = 0, kTypeMission = 1, kTypeGlobal = 2 };
enum : int32_t { kVisible = 0, kOutsideImage = 1, kBehind = 2, kInvalid = 3 };

// R = toRotationMatrixJPL(q), row-major.
MLBA_HD void rot_jpl(const double* q, double* R) {
  const double x = q[0], y = q[1], z = q[2], w = q[3];
  const double xx = x * x, yy = y * y, zz = z * z;
  const double xy = x * y, xz = x * z, yz = y * z, xw = x * w, yw = y * w, zw = z * w;
  R[0] = 1.0 - 2.0 * (yy + zz); R[1] = 2.0 * (xy + zw);        R[2] = 2.0 * (xz - yw);
  R[3] = 2.0 * (xy - zw);       R[4] = 1.0 - 2.0 * (xx + zz);  R[5] = 2.0 * (yz + xw);
  R[6] = 2.0 * (xz + yw);       R[7] = 2.0 * (yz - xw);        R[8] = 1.0 - 2.0 * (xx + yy);
}

MLBA_HD void mat3_vec(const double* R, const double* v, double* out) {
  out[0] = R[0] * v[0] + R[1] * v[1] + R[2] * v[2];
  out[1] = R[3] * v[0] + R[4] * v[1] + R[5] * v[2];
  out[2] = R[6] * v[0] + R[7] * v[1] + R[8] * v[2];
}
MLBA_HD void mat3t_vec(const double* R, const double* v, double* out) {
  out[0] = R[0] * v[0] + R[3] * v[1] + R[6] * v[2];
  out[1] = R[1] * v[0] + R[4] * v[1] + R[7] * v[2];
  out[2] = R[2] * v[0] + R[5] * v[1] + R[8] * v[2];
}
// C = A * B (3x3).
MLBA_HD void mat3_mul(const double* A, const double* B, double* C) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      C[i * 3 + j] = A[i * 3] * B[j] + A[i * 3 + 1] * B[3 + j] + A[i * 3 + 2] * B[6 + j];
}
// C = A * B^T (3x3).
MLBA_HD void mat3_mul_bt(const double* A, const double* B, double* C) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      C[i * 3 + j] = A[i * 3] * B[j * 3] + A[i * 3 + 1] * B[j * 3 + 1] + A[i * 3 + 2] * B[j * 3 + 2];
}

// q' = JplQuaternionParameterization::Plus(q, d): small-angle branch below
// |d| < 0.524 rad, then the left product with positive scalar part.
MLBA_HD void quat_plus_jpl(const double* q, const double* d, double* out) {
  const double sq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  double a0, a1, a2, a3;
  if (sq < 0.524 * 0.524) {
    const double inv = 1.0 / sqrt(0.25 * sq + 1.0);
    a0 = 0.5 * d[0] * inv; a1 = 0.5 * d[1] * inv; a2 = 0.5 * d[2] * inv; a3 = inv;
  } else {
    const double h = 0.5 * sqrt(sq);
    const double s = sin(h) / (2.0 * h);
    a0 = d[0] * s; a1 = d[1] * s; a2 = d[2] * s; a3 = cos(h);
  }
  double p0 = a3 * q[0] + a2 * q[1] - a1 * q[2] + a0 * q[3];
  double p1 = -a2 * q[0] + a3 * q[1] + a0 * q[2] + a1 * q[3];
  double p2 = a1 * q[0] - a0 * q[1] + a3 * q[2] + a2 * q[3];
  double p3 = -a0 * q[0] - a1 * q[1] - a2 * q[2] + a3 * q[3];
  const double n = 1.0 / sqrt(p0 * p0 + p1 * p1 + p2 * p2 + p3 * p3);
  const double sgn = (p3 < 0.0) ? -n : n;
  out[0] = p0 * sgn; out[1] = p1 * sgn; out[2] = p2 * sgn; out[3] = p3 * sgn;
}

// Distortion of a normalized point; jd = d(distorted)/d(point) (2x2).
MLBA_HD void distort(int32_t type, const double* k, double& x, double& y, double* jd) {
  if (type == kDistRadtan) {
    const double k1 = k[0], k2 = k[1], p1 = k[2], p2 = k[3];
    const double xx = x * x, yy = y * y, xy = x * y, r2 = xx + yy;
    const double rad = r2 * (k1 + k2 * r2);
    const double dr = 2.0 * k1 + 4.0 * k2 * r2;  // d(rad)/d(r2) * 2
    jd[0] = 1.0 + rad + dr * xx + 2.0 * p1 * y + 6.0 * p2 * x;
    jd[1] = dr * xy + 2.0 * p1 * x + 2.0 * p2 * y;
    jd[2] = jd[1];
    jd[3] = 1.0 + rad + dr * yy + 2.0 * p2 * x + 6.0 * p1 * y;
    const double xd = x + x * rad + 2.0 * p1 * xy + p2 * (r2 + 2.0 * xx);
    const double yd = y + y * rad + 2.0 * p2 * xy + p1 * (r2 + 2.0 * yy);
    x = xd;
    y = yd;
  } else if (type == kDistEquidistant) {
    const double r2 = x * x + y * y;
    const double r = sqrt(r2);
    if (r < 1e-10) {  // point unchanged, zero Jacobian (distortion-equidistant.cc:38-43)
      jd[0] = jd[1] = jd[2] = jd[3] = 0.0;
      return;
    }
    const double th = atan(r);
    const double t2 = th * th;
    // theta_d = th * poly(t2); its derivative in th is dpoly.
    const double poly = 1.0 + t2 * (k[0] + t2 * (k[1] + t2 * (k[2] + t2 * k[3])));
    const double dpoly = 1.0 + t2 * (3.0 * k[0] + t2 * (5.0 * k[1] + t2 * (7.0 * k[2] + t2 * 9.0 * k[3])));
    const double thd = th * poly;
    const double s = thd / r;
    const double c = (dpoly / (1.0 + r2) - s) / r2;  // (ds/dr) / r
    jd[0] = s + c * x * x;
    jd[1] = c * x * y;
    jd[2] = jd[1];
    jd[3] = s + c * y * y;
    const double scale = (r > 1e-8) ? s : 1.0;
    x *= scale;
    y *= scale;
  } else if (type == kDistFisheye) {
    // FOV model (distortion-fisheye.cc:14-82): the point is scaled by
    // f(r) = atan(2 tan(w / 2) r) / (w r), with the reference's two limits:
    // w^2 < 1e-5 -> f = 1, J = I; r^2 < 1e-5 -> f = 2 tan(w / 2) / w, J = f I.
    const double w = k[0];
    const double r2 = x * x + y * y;
    double f;
    if (w * w < 1e-5) {
      f = 1.0;
      jd[0] = 1.0; jd[1] = 0.0; jd[2] = 0.0; jd[3] = 1.0;
    } else {
      const double t = tan(0.5 * w);
      if (r2 < 1e-5) {
        f = 2.0 * t / w;
        jd[0] = f; jd[1] = 0.0; jd[2] = 0.0; jd[3] = f;
      } else {
        const double r = sqrt(r2);
        const double at = atan(2.0 * t * r);
        f = at / (w * r);
        // J = f I + (f'(r) / r) [x y]' [x y].
        const double c = 2.0 * t / (w * r2 * (4.0 * t * t * r2 + 1.0)) - at / (w * r2 * r);
        jd[0] = f + c * x * x;
        jd[1] = c * x * y;
        jd[2] = jd[1];
        jd[3] = f + c * y * y;
      }
    }
    x *= f;
    y *= f;
  } else {
    jd[0] = 1.0; jd[1] = 0.0; jd[2] = 0.0; jd[3] = 1.0;
  }
}

// d(distorted point) / d(distortion parameters) at the normalized point
// (x, y), 2 x 4 row-major, the parameters a model does not have left zero:
// distortParameterJacobian of distortion-radtan.cc:67-91,
// distortion-equidistant.cc:109-142 and distortion-fisheye.cc:84-120. For
// the fisheye limit r^2 < 1e-5 the reference fills both rows with df/dw
// itself (not x df/dw, y df/dw); that is kept.
MLBA_HD void distortion_parameter_jacobian(int32_t type, const double* k, double x, double y, double* d) {
  for (int i = 0; i < 8; ++i) d[i] = 0.0;
  if (type == kDistRadtan) {
    const double r2 = x * x + y * y, r4 = r2 * r2;
    d[0] = x * r2; d[1] = x * r4; d[2] = 2.0 * x * y; d[3] = r2 + 2.0 * x * x;
    d[4] = y * r2; d[5] = y * r4; d[6] = r2 + 2.0 * y * y; d[7] = 2.0 * x * y;
  } else if (type == kDistEquidistant) {
    const double r = sqrt(x * x + y * y);
    if (r >= 1e-10) {
      const double th = atan(r);
      const double t3 = th * th * th, t5 = t3 * th * th, t7 = t5 * th * th, t9 = t7 * th * th;
      d[0] = x * t3 / r; d[1] = x * t5 / r; d[2] = x * t7 / r; d[3] = x * t9 / r;
      d[4] = y * t3 / r; d[5] = y * t5 / r; d[6] = y * t7 / r; d[7] = y * t9 / r;
    }
  } else if (type == kDistFisheye) {
    const double w = k[0];
    if (w * w < 1e-5) return;
    const double r2 = x * x + y * y;
    if (r2 < 1e-5) {
      const double cw = cos(0.5 * w);
      d[0] = d[4] = (w - sin(w)) / (w * w * cw * cw);
    } else {
      // df/dw = (1 + t^2) / (w (1 + 4 t^2 r^2)) - atan(2 t r) / (w^2 r).
      const double t = tan(0.5 * w), r = sqrt(r2);
      const double g = (1.0 + t * t) / (w * (4.0 * t * t * r2 + 1.0)) - atan(2.0 * t * r) / (w * w * r);
      d[0] = x * g;
      d[4] = y * g;
    }
  }
}

// Projection of p_C. kp in pixels; jp = d kp / d p_C (2x3). Returns the
// ProjectionResult status.
MLBA_HD int32_t project(int32_t dist_type, const double* intr, const double* dist, double width,
                        double height, const double* p, double* kp, double* jp) {
  const double rz = 1.0 / p[2];
  double u = p[0] * rz, v = p[1] * rz;
  double jd[4];
  distort(dist_type, dist, u, v, jd);
  const double fu = intr[0], fv = intr[1];
  const double rz2 = rz * rz;
  jp[0] = fu * jd[0] * rz;
  jp[1] = fu * jd[1] * rz;
  jp[2] = -fu * (p[0] * jd[0] + p[1] * jd[1]) * rz2;
  jp[3] = fv * jd[2] * rz;
  jp[4] = fv * jd[3] * rz;
  jp[5] = -fv * (p[0] * jd[2] + p[1] * jd[3]) * rz2;
  kp[0] = fu * u + intr[2];
  kp[1] = fv * v + intr[3];
  const bool inside = kp[0] >= 0.0 && kp[1] >= 0.0 && kp[0] < width && kp[1] < height;
  if (p[2] > 1e-10) return inside ? kVisible : kOutsideImage;
  return p[2] < 0.0 ? kBehind : kInvalid;
}

// Unified projection (camera-unified-projection.cc:74-196): the point is
// projected through the unit sphere with mirror parameter xi,
// (x, y) / (z + xi |p|), then distorted and scaled like the pinhole. Valid
// only for z > -fov(xi) |p| with fov(xi) = xi (xi <= 1) or 1 / xi; else the
// status is invalid and kp = 0. Visible / outside by the image box when
// |p|^2 > 1e-20, else invalid (no behind-camera status in this model).
// jxi (nullable) = d kp / d xi.
MLBA_HD int32_t project_unified(double xi, int32_t dist_type, const double* intr, const double* dist, double width,
                                double height, const double* p, double* kp, double* jp, double* jxi) {
  const double x = p[0], y = p[1], z = p[2];
  const double d2 = x * x + y * y + z * z;
  const double d = sqrt(d2);
  const double fov = (xi <= 1.0) ? xi : 1.0 / xi;
  if (!(z > -(fov * d))) {
    kp[0] = kp[1] = 0.0;
    for (int k = 0; k < 6; ++k) jp[k] = 0.0;
    if (jxi) jxi[0] = jxi[1] = 0.0;
    return kInvalid;
  }
  const double rz = 1.0 / (z + xi * d);
  double u = x * rz, v = y * rz;
  double jd[4];
  distort(dist_type, dist, u, v, jd);
  // d(x rz, y rz) / dp, 2 x 3 (row-major a).
  double a[6];
  const double s = rz * rz / d;
  a[0] = s * (d * z + xi * (y * y + z * z));
  a[1] = -s * xi * x * y;
  a[3] = a[1];
  a[4] = s * (d * z + xi * (x * x + z * z));
  const double s2 = s * (-xi * z - d);
  a[2] = x * s2;
  a[5] = y * s2;
  const double fu = intr[0], fv = intr[1];
  for (int c = 0; c < 3; ++c) {
    jp[c] = fu * (jd[0] * a[c] + jd[1] * a[3 + c]);
    jp[3 + c] = fv * (jd[2] * a[c] + jd[3] * a[3 + c]);
  }
  if (jxi) {
    const double gx = -x * rz * d * rz, gy = -y * rz * d * rz;
    jxi[0] = fu * (jd[0] * gx + jd[1] * gy);
    jxi[1] = fv * (jd[2] * gx + jd[3] * gy);
  }
  kp[0] = fu * u + intr[2];
  kp[1] = fv * v + intr[3];
  const bool inside = kp[0] >= 0.0 && kp[1] >= 0.0 && kp[0] < width && kp[1] < height;
  if (d2 > 1e-20) return inside ? kVisible : kOutsideImage;
  return kInvalid;
}

// Projection by camera model.
MLBA_HD int32_t project_camera(int32_t model, double xi, int32_t dist_type, const double* intr, const double* dist,
                               double width, double height, const double* p, double* kp, double* jp) {
  if (model == kCameraUnifiedProjection)
    return project_unified(xi, dist_type, intr, dist, width, height, p, kp, jp, nullptr);
  return project(dist_type, intr, dist, width, height, p, kp, jp);
}

// The projection-failure rule of visual-error-term-inl.h:159-166.
MLBA_HD bool projection_failed(int32_t status, const double* p_C, const double* kp) {
  return status == kBehind || status == kInvalid || p_C[2] < 0.05 ||
         (kp[0] * kp[0] + kp[1] * kp[1]) > 1.0e10;
}

// Robust loss rho(s), rho'(s), rho''(s) on the squared whitened residual.
// loss 0: Huber(a), 1: Cauchy(a), 2: trivial.
MLBA_HD void robust_loss(int32_t loss, double a, double s, double* rho) {
  if (loss == 0) {
    const double b = a * a;
    if (s > b) {
      const double r = sqrt(s);
      rho[0] = 2.0 * a * r - b;
      rho[1] = fmax(2.2250738585072014e-308, a / r);
      rho[2] = -rho[1] / (2.0 * s);
    } else {
      rho[0] = s; rho[1] = 1.0; rho[2] = 0.0;
    }
  } else if (loss == 1) {
    const double b = a * a;
    const double inv = 1.0 / (1.0 + s / b);
    rho[0] = b * log1p(s / b);
    rho[1] = fmax(2.2250738585072014e-308, inv);
    rho[2] = -(inv * inv) / b;
  } else {
    rho[0] = s; rho[1] = 1.0; rho[2] = 0.0;
  }
}

}  // namespace mlba

#endif  // MAPLAB_AMD_BA_MATH_H_
