// inertial_oracle.cc -- TEST INFRASTRUCTURE ONLY.
//
// FP64 CPU restatement of maplab's inertial error term (the IMU
// pre-integration residual of optvi): the RK4 state, transition and
// covariance propagation of imu_integrator::ImuIntegratorRK4
// (algorithms/imu-integrator-rk4/include/imu-integrator/imu-integrator-inl.h:17-338)
// and InertialErrorTerm::Evaluate
// (algorithms/ceres-error-terms/src/inertial-error-term.cc:9-286).
// Matrices are row-major; the 16-D state is [q (JPL xyzw) | b_g | v_M | b_a |
// p_M_I] and the 15-D error state [dtheta | b_g | v | b_a | p]
// (imu-integrator/common.h:6-30).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <initializer_list>
#include <vector>

#include "ba_oracle.h"

namespace {

constexpr int S = 16, E = 15;

struct Noise {
  double gn2, gb2, an2, ab2, g;
};

void skew3(const double* v, double* s) {
  s[0] = 0;     s[1] = -v[2]; s[2] = v[1];
  s[3] = v[2];  s[4] = 0;     s[5] = -v[0];
  s[6] = -v[1]; s[7] = v[0];  s[8] = 0;
}

// toRotationMatrixJPL of the normalized quaternion (B_R_G), imu-integrator-inl.h:218-222.
void rotNormalized(const double* q, double* R) {
  const double n = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  const double x = q[0] / n, y = q[1] / n, z = q[2] / n, w = q[3] / n;
  R[0] = 1 - 2 * (y * y + z * z); R[1] = 2 * (x * y + z * w);     R[2] = 2 * (x * z - y * w);
  R[3] = 2 * (x * y - z * w);     R[4] = 1 - 2 * (x * x + z * z); R[5] = 2 * (y * z + x * w);
  R[6] = 2 * (x * z + y * w);     R[7] = 2 * (y * z - x * w);     R[8] = 1 - 2 * (x * x + y * y);
}

// getStateDerivativeRungeKutta (imu-integrator-inl.h:206-238).
void stateDerivative(const Noise& nz, const double* imu, const double* x, double* dx) {
  const double n = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2] + x[3] * x[3]);
  const double q[4] = {x[0] / n, x[1] / n, x[2] / n, x[3] / n};
  double BRG[9];
  rotNormalized(x, BRG);
  const double* a = imu;      // accelerometer
  const double* w = imu + 3;  // gyroscope
  // gyroOmegaJPL (:319-333), q_dot = 0.5 Omega q
  const double Om[16] = {0, w[2], -w[1], w[0], -w[2], 0, w[0], w[1], w[1], -w[0], 0, w[2], -w[0], -w[1], -w[2], 0};
  std::memset(dx, 0, sizeof(double) * S);
  for (int r = 0; r < 4; ++r) {
    double s = 0;
    for (int c = 0; c < 4; ++c) s += Om[r * 4 + c] * q[c];
    dx[r] = 0.5 * s;
  }
  for (int r = 0; r < 3; ++r) {  // v_dot = G_R_B a - g e_z, G_R_B = B_R_G^T
    double s = 0;
    for (int c = 0; c < 3; ++c) s += BRG[c * 3 + r] * a[c];
    dx[7 + r] = s - (r == 2 ? nz.g : 0.0);
  }
  for (int r = 0; r < 3; ++r) dx[13 + r] = x[7 + r];
}

// getCovarianceTransitionDerivativesRungeKutta (imu-integrator-inl.h:240-300).
void covTransitionDerivative(const Noise& nz, const double* imu, const double* x, const double* C,
                             const double* T, double* dC, double* dT) {
  double BRG[9], GRB[9], ws[9], as[9], GRBas[9];
  rotNormalized(x, BRG);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) GRB[r * 3 + c] = BRG[c * 3 + r];
  skew3(imu + 3, ws);
  skew3(imu, as);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += GRB[r * 3 + k] * as[k * 3 + c];
      GRBas[r * 3 + c] = s;
    }
  // phi_cont blocks: (0,0) = -[w]x, (0,3) = -I, (6,0) = -G_R_B [a]x,
  // (6,9) = -G_R_B, (12,6) = I; applied to M (T or C) row-block-wise.
  auto apply = [&](const double* M, double* out) {
    std::memset(out, 0, sizeof(double) * E * E);
    for (int c = 0; c < E; ++c) {
      for (int r = 0; r < 3; ++r) {
        double s0 = 0, s6 = 0;
        for (int k = 0; k < 3; ++k) {
          s0 += -ws[r * 3 + k] * M[k * E + c];
          s6 += -GRBas[r * 3 + k] * M[k * E + c] - GRB[r * 3 + k] * M[(9 + k) * E + c];
        }
        out[r * E + c] = s0 - M[(3 + r) * E + c];
        out[(6 + r) * E + c] = s6;
        out[(12 + r) * E + c] = M[(6 + r) * E + c];
      }
    }
  };
  apply(T, dT);
  double pc[E * E];
  apply(C, pc);
  for (int r = 0; r < E; ++r)
    for (int c = 0; c < E; ++c) dC[r * E + c] = pc[r * E + c] + pc[c * E + r];
  for (int k = 0; k < 3; ++k) {
    dC[k * E + k] += nz.gn2;
    dC[(3 + k) * E + 3 + k] += nz.gb2;
    dC[(6 + k) * E + 6 + k] += nz.an2;
    dC[(9 + k) * E + 9 + k] += nz.ab2;
  }
}

// ImuIntegratorRK4::integrate (imu-integrator-inl.h:17-204).
void integrateStep(const Noise& nz, const double* x, const double* rd, double dt, double* xn, double* phi,
                   double* Q) {
  double k1[6], k23[6], k4[6];
  for (int i = 0; i < 6; ++i) {
    k1[i] = rd[i];
    k4[i] = rd[6 + i];
    // interpolateImuReadings (:302-317) at dt / 2
    k23[i] = dt < 2.220446049250313e-16 ? rd[i] : rd[i] + (rd[6 + i] - rd[i]) * ((0.5 * dt) / dt);
  }
  double d1[S], d2[S], d3[S], d4[S], xi[S];
  stateDerivative(nz, k1, x, d1);
  for (int i = 0; i < S; ++i) xi[i] = x[i] + 0.5 * dt * d1[i];
  stateDerivative(nz, k23, xi, d2);
  for (int i = 0; i < S; ++i) xi[i] = x[i] + 0.5 * dt * d2[i];
  stateDerivative(nz, k23, xi, d3);
  for (int i = 0; i < S; ++i) xi[i] = x[i] + dt * d3[i];
  stateDerivative(nz, k4, xi, d4);
  for (int i = 0; i < S; ++i) xn[i] = x[i] + dt * (d1[i] + 2.0 * d2[i] + 2.0 * d3[i] + d4[i]) / 6.0;

  static thread_local double C0[E * E], T0[E * E], Ci[E * E], Ti[E * E];
  double cd[4][E * E], td[4][E * E];
  std::memset(C0, 0, sizeof(C0));
  std::memset(T0, 0, sizeof(T0));
  for (int k = 0; k < E; ++k) T0[k * E + k] = 1.0;
  covTransitionDerivative(nz, k1, x, C0, T0, cd[0], td[0]);
  const double* dsd[3] = {d1, d2, d3};
  const double frac[3] = {0.5, 0.5, 1.0};
  const double* rk[3] = {k23, k23, k4};
  for (int st = 0; st < 3; ++st) {
    for (int i = 0; i < S; ++i) xi[i] = x[i] + frac[st] * dt * dsd[st][i];
    for (int i = 0; i < E * E; ++i) {
      Ci[i] = C0[i] + frac[st] * dt * cd[st][i];
      Ti[i] = T0[i] + frac[st] * dt * td[st][i];
    }
    covTransitionDerivative(nz, rk[st], xi, Ci, Ti, cd[st + 1], td[st + 1]);
  }
  for (int i = 0; i < E * E; ++i) {
    Q[i] = C0[i] + dt * (cd[0][i] + 2.0 * cd[1][i] + 2.0 * cd[2][i] + cd[3][i]) / 6.0;
    phi[i] = T0[i];
  }
  for (int rb : {0, 6, 12})
    for (int r = rb; r < rb + 3; ++r)
      for (int c = 0; c < E; ++c)
        phi[r * E + c] += dt * (td[0][r * E + c] + 2.0 * td[1][r * E + c] + 2.0 * td[2][r * E + c] +
                                td[3][r * E + c]) / 6.0;
}

void matmul15(const double* A, const double* B, double* C) {
  for (int r = 0; r < E; ++r)
    for (int c = 0; c < E; ++c) {
      double s = 0;
      for (int k = 0; k < E; ++k) s += A[r * E + k] * B[k * E + c];
      C[r * E + c] = s;
    }
}

// In-place lower Cholesky (Eigen::LLT); false if not positive definite.
bool cholesky15(double* A) {
  for (int j = 0; j < E; ++j) {
    double d = A[j * E + j];
    for (int k = 0; k < j; ++k) d -= A[j * E + k] * A[j * E + k];
    if (!(d > 0.0)) return false;
    const double ljj = std::sqrt(d);
    A[j * E + j] = ljj;
    for (int i = j + 1; i < E; ++i) {
      double s = A[i * E + j];
      for (int k = 0; k < j; ++k) s -= A[i * E + k] * A[j * E + k];
      A[i * E + j] = s / ljj;
    }
    for (int k = j + 1; k < E; ++k) A[j * E + k] = 0.0;
  }
  return true;
}

// L^-1 X for X with `cols` columns (row-major, leading dimension cols).
void lowerSolve(const double* L, double* X, int cols) {
  for (int c = 0; c < cols; ++c)
    for (int r = 0; r < E; ++r) {
      double s = X[r * cols + c];
      for (int k = 0; k < r; ++k) s -= L[r * E + k] * X[k * cols + c];
      X[r * cols + c] = s / L[r * E + r];
    }
}

void quatJac43(const double* q, double* J) {  // JplQuaternionParameterization::ComputeJacobian
  J[0] = q[3];  J[1] = -q[2]; J[2] = q[1];
  J[3] = q[2];  J[4] = q[3];  J[5] = -q[0];
  J[6] = -q[1]; J[7] = q[0];  J[8] = q[3];
  J[9] = -q[0]; J[10] = -q[1]; J[11] = -q[2];
  for (int i = 0; i < 12; ++i) J[i] *= 0.5;
}

}  // namespace

extern "C" {

int32_t oracle_inertial_evaluate(int32_t n, const double* imu, const int64_t* t_ns, const double* noise,
                                 const double* const* params, double* residual, double* const* jacobians,
                                 double* phi_out, double* Q_out, double* end_state_out) {
  if (n < 1 || !imu || !t_ns || !noise || !params) return MLBA_EINVAL;
  const Noise nz{noise[0] * noise[0], noise[1] * noise[1], noise[2] * noise[2], noise[3] * noise[3], noise[4]};
  // begin state: [q | b_g | v | b_a | p]
  double x[S], xn[S];
  for (int i = 0; i < 4; ++i) x[i] = params[0][i];
  for (int i = 0; i < 3; ++i) {
    x[4 + i] = params[1][i];
    x[7 + i] = params[2][i];
    x[10 + i] = params[3][i];
    x[13 + i] = params[0][4 + i];
  }
  double phiA[E * E], QA[E * E], phi[E * E], Q[E * E], tmp[E * E], tmp2[E * E];
  std::memset(QA, 0, sizeof(QA));
  std::memset(phiA, 0, sizeof(phiA));
  for (int k = 0; k < E; ++k) phiA[k * E + k] = 1.0;
  std::memcpy(xn, x, sizeof(x));
  // IntegrateStateAndCovariance (inertial-error-term.cc:9-71)
  for (int i = 0; i + 1 < n; ++i) {
    if (t_ns[i + 1] < t_ns[i]) return MLBA_EINVAL;
    double rd[12];
    for (int k = 0; k < 3; ++k) {
      rd[k] = imu[6 * i + k] - x[10 + k];
      rd[3 + k] = imu[6 * i + 3 + k] - x[4 + k];
      rd[6 + k] = imu[6 * (i + 1) + k] - x[10 + k];
      rd[9 + k] = imu[6 * (i + 1) + 3 + k] - x[4 + k];
    }
    const double dt = (double)(t_ns[i + 1] - t_ns[i]) * 1e-9;
    integrateStep(nz, x, rd, dt, xn, phi, Q);
    std::memcpy(x, xn, sizeof(x));
    // Q_accum = phi Q_accum phi^T + Q; phi_accum = phi phi_accum
    matmul15(phi, QA, tmp);
    for (int r = 0; r < E; ++r)
      for (int c = 0; c < E; ++c) {
        double s = 0;
        for (int k = 0; k < E; ++k) s += tmp[r * E + k] * phi[c * E + k];
        tmp2[r * E + c] = s + Q[r * E + c];
      }
    std::memcpy(QA, tmp2, sizeof(QA));
    matmul15(phi, phiA, tmp);
    std::memcpy(phiA, tmp, sizeof(phiA));
  }
  if (phi_out) std::memcpy(phi_out, phiA, sizeof(phiA));
  if (Q_out) std::memcpy(Q_out, QA, sizeof(QA));
  if (end_state_out) std::memcpy(end_state_out, xn, sizeof(xn));
  double L[E * E];
  std::memcpy(L, QA, sizeof(L));
  if (!cholesky15(L)) return MLBA_ESTATE;
  const double* qt = params[4];
  if (residual) {
    // delta_q = positive JPL product of q_to and the inverse of the integrated q (:136-147)
    const double* qe = xn;
    const double n2 = qe[0] * qe[0] + qe[1] * qe[1] + qe[2] * qe[2] + qe[3] * qe[3];
    const double qi[4] = {-qe[0] / n2, -qe[1] / n2, -qe[2] / n2, qe[3] / n2};
    // quaternion-math-inl.h:113-138: L(q1) = [[w I - [v]x, v], [-v^T, w]]
    double sk[9];
    skew3(qt, sk);
    double dq[4];
    for (int r = 0; r < 3; ++r) {
      double s = 0;
      for (int c = 0; c < 3; ++c) s += ((r == c ? qt[3] : 0.0) - sk[r * 3 + c]) * qi[c];
      dq[r] = s + qt[r] * qi[3];
    }
    dq[3] = -(qt[0] * qi[0] + qt[1] * qi[1] + qt[2] * qi[2]) + qt[3] * qi[3];
    const double dn = std::sqrt(dq[0] * dq[0] + dq[1] * dq[1] + dq[2] * dq[2] + dq[3] * dq[3]);
    for (int k = 0; k < 4; ++k) dq[k] /= dn;
    if (dq[3] < 0)
      for (int k = 0; k < 4; ++k) dq[k] = -dq[k];
    for (int k = 0; k < 3; ++k) {
      residual[k] = 2.0 * dq[k];
      residual[3 + k] = params[5][k] - xn[4 + k];
      residual[6 + k] = params[6][k] - xn[7 + k];
      residual[9 + k] = params[7][k] - xn[10 + k];
      residual[12 + k] = params[4][4 + k] - xn[13 + k];
    }
    lowerSolve(L, residual, 1);
  }
  if (jacobians) {
    // J_end, J_begin: 15 x 16 over the state [q | b_g | v | b_a | p] (:164-205)
    double Jend[E * S], Jbeg[E * S], te[12], tb[12];
    quatJac43(qt, te);
    quatJac43(params[0], tb);
    std::memset(Jend, 0, sizeof(Jend));
    std::memset(Jbeg, 0, sizeof(Jbeg));
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 4; ++c) Jend[r * S + c] = 4.0 * te[c * 3 + r];
    for (int k = 0; k < 12; ++k) Jend[(3 + k) * S + 4 + k] = 1.0;
    for (int r = 0; r < E; ++r) {
      for (int c = 0; c < 4; ++c) {
        double s = 0;
        for (int k = 0; k < 3; ++k) s += phiA[r * E + k] * tb[c * 3 + k];
        Jbeg[r * S + c] = -4.0 * s;
      }
      for (int c = 0; c < 12; ++c) Jbeg[r * S + 4 + c] = -phiA[r * E + 3 + c];
    }
    lowerSolve(L, Jend, S);
    lowerSolve(L, Jbeg, S);
    // parameter blocks: pose (q 0-3, p 13-15), b_g 4-6, v 7-9, b_a 10-12 (:207-283)
    for (int side = 0; side < 2; ++side) {
      const double* J = side ? Jend : Jbeg;
      double* const* out = jacobians + 4 * side;
      for (int r = 0; r < E; ++r) {
        if (out[0]) {
          for (int c = 0; c < 4; ++c) out[0][r * 7 + c] = J[r * S + c];
          for (int c = 0; c < 3; ++c) out[0][r * 7 + 4 + c] = J[r * S + 13 + c];
        }
        for (int b = 1; b < 4; ++b)
          if (out[b])
            for (int c = 0; c < 3; ++c) out[b][r * 3 + c] = J[r * S + 1 + 3 * b + c];
      }
    }
  }
  return MLBA_OK;
}


// Speed and bias states with the poses constant (addInertialTermsForEdges,
// optimization-terms-addition.cc:420-490, fix flags :468-485): dense LM over
// the free [b_g | v_M | b_a] columns of the edges' vertices, with the ceres
// rules of the oracle's minimize(). edge_from / edge_to index vertices;
// samples of edge e are [sample_begin[e], sample_begin[e + 1]).
int32_t oracle_solve_inertial_states(int32_t P, int32_t E, const double* vertices, const double* gyro_bias,
                                     const double* velocity, const double* accel_bias, const int32_t* edge_from,
                                     const int32_t* edge_to, const int32_t* sample_begin, const double* imu,
                                     const int64_t* t_ns, const double* noise, const int32_t* fix,
                                     const mlba_options* options, double* gyro_bias_out, double* velocity_out,
                                     double* accel_bias_out, mlba_summary* summary) {
  if (P <= 0 || E <= 0 || !options || !fix) return MLBA_EINVAL;
  const mlba_options& o = *options;
  // Columns: vertex v, group k (0 b_g, 1 v_M, 2 b_a) -> index or -1.
  std::vector<int> col((size_t)P * 9, -1);
  std::vector<uint8_t> in(P, 0);
  for (int e = 0; e < E; ++e) in[edge_from[e]] = in[edge_to[e]] = 1;
  const bool fixed_group[3] = {fix[0] != 0, fix[2] != 0, fix[1] != 0};  // fix = {gyro, accel, velocity}
  int n = 0;
  for (int v = 0; v < P; ++v)
    if (in[v])
      for (int k = 0; k < 9; ++k)
        if (!fixed_group[k / 3]) col[(size_t)v * 9 + k] = n++;
  if (n == 0) return MLBA_EINVAL;
  std::vector<double> st[3] = {std::vector<double>(gyro_bias, gyro_bias + 3 * P),
                               std::vector<double>(velocity, velocity + 3 * P),
                               std::vector<double>(accel_bias, accel_bias + 3 * P)};
  auto linearize = [&](const std::vector<double>* s, double* cost, std::vector<double>* A, std::vector<double>* g) {
    *cost = 0.0;
    if (A) { A->assign((size_t)n * n, 0.0); g->assign(n, 0.0); }
    for (int e = 0; e < E; ++e) {
      const int a = edge_from[e], b = edge_to[e];
      const double* params[8] = {vertices + 7 * a, &s[0][3 * a], &s[1][3 * a], &s[2][3 * a],
                                 vertices + 7 * b, &s[0][3 * b], &s[1][3 * b], &s[2][3 * b]};
      double r[15], Jb[8][45];
      double* jac[8] = {nullptr, A ? Jb[1] : nullptr, A ? Jb[2] : nullptr, A ? Jb[3] : nullptr,
                        nullptr, A ? Jb[5] : nullptr, A ? Jb[6] : nullptr, A ? Jb[7] : nullptr};
      const int rc = oracle_inertial_evaluate(sample_begin[e + 1] - sample_begin[e], imu + 6 * sample_begin[e],
                                              t_ns + sample_begin[e], noise, params, r, jac, nullptr, nullptr,
                                              nullptr);
      if (rc) return rc;
      for (int k = 0; k < 15; ++k) *cost += 0.5 * r[k] * r[k];
      if (!A) continue;
      // Free columns of this edge: (vertex, group, component) -> (block, column).
      int gc[18], blk[18], cc[18], m = 0;
      for (int side = 0; side < 2; ++side)
        for (int k = 0; k < 9; ++k) {
          const int v = side ? b : a;
          const int c = col[(size_t)v * 9 + k];
          if (c < 0) continue;
          gc[m] = c;
          blk[m] = (side ? 5 : 1) + k / 3;
          cc[m] = k % 3;
          ++m;
        }
      for (int i = 0; i < m; ++i) {
        for (int rr = 0; rr < 15; ++rr) (*g)[gc[i]] += Jb[blk[i]][3 * rr + cc[i]] * r[rr];
        for (int j = 0; j < m; ++j) {
          double v = 0.0;
          for (int rr = 0; rr < 15; ++rr) v += Jb[blk[i]][3 * rr + cc[i]] * Jb[blk[j]][3 * rr + cc[j]];
          (*A)[(size_t)gc[i] * n + gc[j]] += v;
        }
      }
    }
    return (int)MLBA_OK;
  };
  auto push = [&](const mlba_iteration_summary& it) {
    if (!summary) return;
    if (summary->iterations && summary->num_iterations < summary->iterations_capacity)
      summary->iterations[summary->num_iterations] = it;
    summary->num_iterations++;
  };
  if (summary) {
    summary->num_iterations = summary->num_successful_steps = summary->num_unsuccessful_steps = 0;
    summary->num_residual_blocks = E;
  }
  std::vector<double> A, g;
  double cost;
  if (linearize(st, &cost, &A, &g)) return MLBA_ESTATE;
  auto gmax = [&] {
    double m = 0.0;
    for (double x : g) m = std::max(m, std::fabs(x));
    return m;
  };
  double gm = gmax();
  if (summary) { summary->initial_cost = cost; summary->fixed_cost = 0.0; }
  mlba_iteration_summary it0{};
  it0.step_is_valid = it0.step_is_successful = 1;
  it0.cost = cost;
  it0.gradient_max_norm = gm;
  double radius = o.initial_trust_region_radius, decrease = 2.0;
  it0.trust_region_radius = radius;
  push(it0);
  int term = gm <= o.gradient_tolerance ? MLBA_CONVERGENCE : MLBA_NO_CONVERGENCE;
  int invalid = 0, iteration = 0;
  while (term == MLBA_NO_CONVERGENCE && o.max_num_iterations > 0) {
    ++iteration;
    mlba_iteration_summary cur{};
    cur.iteration = iteration;
    std::vector<double> sc(n), M((size_t)n * n), w(n);
    for (int j = 0; j < n; ++j) sc[j] = o.use_jacobi_scaling ? 1.0 / (1.0 + std::sqrt(A[(size_t)j * n + j])) : 1.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) M[(size_t)i * n + j] = sc[i] * A[(size_t)i * n + j] * sc[j];
    for (int j = 0; j < n; ++j)
      M[(size_t)j * n + j] += std::min(std::max(M[(size_t)j * n + j], o.min_lm_diagonal), o.max_lm_diagonal) / radius;
    bool ok = true;
    for (int j = 0; j < n && ok; ++j) {  // dense Cholesky in place (lower)
      double d = M[(size_t)j * n + j];
      for (int k = 0; k < j; ++k) d -= M[(size_t)j * n + k] * M[(size_t)j * n + k];
      if (!(d > 0.0)) { ok = false; break; }
      d = std::sqrt(d);
      M[(size_t)j * n + j] = d;
      for (int i = j + 1; i < n; ++i) {
        double t = M[(size_t)i * n + j];
        for (int k = 0; k < j; ++k) t -= M[(size_t)i * n + k] * M[(size_t)j * n + k];
        M[(size_t)i * n + j] = t / d;
      }
    }
    double mcc = 0.0;
    if (ok) {
      for (int i = 0; i < n; ++i) {
        double t = sc[i] * g[i];
        for (int k = 0; k < i; ++k) t -= M[(size_t)i * n + k] * w[k];
        w[i] = t / M[(size_t)i * n + i];
      }
      for (int i = n - 1; i >= 0; --i) {
        double t = w[i];
        for (int k = i + 1; k < n; ++k) t -= M[(size_t)k * n + i] * w[k];
        w[i] = t / M[(size_t)i * n + i];
      }
      for (int j = 0; j < n; ++j) w[j] *= sc[j];
      double gw = 0.0, wAw = 0.0;
      for (int i = 0; i < n; ++i) {
        gw += g[i] * w[i];
        for (int j = 0; j < n; ++j) wAw += w[i] * A[(size_t)i * n + j] * w[j];
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
      cur.cost = cost;
      cur.gradient_max_norm = gm;
      cur.trust_region_radius = radius;
      if (summary) summary->num_unsuccessful_steps++;
    } else {
      invalid = 0;
      std::vector<double> s2[3] = {st[0], st[1], st[2]};
      double ss = 0.0, xx = 0.0;
      for (int v = 0; v < P; ++v)
        for (int k = 0; k < 9; ++k) {
          const int c = col[(size_t)v * 9 + k];
          if (c < 0) continue;
          const double x = st[k / 3][3 * v + k % 3];
          s2[k / 3][3 * v + k % 3] = x - w[c];
          ss += w[c] * w[c];
          xx += x * x;
        }
      cur.step_norm = std::sqrt(ss);
      if (cur.step_norm <= o.parameter_tolerance * (std::sqrt(xx) + o.parameter_tolerance)) {
        term = MLBA_CONVERGENCE;
        break;
      }
      double ccost;
      if (linearize(s2, &ccost, nullptr, nullptr)) return MLBA_ESTATE;
      cur.cost_change = cost - ccost;
      if (std::fabs(cur.cost_change) <= o.function_tolerance * cost) { term = MLBA_CONVERGENCE; break; }
      cur.relative_decrease = (cost - ccost) / mcc;
      if (cur.relative_decrease > o.min_relative_decrease) {
        for (int k = 0; k < 3; ++k) st[k] = s2[k];
        if (linearize(st, &cost, &A, &g)) return MLBA_ESTATE;
        gm = gmax();
        cur.step_is_successful = 1;
        cur.cost = cost;
        cur.gradient_max_norm = gm;
        const double qq = 2.0 * cur.relative_decrease - 1.0;
        radius = std::min(o.max_trust_region_radius, radius / std::max(1.0 / 3.0, 1.0 - qq * qq * qq));
        decrease = 2.0;
        if (summary) summary->num_successful_steps++;
      } else {
        radius /= decrease;
        decrease *= 2.0;
        cur.cost = ccost;
        cur.gradient_max_norm = gm;
        if (summary) summary->num_unsuccessful_steps++;
      }
      cur.trust_region_radius = radius;
    }
    push(cur);
    if (iteration >= o.max_num_iterations) break;
    if (cur.step_is_successful && cur.gradient_max_norm <= o.gradient_tolerance) { term = MLBA_CONVERGENCE; break; }
    if (radius <= o.min_trust_region_radius) { term = MLBA_CONVERGENCE; break; }
  }
  if (summary) { summary->termination_type = term; summary->final_cost = cost; }
  if (gyro_bias_out) std::memcpy(gyro_bias_out, st[0].data(), sizeof(double) * 3 * P);
  if (velocity_out) std::memcpy(velocity_out, st[1].data(), sizeof(double) * 3 * P);
  if (accel_bias_out) std::memcpy(accel_bias_out, st[2].data(), sizeof(double) * 3 * P);
  return MLBA_OK;
}

}  // extern "C"
