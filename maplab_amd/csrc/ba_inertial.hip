// ba_inertial.hip -- the inertial (IMU pre-integration) error term of optvi
// on the device: RK4 integration of the 16-D IMU state with the 15x15
// transition and covariance propagation (imu_integrator::ImuIntegratorRK4,
// algorithms/imu-integrator-rk4/include/imu-integrator/imu-integrator-inl.h:
// 17-338), whitening with the Cholesky factor of the accumulated covariance,
// and the residual and ambient Jacobians of InertialErrorTerm::Evaluate
// (algorithms/ceres-error-terms/src/inertial-error-term.cc:73-286).
//
// One 256-thread work-group per edge. The state is advanced redundantly by
// every thread (a few hundred flops per stage); each 15x15 matrix lives in
// LDS with thread t < 225 owning entry (t / 15, t % 15).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/maplab_ba_inertial.h"

namespace mlba {
namespace {

constexpr int kS = 16, kE = 15, kEE = kE * kE;
constexpr int kInertialThreads = 256;

struct ImuNoise {
  double gn2, gb2, an2, ab2, g;
};

struct InertialArgs {
  int32_t num_edges;
  const double* vert;      // 7 P
  const double* bg;        // 3 P
  const double* vel;       // 3 P
  const double* ba;        // 3 P
  const int32_t* from;
  const int32_t* to;
  const int32_t* sample_begin;
  const double* imu;       // 6 per sample
  const int64_t* t_ns;
  ImuNoise nz;
  double* residual;        // 15 per edge
  double* jac[8];          // 15 x {7,3,3,3,7,3,3,3} per edge (nullable)
  int32_t* fail;
};

__device__ inline void rot_normalized(const double* q, double* R) {
  const double n = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  const double x = q[0] / n, y = q[1] / n, z = q[2] / n, w = q[3] / n;
  R[0] = 1 - 2 * (y * y + z * z); R[1] = 2 * (x * y + z * w);     R[2] = 2 * (x * z - y * w);
  R[3] = 2 * (x * y - z * w);     R[4] = 1 - 2 * (x * x + z * z); R[5] = 2 * (y * z + x * w);
  R[6] = 2 * (x * z + y * w);     R[7] = 2 * (y * z - x * w);     R[8] = 1 - 2 * (x * x + y * y);
}

// d/dt of the IMU state at debiased readings (accelerometer, gyroscope).
__device__ inline void state_derivative(const ImuNoise& nz, const double* imu, const double* x, double* dx) {
  const double n = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2] + x[3] * x[3]);
  const double q[4] = {x[0] / n, x[1] / n, x[2] / n, x[3] / n};
  const double* a = imu;
  const double* w = imu + 3;
  double R[9];
  rot_normalized(x, R);
  // q_dot = 0.5 Omega(w) q with the JPL Omega
  dx[0] = 0.5 * (w[2] * q[1] - w[1] * q[2] + w[0] * q[3]);
  dx[1] = 0.5 * (-w[2] * q[0] + w[0] * q[2] + w[1] * q[3]);
  dx[2] = 0.5 * (w[1] * q[0] - w[0] * q[1] + w[2] * q[3]);
  dx[3] = 0.5 * (-w[0] * q[0] - w[1] * q[1] - w[2] * q[2]);
  for (int k = 4; k < 7; ++k) dx[k] = 0.0;
  for (int r = 0; r < 3; ++r) dx[7 + r] = R[r] * a[0] + R[3 + r] * a[1] + R[6 + r] * a[2] - (r == 2 ? nz.g : 0.0);
  for (int k = 10; k < 13; ++k) dx[k] = 0.0;
  for (int r = 0; r < 3; ++r) dx[13 + r] = x[7 + r];
}

// Entry (r, c) of phi_cont * M for the continuous error-state transition
// (blocks (0,0) = -[w]x, (0,3) = -I, (6,0) = -G_R_B [a]x, (6,9) = -G_R_B,
// (12,6) = I); M in LDS.
__device__ inline double phi_cont_times(const double* ws, const double* GRBas, const double* GRB, const double* M,
                                        int r, int c) {
  if (r < 3) {
    double s = -M[(3 + r) * kE + c];
    for (int k = 0; k < 3; ++k) s -= ws[r * 3 + k] * M[k * kE + c];
    return s;
  }
  if (r >= 6 && r < 9) {
    const int rr = r - 6;
    double s = 0.0;
    for (int k = 0; k < 3; ++k) s -= GRBas[rr * 3 + k] * M[k * kE + c] + GRB[rr * 3 + k] * M[(9 + k) * kE + c];
    return s;
  }
  if (r >= 12) return M[(r - 6) * kE + c];
  return 0.0;
}

__global__ __launch_bounds__(kInertialThreads) void k_inertial_evaluate(InertialArgs a) {
  __shared__ double Ti[kEE], Ci[kEE], pc[kEE];
  __shared__ double td[4][kEE], cd[4][kEE];
  __shared__ double phiA[kEE], QA[kEE], phi[kEE], Q[kEE], tmp[kEE];
  __shared__ double Jb[kE * kS], Je[kE * kS];
  __shared__ double res[kE];
  __shared__ int bad;
  const int e = blockIdx.x, tid = threadIdx.x;
  const int r = tid / kE, c = tid - (tid / kE) * kE;
  const bool owner = tid < kEE;
  const int vf = a.from[e], vt = a.to[e];
  double x[kS];
  for (int k = 0; k < 4; ++k) x[k] = a.vert[7 * vf + k];
  for (int k = 0; k < 3; ++k) {
    x[4 + k] = a.bg[3 * vf + k];
    x[7 + k] = a.vel[3 * vf + k];
    x[10 + k] = a.ba[3 * vf + k];
    x[13 + k] = a.vert[7 * vf + 4 + k];
  }
  if (owner) {
    phiA[tid] = r == c ? 1.0 : 0.0;
    QA[tid] = 0.0;
  }
  if (tid == 0) bad = 0;
  const int sb = a.sample_begin[e], se = a.sample_begin[e + 1];
  for (int i = sb; i + 1 < se; ++i) {
    // Debiased readings at both ends and at the midpoint.
    double k1[6], k4[6], k23[6];
    const double dt = (double)(a.t_ns[i + 1] - a.t_ns[i]) * 1e-9;
    for (int k = 0; k < 3; ++k) {
      k1[k] = a.imu[6 * i + k] - x[10 + k];
      k1[3 + k] = a.imu[6 * i + 3 + k] - x[4 + k];
      k4[k] = a.imu[6 * (i + 1) + k] - x[10 + k];
      k4[3 + k] = a.imu[6 * (i + 1) + 3 + k] - x[4 + k];
    }
    for (int k = 0; k < 6; ++k)
      k23[k] = dt < 2.220446049250313e-16 ? k1[k] : k1[k] + (k4[k] - k1[k]) * ((0.5 * dt) / dt);
    // State RK4.
    double d[4][kS], xi[kS];
    state_derivative(a.nz, k1, x, d[0]);
    for (int k = 0; k < kS; ++k) xi[k] = x[k] + 0.5 * dt * d[0][k];
    state_derivative(a.nz, k23, xi, d[1]);
    for (int k = 0; k < kS; ++k) xi[k] = x[k] + 0.5 * dt * d[1][k];
    state_derivative(a.nz, k23, xi, d[2]);
    for (int k = 0; k < kS; ++k) xi[k] = x[k] + dt * d[2][k];
    state_derivative(a.nz, k4, xi, d[3]);
    // Transition and covariance RK4 from (I, 0) over the step.
    const double frac[4] = {0.0, 0.5, 0.5, 1.0};
    const double* rk[4] = {k1, k23, k23, k4};
#pragma unroll
    for (int st = 0; st < 4; ++st) {
      for (int k = 0; k < kS; ++k) xi[k] = st == 0 ? x[k] : x[k] + frac[st] * dt * d[st - 1][k];
      __syncthreads();
      if (owner) {
        Ti[tid] = (r == c ? 1.0 : 0.0) + (st ? frac[st] * dt * td[st - 1][tid] : 0.0);
        Ci[tid] = st ? frac[st] * dt * cd[st - 1][tid] : 0.0;
      }
      __syncthreads();
      double R[9], GRB[9], ws[9], as[9], GRBas[9];
      rot_normalized(xi, R);
      for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q) GRB[p * 3 + q] = R[q * 3 + p];
      const double* rd = rk[st];
      ws[0] = 0; ws[1] = -rd[5]; ws[2] = rd[4]; ws[3] = rd[5]; ws[4] = 0; ws[5] = -rd[3];
      ws[6] = -rd[4]; ws[7] = rd[3]; ws[8] = 0;
      as[0] = 0; as[1] = -rd[2]; as[2] = rd[1]; as[3] = rd[2]; as[4] = 0; as[5] = -rd[0];
      as[6] = -rd[1]; as[7] = rd[0]; as[8] = 0;
      for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q)
          GRBas[p * 3 + q] = GRB[p * 3] * as[q] + GRB[p * 3 + 1] * as[3 + q] + GRB[p * 3 + 2] * as[6 + q];
      if (owner) {
        td[st][tid] = phi_cont_times(ws, GRBas, GRB, Ti, r, c);
        pc[tid] = phi_cont_times(ws, GRBas, GRB, Ci, r, c);
      }
      __syncthreads();
      if (owner) {
        double v = pc[tid] + pc[c * kE + r];
        if (r == c) v += r < 3 ? a.nz.gn2 : r < 6 ? a.nz.gb2 : r < 9 ? a.nz.an2 : r < 12 ? a.nz.ab2 : 0.0;
        cd[st][tid] = v;
      }
    }
    __syncthreads();
    if (owner) {
      Q[tid] = dt * (cd[0][tid] + 2.0 * cd[1][tid] + 2.0 * cd[2][tid] + cd[3][tid]) / 6.0;
      double p = r == c ? 1.0 : 0.0;
      if (r < 3 || (r >= 6 && r < 9) || r >= 12)
        p += dt * (td[0][tid] + 2.0 * td[1][tid] + 2.0 * td[2][tid] + td[3][tid]) / 6.0;
      phi[tid] = p;
    }
    for (int k = 0; k < kS; ++k) x[k] += dt * (d[0][k] + 2.0 * d[1][k] + 2.0 * d[2][k] + d[3][k]) / 6.0;
    __syncthreads();
    // Q_accum = phi Q_accum phi^T + Q; phi_accum = phi phi_accum.
    if (owner) {
      double s = 0.0;
      for (int k = 0; k < kE; ++k) s += phi[r * kE + k] * QA[k * kE + c];
      tmp[tid] = s;
    }
    __syncthreads();
    double qa = 0.0, pa = 0.0;
    if (owner) {
      for (int k = 0; k < kE; ++k) {
        qa += tmp[r * kE + k] * phi[c * kE + k];
        pa += phi[r * kE + k] * phiA[k * kE + c];
      }
      qa += Q[tid];
    }
    __syncthreads();
    if (owner) {
      QA[tid] = qa;
      phiA[tid] = pa;
    }
  }
  __syncthreads();
  // Cholesky of Q_accum in place (lower), column by column.
  for (int j = 0; j < kE; ++j) {
    if (tid == 0) {
      double dj = QA[j * kE + j];
      for (int k = 0; k < j; ++k) dj -= QA[j * kE + k] * QA[j * kE + k];
      if (!(dj > 0.0)) bad = 1;
      QA[j * kE + j] = dj > 0.0 ? sqrt(dj) : 1.0;
    }
    __syncthreads();
    if (tid > j && tid < kE) {
      double s = QA[tid * kE + j];
      for (int k = 0; k < j; ++k) s -= QA[tid * kE + k] * QA[j * kE + k];
      QA[tid * kE + j] = s / QA[j * kE + j];
    }
    __syncthreads();
  }
  if (bad) {
    if (tid == 0) atomicAdd(a.fail, 1);
    return;
  }
  // Residual (thread 0) and the unwhitened Jacobians J_end, J_begin (15 x 16).
  const double* qt = a.vert + 7 * vt;
  if (tid == 0) {
    const double n2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2] + x[3] * x[3];
    const double qi[4] = {-x[0] / n2, -x[1] / n2, -x[2] / n2, x[3] / n2};
    // JPL product q_to (x) q_int^-1: [w1 v2 + w2 v1 - v1 x v2 ; w1 w2 - v1.v2]
    double dq[4];
    dq[0] = qt[3] * qi[0] + qi[3] * qt[0] - (qt[1] * qi[2] - qt[2] * qi[1]);
    dq[1] = qt[3] * qi[1] + qi[3] * qt[1] - (qt[2] * qi[0] - qt[0] * qi[2]);
    dq[2] = qt[3] * qi[2] + qi[3] * qt[2] - (qt[0] * qi[1] - qt[1] * qi[0]);
    dq[3] = qt[3] * qi[3] - (qt[0] * qi[0] + qt[1] * qi[1] + qt[2] * qi[2]);
    const double dn = sqrt(dq[0] * dq[0] + dq[1] * dq[1] + dq[2] * dq[2] + dq[3] * dq[3]);
    const double sg = dq[3] < 0.0 ? -1.0 : 1.0;
    for (int k = 0; k < 3; ++k) {
      res[k] = 2.0 * sg * dq[k] / dn;
      res[3 + k] = a.bg[3 * vt + k] - x[4 + k];
      res[6 + k] = a.vel[3 * vt + k] - x[7 + k];
      res[9 + k] = a.ba[3 * vt + k] - x[10 + k];
      res[12 + k] = a.vert[7 * vt + 4 + k] - x[13 + k];
    }
  }
  // JPL parameterization Jacobians (4 x 3) of q_to and q_from:
  // 0.5 [[w,-z,y],[z,w,-x],[-y,x,w],[-x,-y,-z]].
  __shared__ double th[2][12];
  if (tid < 2) {
    const double* q = a.vert + 7 * (tid ? vf : vt);
    const double m[12] = {q[3], -q[2], q[1], q[2], q[3], -q[0], -q[1], q[0], q[3], -q[0], -q[1], -q[2]};
    for (int k = 0; k < 12; ++k) th[tid][k] = 0.5 * m[k];
  }
  __syncthreads();
  for (int k = tid; k < kE * kS; k += blockDim.x) {
    const int rr = k / kS, cc = k % kS;
    double je = 0.0, jb = 0.0;
    if (rr < 3 && cc < 4) je = 4.0 * th[0][cc * 3 + rr];
    if (rr >= 3 && cc == rr + 1) je = 1.0;
    if (cc < 4) {
      double s = 0.0;
      for (int m = 0; m < 3; ++m) s += phiA[rr * kE + m] * th[1][cc * 3 + m];
      jb = -4.0 * s;
    } else {
      jb = -phiA[rr * kE + cc - 1];
    }
    Je[k] = je;
    Jb[k] = jb;
  }
  __syncthreads();
  // Whitening: forward substitution with L, one thread per column
  // (32 Jacobian columns and the residual).
  if (tid < 2 * kS + 1) {
    double* col = tid < kS ? Je + tid : (tid < 2 * kS ? Jb + (tid - kS) : res);
    const int ld = tid < 2 * kS ? kS : 1;
    for (int i = 0; i < kE; ++i) {
      double s = col[i * ld];
      for (int k = 0; k < i; ++k) s -= QA[i * kE + k] * col[k * ld];
      col[i * ld] = s / QA[i * kE + i];
    }
  }
  __syncthreads();
  if (tid < kE) a.residual[(size_t)kE * e + tid] = res[tid];
  // Parameter blocks: pose (q 0-3 | p 13-15), b_g 4-6, v 7-9, b_a 10-12.
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const int sizes[8] = {7, 3, 3, 3, 7, 3, 3, 3};
    double* out = a.jac[b];
    if (!out) continue;
    const double* J = b < 4 ? Jb : Je;
    const int kb = b & 3, sz = sizes[b];
    for (int k = tid; k < kE * sz; k += blockDim.x) {
      const int rr = k / sz, cc = k % sz;
      const int sc = kb == 0 ? (cc < 4 ? cc : 13 + cc - 4) : 1 + 3 * kb + cc;
      out[(size_t)kE * sz * e + k] = J[rr * kS + sc];
    }
  }
}

thread_local std::string g_inertial_error;

int fail(int code, const char* msg) {
  g_inertial_error = msg;
  return code;
}

template <typename T>
struct DevArray {
  T* p = nullptr;
  ~DevArray() {
    if (p) (void)hipFree(p);
  }
  bool alloc(size_t n) { return hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) == hipSuccess; }
  bool upload(const T* src, size_t n) {
    return alloc(n) && (n == 0 || hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice) == hipSuccess);
  }
};

}  // namespace
}  // namespace mlba

extern "C" {

const char* mlba_inertial_last_error(void) { return mlba::g_inertial_error.c_str(); }

int32_t mlba_inertial_evaluate(const mlba_inertial_problem* p, double* residuals, double* const* jacobians) {
  using namespace mlba;
  if (!p || !residuals) return fail(MLBA_EINVAL, "NULL argument");
  const int P = p->num_vertices, E = p->num_edges, S = p->num_samples;
  if (P < 0 || E < 0 || S < 0) return fail(MLBA_EINVAL, "negative size");
  if (E == 0) return MLBA_OK;
  if (!p->vertex_q_IM__M_p_MI || !p->vertex_gyro_bias || !p->vertex_velocity_M || !p->vertex_accel_bias ||
      !p->edge_from || !p->edge_to || !p->edge_sample_begin || !p->imu || !p->timestamps_ns)
    return fail(MLBA_EINVAL, "NULL array in inertial problem");
  if (!(p->gyro_noise_sigma > 0.0) || !(p->gyro_bias_sigma > 0.0) || !(p->acc_noise_sigma > 0.0) ||
      !(p->acc_bias_sigma > 0.0))
    return fail(MLBA_EINVAL, "IMU noise sigmas must be > 0");
  if (p->edge_sample_begin[0] != 0 || p->edge_sample_begin[E] > S) return fail(MLBA_EINVAL, "edge samples out of range");
  for (int e = 0; e < E; ++e) {
    if (p->edge_from[e] < 0 || p->edge_from[e] >= P || p->edge_to[e] < 0 || p->edge_to[e] >= P)
      return fail(MLBA_EINVAL, "edge vertex out of range");
    if (p->edge_sample_begin[e + 1] <= p->edge_sample_begin[e]) return fail(MLBA_EINVAL, "edge without IMU data");
    for (int i = p->edge_sample_begin[e]; i + 1 < p->edge_sample_begin[e + 1]; ++i)
      if (p->timestamps_ns[i + 1] < p->timestamps_ns[i]) return fail(MLBA_EINVAL, "IMU measurements not properly ordered");
  }
  DevArray<double> vert, bg, vel, ba, imu, res, jac[8];
  DevArray<int32_t> from, to, sb, failc;
  DevArray<int64_t> tns;
  if (!vert.upload(p->vertex_q_IM__M_p_MI, 7 * (size_t)P) || !bg.upload(p->vertex_gyro_bias, 3 * (size_t)P) ||
      !vel.upload(p->vertex_velocity_M, 3 * (size_t)P) || !ba.upload(p->vertex_accel_bias, 3 * (size_t)P) ||
      !from.upload(p->edge_from, E) || !to.upload(p->edge_to, E) || !sb.upload(p->edge_sample_begin, E + 1) ||
      !imu.upload(p->imu, 6 * (size_t)S) || !tns.upload(p->timestamps_ns, S) || !res.alloc(15 * (size_t)E) ||
      !failc.alloc(1))
    return fail(MLBA_ENOMEM, "device allocation failed");
  const int sizes[8] = {7, 3, 3, 3, 7, 3, 3, 3};
  InertialArgs a{};
  a.num_edges = E;
  a.vert = vert.p; a.bg = bg.p; a.vel = vel.p; a.ba = ba.p;
  a.from = from.p; a.to = to.p; a.sample_begin = sb.p; a.imu = imu.p; a.t_ns = tns.p;
  a.nz = ImuNoise{p->gyro_noise_sigma * p->gyro_noise_sigma, p->gyro_bias_sigma * p->gyro_bias_sigma,
                  p->acc_noise_sigma * p->acc_noise_sigma, p->acc_bias_sigma * p->acc_bias_sigma,
                  p->gravity_magnitude};
  a.residual = res.p;
  for (int b = 0; b < 8; ++b) {
    a.jac[b] = nullptr;
    if (jacobians && jacobians[b]) {
      if (!jac[b].alloc(15 * (size_t)sizes[b] * E)) return fail(MLBA_ENOMEM, "device allocation failed");
      a.jac[b] = jac[b].p;
    }
  }
  a.fail = failc.p;
  if (hipMemset(failc.p, 0, sizeof(int32_t)) != hipSuccess) return fail(MLBA_EDEVICE, "hipMemset failed");
  k_inertial_evaluate<<<E, kInertialThreads>>>(a);
  if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
    return fail(MLBA_EDEVICE, "inertial kernel failed");
  int32_t nfail = 0;
  if (hipMemcpy(&nfail, failc.p, sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(residuals, res.p, 15 * sizeof(double) * E, hipMemcpyDeviceToHost) != hipSuccess)
    return fail(MLBA_EDEVICE, "copy failed");
  for (int b = 0; b < 8; ++b)
    if (a.jac[b] && hipMemcpy(jacobians[b], a.jac[b], 15 * sizeof(double) * sizes[b] * E, hipMemcpyDeviceToHost) !=
                        hipSuccess)
      return fail(MLBA_EDEVICE, "copy failed");
  if (nfail) return fail(MLBA_ESTATE, "accumulated IMU covariance is not positive definite");
  return MLBA_OK;
}

void mlba_inertial_options_init(mlba_inertial_options* o) {
  if (!o) return;
  o->fix_gyro_bias = 0;
  o->fix_accel_bias = 0;
  o->fix_velocity = 0;
  o->reserved0 = 0;
}

// Speed and bias states of a keyframe chain with the poses constant: the
// inertial residuals of every edge (device, mlba_inertial_evaluate) and an
// LM with the Ceres rules of mlba_solve on the block-tridiagonal normal
// equations of the free [b_g | v_M | b_a] columns of each vertex.
int32_t mlba_solve_inertial_states(const mlba_inertial_problem* p, const mlba_options* o,
                                   const mlba_inertial_options* io, double* gyro_bias_out, double* velocity_out,
                                   double* accel_bias_out, mlba_summary* sm) {
  using namespace mlba;
  if (!p || !o || !io) return fail(MLBA_EINVAL, "NULL argument");
  const int P = p->num_vertices, E = p->num_edges;
  if (P <= 0 || E <= 0) return fail(MLBA_EINVAL, "no inertial edges");
  for (int e = 0; e < E; ++e)
    if (p->edge_from[e] < 0 || p->edge_from[e] >= P || p->edge_to[e] != p->edge_from[e] + 1)
      return fail(MLBA_EINVAL, "inertial edges must join consecutive vertices (a keyframe chain)");
  // Free columns of a vertex: b_g (0..2), v_M (3..5), b_a (6..8).
  int cols[9], nd = 0;
  for (int g = 0; g < 3; ++g) {
    const bool fixed = g == 0 ? io->fix_gyro_bias : (g == 1 ? io->fix_velocity : io->fix_accel_bias);
    if (!fixed)
      for (int k = 0; k < 3; ++k) cols[nd++] = 3 * g + k;
  }
  if (nd == 0) return fail(MLBA_EINVAL, "no speed or bias block is free");
  std::vector<uint8_t> in(P, 0);
  for (int e = 0; e < E; ++e) in[p->edge_from[e]] = in[p->edge_to[e]] = 1;
  std::vector<double> bg(p->vertex_gyro_bias, p->vertex_gyro_bias + 3 * P);
  std::vector<double> vel(p->vertex_velocity_M, p->vertex_velocity_M + 3 * P);
  std::vector<double> ba(p->vertex_accel_bias, p->vertex_accel_bias + 3 * P);
  // Evaluate: cost, and with A the blocks D (P x nd x nd), C (coupling of v
  // and v + 1) and g.
  std::vector<double> res(15 * (size_t)E), J[8];
  const int sizes[8] = {7, 3, 3, 3, 7, 3, 3, 3};
  auto evaluate = [&](const std::vector<double>& gb, const std::vector<double>& vv, const std::vector<double>& ab,
                      bool jac, double* cost, std::vector<double>* D, std::vector<double>* C,
                      std::vector<double>* g) -> int {
    mlba_inertial_problem q = *p;
    q.vertex_gyro_bias = gb.data();
    q.vertex_velocity_M = vv.data();
    q.vertex_accel_bias = ab.data();
    double* jp[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    if (jac)
      for (int b : {1, 2, 3, 5, 6, 7}) {
        J[b].resize(15 * (size_t)sizes[b] * E);
        jp[b] = J[b].data();
      }
    const int rc = mlba_inertial_evaluate(&q, res.data(), jac ? jp : nullptr);
    if (rc) return rc;
    double c = 0.0;
    for (double r : res) c += r * r;
    *cost = 0.5 * c;
    if (!jac) return MLBA_OK;
    D->assign((size_t)P * nd * nd, 0.0);
    C->assign((size_t)P * nd * nd, 0.0);
    g->assign((size_t)P * nd, 0.0);
    double Ja[15 * 9], Jb[15 * 9];
    for (int e = 0; e < E; ++e) {
      const int a = p->edge_from[e], b = a + 1;
      for (int r = 0; r < 15; ++r)
        for (int k = 0; k < nd; ++k) {
          const int grp = cols[k] / 3, cc = cols[k] % 3;  // parameter blocks 1..3 (from), 5..7 (to)
          Ja[r * nd + k] = J[1 + grp][(size_t)45 * e + 3 * r + cc];
          Jb[r * nd + k] = J[5 + grp][(size_t)45 * e + 3 * r + cc];
        }
      const double* re = &res[15 * (size_t)e];
      for (int i = 0; i < nd; ++i) {
        double ga = 0.0, gbv = 0.0;
        for (int r = 0; r < 15; ++r) { ga += Ja[r * nd + i] * re[r]; gbv += Jb[r * nd + i] * re[r]; }
        (*g)[(size_t)a * nd + i] += ga;
        (*g)[(size_t)b * nd + i] += gbv;
        for (int j = 0; j < nd; ++j) {
          double daa = 0.0, dbb = 0.0, dab = 0.0;
          for (int r = 0; r < 15; ++r) {
            daa += Ja[r * nd + i] * Ja[r * nd + j];
            dbb += Jb[r * nd + i] * Jb[r * nd + j];
            dab += Ja[r * nd + i] * Jb[r * nd + j];
          }
          (*D)[((size_t)a * nd + i) * nd + j] += daa;
          (*D)[((size_t)b * nd + i) * nd + j] += dbb;
          (*C)[((size_t)a * nd + i) * nd + j] += dab;
        }
      }
    }
    return MLBA_OK;
  };
  if (sm) {
    sm->num_iterations = sm->num_successful_steps = sm->num_unsuccessful_steps = 0;
    sm->num_residual_blocks = E;
    sm->num_removed_landmarks = 0;
    sm->total_linear_solver_iterations = 0;
    sm->fixed_cost = 0.0;
  }
  auto push = [&](const mlba_iteration_summary& it) {
    if (!sm) return;
    if (sm->iterations && sm->num_iterations < sm->iterations_capacity) sm->iterations[sm->num_iterations] = it;
    sm->num_iterations++;
  };
  const auto t0 = std::chrono::steady_clock::now();
  auto elapsed = [&] { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
  std::vector<double> D, C, g;
  double cost = 0.0;
  int rc = evaluate(bg, vel, ba, true, &cost, &D, &C, &g);
  if (rc) return rc;
  auto gmax = [&](const std::vector<double>& gv) {
    double m = 0.0;
    for (int v = 0; v < P; ++v)
      if (in[v])
        for (int k = 0; k < nd; ++k) m = std::max(m, std::fabs(gv[(size_t)v * nd + k]));
    return m;
  };
  double grad = gmax(g);
  if (sm) sm->initial_cost = cost;
  mlba_iteration_summary it0{};
  it0.step_is_valid = it0.step_is_successful = 1;
  it0.cost = cost;
  it0.gradient_max_norm = grad;
  double radius = o->initial_trust_region_radius, decrease = 2.0;
  it0.trust_region_radius = radius;
  push(it0);
  int term = grad <= o->gradient_tolerance ? MLBA_CONVERGENCE : MLBA_NO_CONVERGENCE;
  int invalid = 0, iteration = 0;
  std::vector<double> sc((size_t)P * nd), w((size_t)P * nd), L((size_t)P * nd * nd), B((size_t)P * nd * nd);
  while (term == MLBA_NO_CONVERGENCE && o->max_num_iterations > 0) {
    ++iteration;
    mlba_iteration_summary cur{};
    cur.iteration = iteration;
    for (int v = 0; v < P; ++v)
      for (int k = 0; k < nd; ++k) {
        const double d = D[((size_t)v * nd + k) * nd + k];
        sc[(size_t)v * nd + k] = o->use_jacobi_scaling ? 1.0 / (1.0 + std::sqrt(d)) : 1.0;
      }
    // Block-tridiagonal Cholesky of the scaled, damped system:
    // L_v L_v' = D'_v - B_{v-1}' B_{v-1}, B_v = L_v^-1 C'_v.
    bool ok = true;
    std::vector<double> y((size_t)P * nd, 0.0);
    for (int v = 0; v < P && ok; ++v) {
      if (!in[v]) continue;
      double M[81];
      for (int i = 0; i < nd; ++i)
        for (int j = 0; j < nd; ++j) {
          double m = sc[(size_t)v * nd + i] * D[((size_t)v * nd + i) * nd + j] * sc[(size_t)v * nd + j];
          if (i == j) m += std::min(std::max(m, o->min_lm_diagonal), o->max_lm_diagonal) / radius;
          M[i * nd + j] = m;
        }
      const bool prev = v > 0 && in[v - 1];
      if (prev)
        for (int i = 0; i < nd; ++i)
          for (int j = 0; j < nd; ++j) {
            double s2 = 0.0;
            for (int k = 0; k < nd; ++k) s2 += B[((size_t)(v - 1) * nd + k) * nd + i] * B[((size_t)(v - 1) * nd + k) * nd + j];
            M[i * nd + j] -= s2;
          }
      double* Lv = &L[(size_t)v * nd * nd];
      for (int j = 0; j < nd && ok; ++j) {
        double dj = M[j * nd + j];
        for (int k = 0; k < j; ++k) dj -= Lv[j * nd + k] * Lv[j * nd + k];
        if (!(dj > 0.0) || !std::isfinite(dj)) { ok = false; break; }
        Lv[j * nd + j] = std::sqrt(dj);
        for (int i = j + 1; i < nd; ++i) {
          double t = M[i * nd + j];
          for (int k = 0; k < j; ++k) t -= Lv[i * nd + k] * Lv[j * nd + k];
          Lv[i * nd + j] = t / Lv[j * nd + j];
        }
        for (int i = 0; i < j; ++i) Lv[i * nd + j] = 0.0;
      }
      if (!ok) break;
      // Forward substitution of the right-hand side.
      double rhs[9];
      for (int i = 0; i < nd; ++i) rhs[i] = sc[(size_t)v * nd + i] * g[(size_t)v * nd + i];
      if (prev)
        for (int i = 0; i < nd; ++i)
          for (int k = 0; k < nd; ++k) rhs[i] -= B[((size_t)(v - 1) * nd + k) * nd + i] * y[(size_t)(v - 1) * nd + k];
      for (int i = 0; i < nd; ++i) {
        double t = rhs[i];
        for (int k = 0; k < i; ++k) t -= Lv[i * nd + k] * y[(size_t)v * nd + k];
        y[(size_t)v * nd + i] = t / Lv[i * nd + i];
      }
      // B_v = L_v^-1 (S_v C_v S_{v+1}) when v + 1 is in the chain.
      if (v + 1 < P && in[v + 1])
        for (int j = 0; j < nd; ++j)
          for (int i = 0; i < nd; ++i) {
            double t = sc[(size_t)v * nd + i] * C[((size_t)v * nd + i) * nd + j] * sc[(size_t)(v + 1) * nd + j];
            for (int k = 0; k < i; ++k) t -= Lv[i * nd + k] * B[((size_t)v * nd + k) * nd + j];
            B[((size_t)v * nd + i) * nd + j] = t / Lv[i * nd + i];
          }
    }
    double mcc = 0.0;
    if (ok) {
      for (int v = P - 1; v >= 0; --v) {
        if (!in[v]) continue;
        const double* Lv = &L[(size_t)v * nd * nd];
        double rhs[9];
        for (int i = 0; i < nd; ++i) rhs[i] = y[(size_t)v * nd + i];
        if (v + 1 < P && in[v + 1])
          for (int i = 0; i < nd; ++i)
            for (int j = 0; j < nd; ++j) rhs[i] -= B[((size_t)v * nd + i) * nd + j] * w[(size_t)(v + 1) * nd + j];
        for (int i = nd - 1; i >= 0; --i) {
          double t = rhs[i];
          for (int k = i + 1; k < nd; ++k) t -= Lv[k * nd + i] * w[(size_t)v * nd + k];
          w[(size_t)v * nd + i] = t / Lv[i * nd + i];
        }
      }
      for (size_t k = 0; k < w.size(); ++k) w[k] *= sc[k];
      // Model cost change of the step -w: g'w - w'Aw / 2.
      double gw = 0.0, wAw = 0.0;
      for (int v = 0; v < P; ++v) {
        if (!in[v]) continue;
        for (int i = 0; i < nd; ++i) {
          gw += g[(size_t)v * nd + i] * w[(size_t)v * nd + i];
          double aw = 0.0;
          for (int j = 0; j < nd; ++j) {
            aw += D[((size_t)v * nd + i) * nd + j] * w[(size_t)v * nd + j];
            if (v + 1 < P && in[v + 1]) aw += C[((size_t)v * nd + i) * nd + j] * w[(size_t)(v + 1) * nd + j];
            if (v > 0 && in[v - 1]) aw += C[((size_t)(v - 1) * nd + j) * nd + i] * w[(size_t)(v - 1) * nd + j];
          }
          wAw += w[(size_t)v * nd + i] * aw;
        }
      }
      mcc = gw - 0.5 * wAw;
      ok = std::isfinite(mcc) && mcc > 0.0;
    }
    cur.step_is_valid = ok;
    cur.linear_solver_iterations = 1;
    if (!ok) {
      if (++invalid >= o->max_num_consecutive_invalid_steps) { term = MLBA_FAILURE; break; }
      radius /= decrease;
      decrease *= 2.0;
      cur.cost = cost;
      cur.gradient_max_norm = grad;
      cur.trust_region_radius = radius;
      if (sm) sm->num_unsuccessful_steps++;
    } else {
      invalid = 0;
      std::vector<double> s3[3] = {bg, vel, ba};
      const std::vector<double>* cur3[3] = {&bg, &vel, &ba};
      double ss = 0.0, xx = 0.0;
      for (int v = 0; v < P; ++v) {
        if (!in[v]) continue;
        for (int k = 0; k < nd; ++k) {
          const int c = cols[k];
          const double x = (*cur3[c / 3])[3 * v + c % 3];
          s3[c / 3][3 * v + c % 3] = x - w[(size_t)v * nd + k];
          ss += w[(size_t)v * nd + k] * w[(size_t)v * nd + k];
          xx += x * x;
        }
      }
      cur.step_norm = std::sqrt(ss);
      if (cur.step_norm <= o->parameter_tolerance * (std::sqrt(xx) + o->parameter_tolerance)) {
        term = MLBA_CONVERGENCE;
        break;
      }
      double ccost = 0.0;
      if ((rc = evaluate(s3[0], s3[1], s3[2], false, &ccost, nullptr, nullptr, nullptr))) return rc;
      cur.cost_change = cost - ccost;
      if (std::fabs(cur.cost_change) <= o->function_tolerance * cost) { term = MLBA_CONVERGENCE; break; }
      cur.relative_decrease = (cost - ccost) / mcc;
      if (cur.relative_decrease > o->min_relative_decrease) {
        bg = s3[0]; vel = s3[1]; ba = s3[2];
        if ((rc = evaluate(bg, vel, ba, true, &cost, &D, &C, &g))) return rc;
        grad = gmax(g);
        cur.step_is_successful = 1;
        cur.cost = cost;
        cur.gradient_max_norm = grad;
        const double qq = 2.0 * cur.relative_decrease - 1.0;
        radius = std::min(o->max_trust_region_radius, radius / std::max(1.0 / 3.0, 1.0 - qq * qq * qq));
        decrease = 2.0;
        if (sm) sm->num_successful_steps++;
      } else {
        radius /= decrease;
        decrease *= 2.0;
        cur.cost = ccost;
        cur.gradient_max_norm = grad;
        if (sm) sm->num_unsuccessful_steps++;
      }
      cur.trust_region_radius = radius;
    }
    cur.cumulative_time_in_seconds = elapsed();
    push(cur);
    if (elapsed() > o->max_solver_time_in_seconds) break;
    if (iteration >= o->max_num_iterations) break;
    if (cur.step_is_successful && cur.gradient_max_norm <= o->gradient_tolerance) { term = MLBA_CONVERGENCE; break; }
    if (radius <= o->min_trust_region_radius) { term = MLBA_CONVERGENCE; break; }
  }
  if (sm) {
    sm->termination_type = term;
    sm->final_cost = cost;
    sm->lm_loop_seconds = elapsed();
  }
  if (gyro_bias_out) std::memcpy(gyro_bias_out, bg.data(), sizeof(double) * 3 * P);
  if (velocity_out) std::memcpy(velocity_out, vel.data(), sizeof(double) * 3 * P);
  if (accel_bias_out) std::memcpy(accel_bias_out, ba.data(), sizeof(double) * 3 * P);
  return MLBA_OK;
}

}  // extern "C"
