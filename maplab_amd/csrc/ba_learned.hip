// ba_learned.hip -- learned block correction of the block-Jacobi
// preconditioner of the reduced camera system and the PCG that applies it
// (interface and the mathematics in ba_learned.h). No counterpart in the
// reference: maplab's Ceres ITERATIVE_SCHUR uses the fixed SCHUR_JACOBI
// preconditioner (solver-options.h:57-68); this is the preconditioner the
// device solve can use in its place, with plain block-Jacobi as fallback.
//
// Kernels:
//   k_lp_factor      Cholesky of every diagonal block and its inverse factor (FP64)
//   k_lp_normalize   Sh_ij = L_i^-1 S_ij L_j^-T of every block position (FP32 out)
//   k_lp_features    node and edge features of the network (normalized)
//   k_lp_terms       per row: sum over the lower neighbourhood of the left
//                    factor times Sh_kj (T2 from Sh, T3 from T2), one work-group
//                    per row with the row's lower neighbour map and left blocks in LDS
//   k_lpcg_*         PCG: S p, scalar reductions, x/r update with L^-1 r,
//                    (I - K) row sweep, (I - K)' row sweep with L^-T, all
//                    reductions fixed-order; a device flag ends the solve.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "ba_learned.h"
#include "ba_reduce.h"

namespace mlba {

namespace {

using ::mlba::wave_sum_f;
__device__ inline double wave_sum_d(double v) { return wave_sum(v); }

// ------------------------------------------------------------ set-up

__global__ void k_lp_factor(LpSystem sys, double* __restrict__ linv, int32_t* __restrict__ fail) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sys.Pvar) return;
  const double* A = sys.S + 36 * (size_t)sys.diag_pos[i];
  double L[36];
  for (int k = 0; k < 36; ++k) L[k] = 0.0;
  bool ok = true;
  for (int r = 0; r < 6; ++r) {
    for (int c = 0; c <= r; ++c) {
      double s = 0.5 * (A[6 * r + c] + A[6 * c + r]);
      for (int k = 0; k < c; ++k) s -= L[6 * r + k] * L[6 * c + k];
      if (r == c) {
        if (!(s > 0.0)) { ok = false; s = 1.0; }
        L[6 * r + r] = sqrt(s);
      } else {
        L[6 * r + c] = s / L[6 * c + c];
      }
    }
  }
  // Inverse of the lower-triangular factor, column by column.
  double X[36];
  for (int k = 0; k < 36; ++k) X[k] = 0.0;
  for (int c = 0; c < 6; ++c) {
    for (int r = c; r < 6; ++r) {
      double s = (r == c) ? 1.0 : 0.0;
      for (int k = c; k < r; ++k) s -= L[6 * r + k] * X[6 * k + c];
      X[6 * r + c] = s / L[6 * r + r];
    }
  }
  double* out = linv + 36 * (size_t)i;
  for (int k = 0; k < 36; ++k) out[k] = X[k];
  if (!ok) atomicAdd(fail, 1);
}

__global__ void k_lp_normalize(LpSystem sys, const double* __restrict__ linv, float* __restrict__ shat,
                               float* __restrict__ snorm) {
  const int pos = blockIdx.x * blockDim.x + threadIdx.x;
  if (pos >= sys.nnzb) return;
  const int i = sys.pos_row[pos], j = sys.col_var[pos];
  float* out = shat + 36 * (size_t)pos;
  if (i == j) {
    for (int k = 0; k < 36; ++k) out[k] = (k % 7 == 0) ? 1.0f : 0.0f;
    snorm[pos] = 0.0f;
    return;
  }
  const double* Li = linv + 36 * (size_t)i;
  const double* Lj = linv + 36 * (size_t)j;
  const double* B = sys.S + 36 * (size_t)pos;
  double A[36];
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 6; ++c) {
      double s = 0.0;
      for (int m = 0; m <= r; ++m) s += Li[6 * r + m] * B[6 * m + c];
      A[6 * r + c] = s;
    }
  double nrm = 0.0;
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 6; ++c) {
      double s = 0.0;
      for (int m = 0; m <= c; ++m) s += A[6 * r + m] * Lj[6 * c + m];
      out[6 * r + c] = (float)s;
      nrm += s * s;
    }
  snorm[pos] = (float)sqrt(nrm);
}

__global__ void k_lp_node_features(LpSystem sys, const double* __restrict__ ujtj, const double* __restrict__ scale,
                                   const float* __restrict__ norm, float* __restrict__ node_feat) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sys.Pvar) return;
  float f[kLpNodeFeatures];
  const double* U = ujtj + 36 * (size_t)i;
  const double* A = sys.S + 36 * (size_t)sys.diag_pos[i];
  double tr = 0.0;
  for (int c = 0; c < 6; ++c) {
    const double sc = scale[6 * (size_t)i + c];
    f[c] = (float)log(fmax(U[7 * c] * sc * sc, 1e-30));
    tr += A[7 * c];
  }
  f[6] = (float)log(fmax(tr, 1e-30));
  f[7] = (float)log1p((double)(sys.row_ptr[i + 1] - sys.row_ptr[i] - 1));
  f[8] = (float)log1p((double)(sys.diag_pos[i] - sys.row_ptr[i]));
  for (int k = 0; k < kLpNodeFeatures; ++k)
    node_feat[(size_t)kLpNodeFeatures * i + k] = (f[k] - norm[k]) / norm[kLpNodeFeatures + k];
}

__global__ void k_lp_edge_features(LpSystem sys, const float* __restrict__ snorm, const int32_t* __restrict__ shared,
                                   const float* __restrict__ norm, float* __restrict__ edge_feat) {
  const int pos = blockIdx.x * blockDim.x + threadIdx.x;
  if (pos >= sys.nnzb) return;
  const int i = sys.pos_row[pos], j = sys.col_var[pos];
  float* out = edge_feat + (size_t)kLpEdgeFeatures * pos;
  if (i == j) {
    for (int k = 0; k < kLpEdgeFeatures; ++k) out[k] = 0.0f;
    return;
  }
  const float* en = norm + 2 * kLpNodeFeatures;
  const float f[kLpEdgeFeatures] = {(float)log1p((double)shared[pos]), snorm[pos], j < i ? 1.0f : -1.0f};
  for (int k = 0; k < kLpEdgeFeatures; ++k) out[k] = (f[k] - en[k]) / en[kLpEdgeFeatures + k];
}

// One work-group (4 waves) per row i. LDS: map from vertex to its index in
// the row's lower neighbour list (int16, -1 elsewhere), then the row's left
// blocks (Sh_iP, or T2_iP for the T3 traces).
template <int kMode>
__global__ __launch_bounds__(256) void k_lp_terms(LpSystem sys, const float* __restrict__ shat,
                                                  const float* __restrict__ theta, float* __restrict__ kmat,
                                                  float* __restrict__ t2, float* __restrict__ trace,
                                                  int32_t left_offset) {
  extern __shared__ __align__(16) char lds[];
  int16_t* map = reinterpret_cast<int16_t*>(lds);
  float* left = reinterpret_cast<float*>(lds + left_offset);
  __shared__ float red[4][4];
  const int i = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int b = sys.row_ptr[i], nP = sys.diag_pos[i] - b;
  for (int v = tid; v < sys.Pvar; v += 256) map[v] = -1;
  __syncthreads();
  for (int k = tid; k < nP; k += 256) map[sys.col_var[b + k]] = (int16_t)k;
  const float* lsrc = (kMode == kLpTraceT3) ? t2 : shat;
  for (int e = tid; e < 36 * nP; e += 256) left[e] = lsrc[36 * (size_t)b + e];
  __syncthreads();
  float th1 = 0.0f, th2 = 0.0f;
  if (kMode == kLpAssemble) { th1 = theta[kLpTerms * i]; th2 = theta[kLpTerms * i + 1]; }
  float ta = 0.0f, tb = 0.0f, te = 0.0f, td = 0.0f;
  for (int jj = wave; jj < nP; jj += 4) {
    const int j = sys.col_var[b + jj];
    float acc[36];
    for (int k = 0; k < 36; ++k) acc[k] = 0.0f;
    for (int q = sys.row_ptr[j] + lane; q < sys.row_ptr[j + 1]; q += 64) {
      const int k = sys.col_var[q];
      if (k == j) continue;
      const int m = map[k];
      if (m < 0) continue;
      // Sh_kj = Sh_jk', Sh_jk at position q of row j.
      float B[36];
      const float4* bq = reinterpret_cast<const float4*>(shat + 36 * (size_t)q);
      for (int v = 0; v < 9; ++v) {
        const float4 x = bq[v];
        B[4 * v] = x.x; B[4 * v + 1] = x.y; B[4 * v + 2] = x.z; B[4 * v + 3] = x.w;
      }
      const float* A = left + 36 * m;
      for (int r = 0; r < 6; ++r)
        for (int mm = 0; mm < 6; ++mm) {
          const float a = A[6 * r + mm];
          for (int c = 0; c < 6; ++c) acc[6 * r + c] += a * B[6 * c + mm];
        }
    }
    float val = 0.0f;
    for (int k = 0; k < 36; ++k) {
      const float s = wave_sum_f(acc[k]);
      if (lane == k) val = s;
    }
    if (lane < 36) {
      const size_t at = 36 * (size_t)(b + jj) + lane;
      if (kMode == kLpAssemble) {
        kmat[at] = th1 * shat[at] + th2 * val;
      } else if (kMode == kLpStoreT2) {
        const float s1 = shat[at];
        t2[at] = val;
        ta += s1 * s1;
        tb += s1 * val;
        te += val * val;
      } else {
        td += t2[at] * val;
      }
    }
  }
  if (kMode == kLpAssemble) return;
  ta = wave_sum_f(ta); tb = wave_sum_f(tb); te = wave_sum_f(te); td = wave_sum_f(td);
  if (lane == 0) { red[wave][0] = ta; red[wave][1] = tb; red[wave][2] = te; red[wave][3] = td; }
  __syncthreads();
  if (tid < 4) {
    const float s = (red[0][tid] + red[1][tid]) + (red[2][tid] + red[3][tid]);
    if (kMode == kLpStoreT2 && tid < 3) trace[4 * i + tid] = s;
    if (kMode == kLpTraceT3 && tid == 3) trace[4 * i + 3] = s;
  }
}

// ------------------------------------------------------------ PCG

struct LpPcgState {
  double rz, alpha, beta, rr, bb, tol2;
  int32_t iter, done, converged, max_iters;
};

// Scalar step, one work-group: mode 0 after the set-up, 1 after S p, 2 after
// the preconditioner. part: [Pvar][2], slot 0 r'r (mode 0: b'b) or p'q
// (mode 1), slot 1 r'z.
__global__ __launch_bounds__(256) void k_lpcg_reduce(int32_t Pvar, int32_t mode, const double* __restrict__ part,
                                                     LpPcgState* __restrict__ st) {
  __shared__ double red[2][4];
  if (st->done) return;
  double s0 = 0.0, s1 = 0.0;
  for (int i = threadIdx.x; i < Pvar; i += 256) {
    s0 += part[2 * i];
    s1 += part[2 * i + 1];
  }
  s0 = wave_sum_d(s0);
  s1 = wave_sum_d(s1);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) { red[0][wave] = s0; red[1][wave] = s1; }
  __syncthreads();
  if (threadIdx.x != 0) return;
  const double a = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
  const double c = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
  if (mode == 0) {
    st->bb = a;
    st->rr = a;
    st->rz = c;
    st->beta = 0.0;
    st->iter = 0;
    if (!(a > 0.0)) { st->done = 1; st->converged = 1; }
    else if (!(c > 0.0) || !isfinite(c)) { st->done = 1; st->converged = 0; }
  } else if (mode == 1) {
    if (!(a > 0.0) || !isfinite(a)) { st->done = 1; st->converged = 0; return; }
    st->alpha = st->rz / a;
  } else {
    st->iter += 1;
    st->rr = a;
    if (a <= st->tol2 * st->bb) { st->done = 1; st->converged = 1; return; }
    if (!(c > 0.0) || !isfinite(c)) { st->done = 1; st->converged = 0; return; }
    st->beta = c / st->rz;
    st->rz = c;
    if (st->iter >= st->max_iters) { st->done = 1; st->converged = 0; }
  }
}

// x = 0, r = b, t = L^-1 b, p buffers zeroed; part slot 0 = b'b of the row.
// With no K: z = L^-T t and slot 1 = r'z.
__global__ void k_lpcg_init(LpPcgBuffers B, int32_t use_k) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B.sys.Pvar) return;
  const double* Li = B.linv + 36 * (size_t)i;
  double r[6], t[6];
  double bb = 0.0;
  for (int c = 0; c < 6; ++c) {
    r[c] = B.b[6 * i + c];
    bb += r[c] * r[c];
    B.x[6 * i + c] = 0.0;
    B.r[6 * i + c] = r[c];
    B.p0[6 * i + c] = 0.0;
    B.p1[6 * i + c] = 0.0;
  }
  for (int a = 0; a < 6; ++a) {
    double s = 0.0;
    for (int m = 0; m <= a; ++m) s += Li[6 * a + m] * r[m];
    t[a] = s;
    B.t[6 * i + a] = s;
  }
  B.part[2 * i] = bb;
  if (!use_k) {
    double rz = 0.0;
    for (int c = 0; c < 6; ++c) {
      double s = 0.0;
      for (int a = c; a < 6; ++a) s += Li[6 * a + c] * t[a];
      B.z[6 * i + c] = s;
      rz += r[c] * s;
    }
    B.part[2 * i + 1] = rz;
  }
}

// q_i = sum_j S_ij (z_j + beta p_j^old), p_i^new = z_i + beta p_i^old, part = p'q.
__global__ __launch_bounds__(256) void k_lpcg_spmv(LpPcgBuffers B, const double* __restrict__ pold,
                                                   double* __restrict__ pnew, const LpPcgState* __restrict__ st) {
  __shared__ double red[42][6];
  if (st->done) return;
  const int i = blockIdx.x, tid = threadIdx.x;
  const double beta = st->beta;
  const int slot = tid / 6, r = tid % 6;
  double acc = 0.0;
  if (slot < 42) {
    for (int pos = B.sys.row_ptr[i] + slot; pos < B.sys.row_ptr[i + 1]; pos += 42) {
      const int j = B.sys.col_var[pos];
      const double* A = B.sys.S + 36 * (size_t)pos + 6 * r;
      const double* zj = B.z + 6 * (size_t)j;
      const double* pj = pold + 6 * (size_t)j;
      for (int c = 0; c < 6; ++c) acc += A[c] * (zj[c] + beta * pj[c]);
    }
    red[slot][r] = acc;
  }
  __syncthreads();
  if (tid < 6) {
    double s = 0.0;
    for (int k = 0; k < 42; ++k) s += red[k][tid];
    const double pn = B.z[6 * i + tid] + beta * pold[6 * i + tid];
    pnew[6 * i + tid] = pn;
    B.q[6 * i + tid] = s;
    red[0][tid] = pn * s;
  }
  __syncthreads();
  if (tid == 0) {
    B.part[2 * i] = ((red[0][0] + red[0][1]) + (red[0][2] + red[0][3])) + (red[0][4] + red[0][5]);
    B.part[2 * i + 1] = 0.0;
  }
}

// x += alpha p, r -= alpha q, t = L^-1 r; part slot 0 = r'r. With no K:
// z = L^-T t and slot 1 = r'z.
__global__ void k_lpcg_update(LpPcgBuffers B, const double* __restrict__ p, int32_t use_k,
                              const LpPcgState* __restrict__ st) {
  if (st->done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B.sys.Pvar) return;
  const double alpha = st->alpha;
  const double* Li = B.linv + 36 * (size_t)i;
  double r[6], t[6];
  double rr = 0.0;
  for (int c = 0; c < 6; ++c) {
    B.x[6 * i + c] += alpha * p[6 * i + c];
    r[c] = B.r[6 * i + c] - alpha * B.q[6 * i + c];
    B.r[6 * i + c] = r[c];
    rr += r[c] * r[c];
  }
  for (int a = 0; a < 6; ++a) {
    double s = 0.0;
    for (int m = 0; m <= a; ++m) s += Li[6 * a + m] * r[m];
    t[a] = s;
    B.t[6 * i + a] = s;
  }
  B.part[2 * i] = rr;
  if (!use_k) {
    double rz = 0.0;
    for (int c = 0; c < 6; ++c) {
      double s = 0.0;
      for (int a = c; a < 6; ++a) s += Li[6 * a + c] * t[a];
      B.z[6 * i + c] = s;
      rz += r[c] * s;
    }
    B.part[2 * i + 1] = rz;
  }
}

// u_i = t_i - sum_{j < i} K_ij t_j; one wave per row.
__global__ __launch_bounds__(64) void k_lpcg_lower(LpPcgBuffers B, const LpPcgState* __restrict__ st) {
  if (st->done) return;
  const int i = blockIdx.x, lane = threadIdx.x;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int pos = B.sys.row_ptr[i] + lane; pos < B.sys.diag_pos[i]; pos += 64) {
    const int j = B.sys.col_var[pos];
    const float* K = B.kmat + 36 * (size_t)pos;
    const double* tj = B.t + 6 * (size_t)j;
    for (int r = 0; r < 6; ++r) {
      double s = 0.0;
      for (int c = 0; c < 6; ++c) s += (double)K[6 * r + c] * tj[c];
      acc[r] += s;
    }
  }
  for (int r = 0; r < 6; ++r) acc[r] = wave_sum_d(acc[r]);
  if (lane < 6) {
    double v = acc[0];
    for (int r = 1; r < 6; ++r) if (lane == r) v = acc[r];
    B.u[6 * i + lane] = B.t[6 * i + lane] - v;
  }
}

// w_j = u_j - sum_{i > j} K_ij' u_i, z_j = L_j^-T w_j, part slot 1 = r_j'z_j; one wave per row.
__global__ __launch_bounds__(64) void k_lpcg_upper(LpPcgBuffers B, const LpPcgState* __restrict__ st) {
  if (st->done) return;
  const int j = blockIdx.x, lane = threadIdx.x;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int pos = B.sys.diag_pos[j] + 1 + lane; pos < B.sys.row_ptr[j + 1]; pos += 64) {
    const int i = B.sys.col_var[pos];
    const float* K = B.kmat + 36 * (size_t)B.sys.tpos[pos];  // K_ij, stored in row i
    const double* ui = B.u + 6 * (size_t)i;
    for (int c = 0; c < 6; ++c) {
      double s = 0.0;
      for (int r = 0; r < 6; ++r) s += (double)K[6 * r + c] * ui[r];
      acc[c] += s;
    }
  }
  for (int c = 0; c < 6; ++c) acc[c] = wave_sum_d(acc[c]);
  __shared__ double w[6];
  if (lane < 6) {
    double v = acc[0];
    for (int c = 1; c < 6; ++c) if (lane == c) v = acc[c];
    w[lane] = B.u[6 * j + lane] - v;
  }
  __syncthreads();
  if (lane == 0) {
    const double* Lj = B.linv + 36 * (size_t)j;
    double rz = 0.0;
    for (int c = 0; c < 6; ++c) {
      double s = 0.0;
      for (int a = c; a < 6; ++a) s += Lj[6 * a + c] * w[a];
      B.z[6 * j + c] = s;
      rz += B.r[6 * j + c] * s;
    }
    B.part[2 * j + 1] = rz;
  }
}

}  // namespace

void launch_lp_factor(hipStream_t s, const LpSystem& sys, double* linv, int32_t* fail) {
  if (sys.Pvar > 0) k_lp_factor<<<(sys.Pvar + 63) / 64, 64, 0, s>>>(sys, linv, fail);
}

void launch_lp_normalize(hipStream_t s, const LpSystem& sys, const double* linv, float* shat, float* snorm) {
  if (sys.nnzb > 0) k_lp_normalize<<<(sys.nnzb + 127) / 128, 128, 0, s>>>(sys, linv, shat, snorm);
}

void launch_lp_features(hipStream_t s, const LpSystem& sys, const double* ujtj, const double* scale,
                        const float* snorm, const int32_t* shared, const float* norm, float* node_feat,
                        float* edge_feat) {
  if (sys.Pvar > 0) k_lp_node_features<<<(sys.Pvar + 127) / 128, 128, 0, s>>>(sys, ujtj, scale, norm, node_feat);
  if (sys.nnzb > 0)
    k_lp_edge_features<<<(sys.nnzb + 127) / 128, 128, 0, s>>>(sys, snorm, shared, norm, edge_feat);
}

bool launch_lp_terms(hipStream_t s, const LpSystem& sys, int32_t max_lower_deg, LpMode mode, const float* shat,
                     const float* theta, float* kmat, float* t2, float* trace) {
  if (sys.Pvar <= 0) return true;
  if (sys.Pvar >= 32768) return false;
  const int left_offset = ((sys.Pvar * 2 + 15) / 16) * 16;
  const size_t bytes = (size_t)left_offset + 36 * sizeof(float) * (size_t)max_lower_deg;
  if (bytes > 60 * 1024) return false;
  switch (mode) {
    case kLpAssemble:
      k_lp_terms<kLpAssemble><<<sys.Pvar, 256, bytes, s>>>(sys, shat, theta, kmat, t2, trace, left_offset);
      break;
    case kLpStoreT2:
      k_lp_terms<kLpStoreT2><<<sys.Pvar, 256, bytes, s>>>(sys, shat, theta, kmat, t2, trace, left_offset);
      break;
    case kLpTraceT3:
      k_lp_terms<kLpTraceT3><<<sys.Pvar, 256, bytes, s>>>(sys, shat, theta, kmat, t2, trace, left_offset);
      break;
  }
  return true;
}

size_t lp_pcg_state_bytes() { return sizeof(LpPcgState); }

int launch_lp_pcg(hipStream_t s, const LpPcgBuffers& B, double tol, int32_t max_iters, int32_t poll,
                  int32_t* converged) {
  const int P = B.sys.Pvar;
  *converged = 1;
  if (P <= 0) return 0;
  const int use_k = B.kmat != nullptr;
  LpPcgState init{};
  init.tol2 = tol * tol;
  init.max_iters = max_iters;
  LpPcgState* st = static_cast<LpPcgState*>(B.state_dev);
  LpPcgState* host = static_cast<LpPcgState*>(B.state_host);
  *host = init;
  if (hipMemcpyAsync(st, host, sizeof(LpPcgState), hipMemcpyHostToDevice, s) != hipSuccess) return -1;
  const int tb = (P + 127) / 128;
  k_lpcg_init<<<tb, 128, 0, s>>>(B, use_k);
  if (use_k) {
    k_lpcg_lower<<<P, 64, 0, s>>>(B, st);
    k_lpcg_upper<<<P, 64, 0, s>>>(B, st);
  }
  k_lpcg_reduce<<<1, 256, 0, s>>>(P, 0, B.part, st);
  if (hipGetLastError() != hipSuccess) return -1;
  int launched = 0;
  poll = poll > 0 ? poll : 8;
  while (true) {
    const int n = std::min(poll, max_iters - launched);
    for (int k = 0; k < n; ++k, ++launched) {
      const double* pold = (launched & 1) ? B.p0 : B.p1;
      double* pnew = (launched & 1) ? B.p1 : B.p0;
      k_lpcg_spmv<<<P, 256, 0, s>>>(B, pold, pnew, st);
      k_lpcg_reduce<<<1, 256, 0, s>>>(P, 1, B.part, st);
      k_lpcg_update<<<tb, 128, 0, s>>>(B, pnew, use_k, st);
      if (use_k) {
        k_lpcg_lower<<<P, 64, 0, s>>>(B, st);
        k_lpcg_upper<<<P, 64, 0, s>>>(B, st);
      }
      k_lpcg_reduce<<<1, 256, 0, s>>>(P, 2, B.part, st);
    }
    if (hipGetLastError() != hipSuccess) return -1;
    if (hipMemcpyAsync(host, st, sizeof(LpPcgState), hipMemcpyDeviceToHost, s) != hipSuccess) return -1;
    if (hipStreamSynchronize(s) != hipSuccess) return -1;
    if (host->done || launched >= max_iters) break;
  }
  *converged = host->done ? host->converged : 0;
  return host->iter;
}

}  // namespace mlba
