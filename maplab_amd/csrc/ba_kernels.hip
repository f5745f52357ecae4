// ba_kernels.hip -- gfx950 kernels of the visual bundle-adjustment path.
//
// One LM iteration on the device (names match the launchers at the bottom):
//   k_vertex_prep     R_I_M and |q|^2 of every vertex (state or candidate)
//   k_linearize<1>    per observation: residual, tangent Jacobians, robust
//                     correction (VisualReprojectionError::Evaluate + ceres
//                     Corrector); per landmark: V, g reductions in LDS
//   k_linearize<0>    the same chain, cost only (candidate evaluation)
//   k_pose_rows       per variable vertex: unscaled gradient, diag(F'F) and
//                     the direct observation block sum Jo^T Jo
//   k_landmark_prep   per landmark: Z = C (C V C + D^2)^-1 C (Schur pivot)
//   k_schur_factors   per row incidence, per linearization: W_x, r0_x (and
//                     the anchor maps A_l)
//   k_schur_incidences per row incidence, per LM iteration: Q_x and rhs_x
//   k_schur_items     partial sums of the upper blocks of the reduced camera
//                     system S over work items of <= 512 incidence pairs
//   k_schur_finalize  S blocks from the item partials, both triangles
//   k_schur_rhs       rhs of every S row
//   k_block_inverse   block-Jacobi preconditioner (6x6 inverses of diag(S))
//   (ba_pcg.hip)      deflated block-Jacobi PCG on S
//   k_back_substitute landmark step and the LM model cost change
//   k_candidate       x (+) step for vertices (JPL Plus) and landmarks
//   k_gradient_norm   max |x - Plus(x, -g)|
//   k_evaluate_full   every reference Jacobian block (Evaluate parity API)
// All reductions are fixed-order (per-tile partials summed by one block,
// S blocks by one wave each, the tile sums of a landmark split over several
// linearization tiles added in tile order), so results repeat bit for bit
// run to run.
#include "ba_device.h"
#include "ba_math.h"
#include "ba_reduce.h"
#include "ba_tiles.h"

namespace mlba {

// ------------------------------------------------------------ utilities

// Inverse of a symmetric positive definite 3x3 matrix given as
// (xx xy xz yy yz zz); returns false when not positive definite.
__device__ inline bool spd3_inverse(const double* a, double* out) {
  const double l00 = sqrt(a[0]);
  if (!(l00 > 0.0)) return false;
  const double l10 = a[1] / l00, l20 = a[2] / l00;
  const double d11 = a[3] - l10 * l10;
  if (!(d11 > 0.0)) return false;
  const double l11 = sqrt(d11);
  const double l21 = (a[4] - l20 * l10) / l11;
  const double d22 = a[5] - l20 * l20 - l21 * l21;
  if (!(d22 > 0.0)) return false;
  const double l22 = sqrt(d22);
  const double i00 = 1.0 / l00, i11 = 1.0 / l11, i22 = 1.0 / l22;
  const double i10 = -l10 * i00 * i11;
  const double i21 = -l21 * i11 * i22;
  const double i20 = -(l20 * i00 + l21 * i10) * i22;
  // (L^-1)^T L^-1, L^-1 lower with rows (i00), (i10 i11), (i20 i21 i22).
  out[0] = i00 * i00 + i10 * i10 + i20 * i20;
  out[1] = i10 * i11 + i20 * i21;
  out[2] = i20 * i22;
  out[3] = i11 * i11 + i21 * i21;
  out[4] = i21 * i22;
  out[5] = i22 * i22;
  return true;
}

__device__ inline void sym3_full(const double* s, double* m) {
  m[0] = s[0]; m[1] = s[1]; m[2] = s[2];
  m[3] = s[1]; m[4] = s[3]; m[5] = s[4];
  m[6] = s[2]; m[7] = s[4]; m[8] = s[5];
}

// Anchor map M = [-|q_a|^2 [p_B]x , R_B_LM] (3x6): the anchor-vertex
// tangent Jacobian of an observation is J_landmark * M
// (visual-error-term-inl.h:199-224 with the JPL lift).
__device__ inline void anchor_map(const double* pB, const double* Ra, double n2, double* M) {
  // -n2 [p]x = -n2 [[0,-z,y],[z,0,-x],[-y,x,0]]
  M[0] = 0.0;         M[1] = n2 * pB[2];  M[2] = -n2 * pB[1];
  M[6] = -n2 * pB[2]; M[7] = 0.0;         M[8] = n2 * pB[0];
  M[12] = n2 * pB[1]; M[13] = -n2 * pB[0]; M[14] = 0.0;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) M[r * 6 + 3 + c] = Ra[r * 3 + c];
}

// ------------------------------------------------------------ vertices

// The 12-double rotation record of a vertex: R_I_M (JPL, 9), |q|^2, 0, 0.
__device__ inline void vertex_rotation(const double* q, double* __restrict__ out) {
  double R[9];
  rot_jpl(q, R);
  for (int i = 0; i < 9; ++i) out[i] = R[i];
  out[9] = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
  out[10] = 0.0;
  out[11] = 0.0;
}

__global__ void k_vertex_prep(int32_t P, const double* __restrict__ vert, double* __restrict__ vR) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= P) return;
  vertex_rotation(vert + 7 * v, vR + 12 * v);
}

// ------------------------------------------------------------ linearization

struct ObsEval {
  double r[2];
  double Jl[6];   // d r / d p_B, row-major 2x3
  double Jo[12];  // observer tangent [rot 3 | pos 3], row-major 2x6
  double rho0;
  bool failed;
};

// Residual and Jacobians of one observation (visual-error-term-inl.h:13-379,
// evaluated directly in the tangent space of the observer pose).
template <bool kJac>
__device__ inline void eval_observation(const DevTopology& t, int o, const double* __restrict__ vert,
                                        const double* __restrict__ vR, const double* __restrict__ lm,
                                        ObsEval& e) {
  const int32_t meta = t.obs_meta[o];
  const int32_t type = obs_type(meta);
  const DevCamera& cam = t.cams[obs_camera(meta)];
  const int32_t l = t.obs_landmark[o];
  const int32_t v = t.obs_vertex[o];
  const double* pB = lm + 3 * l;
  double pI[3];
  double G[9];  // d p_I / d p_B
  const double* Rv = vR + 12 * v;
  if (type == kTypeKeyframe) {
    pI[0] = pB[0]; pI[1] = pB[1]; pI[2] = pB[2];
  } else {
    const int32_t a = t.lm_anchor[l];
    const double* Ra = vR + 12 * a;        // R_B_LM
    const double* pa = vert + 7 * a + 4;   // p_LM_B
    double pM[3];
    mat3t_vec(Ra, pB, pM);
    pM[0] += pa[0]; pM[1] += pa[1]; pM[2] += pa[2];
    double H[9];  // d p_M / d p_B
    if (kJac) for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) H[i * 3 + j] = Ra[j * 3 + i];
    if (type == kTypeGlobal) {
      const double* TL = t.mission_T + 12 * t.vertex_mission[a];
      const double* TM = t.mission_T + 12 * t.vertex_mission[v];
      double pG[3], d[3];
      mat3_vec(TL, pM, pG);
      for (int i = 0; i < 3; ++i) d[i] = pG[i] + TL[9 + i] - TM[9 + i];
      mat3t_vec(TM, d, pM);
      if (kJac) {
        double T1[9], T2[9];
        mat3_mul(TL, H, T1);  // R_G_LM R_LM_B
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j)
            T2[i * 3 + j] = TM[i] * T1[j] + TM[3 + i] * T1[3 + j] + TM[6 + i] * T1[6 + j];
        for (int i = 0; i < 9; ++i) H[i] = T2[i];
      }
    }
    const double* pv = vert + 7 * v + 4;
    const double dm[3] = {pM[0] - pv[0], pM[1] - pv[1], pM[2] - pv[2]};
    mat3_vec(Rv, dm, pI);
    if (kJac) mat3_mul(Rv, H, G);
  }
  double pC[3];
  mat3_vec(cam.R_C_I, pI, pC);
  pC[0] += cam.p_C_I[0]; pC[1] += cam.p_C_I[1]; pC[2] += cam.p_C_I[2];
  double kp[2], jp[6];
  const int32_t status = project_camera(cam.model, cam.xi, cam.dist_type, cam.intr, cam.dist, cam.width, cam.height, pC, kp, jp);
  e.failed = projection_failed(status, pC, kp);
  const double sinv = t.obs_sinv[o];
  const double* meas = t.obs_kp + 2 * o;
  if (e.failed) {
    e.r[0] = e.r[1] = 0.0;
  } else {
    e.r[0] = (kp[0] - meas[0]) * sinv;
    e.r[1] = (kp[1] - meas[1]) * sinv;
  }
  if (!kJac) return;
  if (e.failed) {
    for (int i = 0; i < 6; ++i) e.Jl[i] = 0.0;
    for (int i = 0; i < 12; ++i) e.Jo[i] = 0.0;
    return;
  }
  // K = (d kp / d p_C) * sigma^-1 * R_C_I : d r / d p_I (2x3)
  double K[6];
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 3; ++c)
      K[r * 3 + c] = sinv * (jp[r * 3] * cam.R_C_I[c] + jp[r * 3 + 1] * cam.R_C_I[3 + c] +
                             jp[r * 3 + 2] * cam.R_C_I[6 + c]);
  if (type == kTypeKeyframe) {
    for (int i = 0; i < 6; ++i) e.Jl[i] = K[i];
    for (int i = 0; i < 12; ++i) e.Jo[i] = 0.0;
    return;
  }
  for (int r = 0; r < 2; ++r) {
    const double* k = K + 3 * r;
    for (int c = 0; c < 3; ++c) e.Jl[r * 3 + c] = k[0] * G[c] + k[1] * G[3 + c] + k[2] * G[6 + c];
    // rotation: |q|^2 * (K_r x p_I); position: -K_r R_I_M
    const double n2 = Rv[9];
    e.Jo[r * 6 + 0] = n2 * (k[1] * pI[2] - k[2] * pI[1]);
    e.Jo[r * 6 + 1] = n2 * (k[2] * pI[0] - k[0] * pI[2]);
    e.Jo[r * 6 + 2] = n2 * (k[0] * pI[1] - k[1] * pI[0]);
    for (int c = 0; c < 3; ++c) e.Jo[r * 6 + 3 + c] = -(k[0] * Rv[c] + k[1] * Rv[3 + c] + k[2] * Rv[6 + c]);
  }
}

// Loss scale of one observation: HuberLoss(delta * sigma) with delta 3 / 10
// (optimization-terms-addition.cc:189-198), Cauchy(scale_px / sigma).
__device__ inline double loss_scale(const DevLoss& loss, int32_t type, double sinv) {
  if (loss.type == 0) return (type == kTypeGlobal ? 10.0 : 3.0) / sinv;
  if (loss.type == 1) return loss.scale_px * sinv;
  return 1.0;
}

// Per observation: robust-corrected residual and Jacobians into obs_rec;
// per landmark: V = sum Jl'Jl, V_mission (pose-coupled observations only),
// g = sum Jl'r, g_mission; per tile: cost and fixed cost.
//
// With kF32 the Jacobians are rounded to FP32 right after the correction and
// every later product (V, g here; J'J, W, the model cost in the readers)
// uses the rounded values, so S stays the Schur complement of J'J of one
// Jacobian; records are Rec32 (residual in FP64).
template <bool kJac, bool kF32 = false>
__global__ __launch_bounds__(kObsTile) void k_linearize(DevTopology t, const double* __restrict__ vert,
                                                         const double* __restrict__ vR,
                                                         const double* __restrict__ lm, DevLoss loss,
                                                         const int32_t* __restrict__ tile_begin,
                                                         void* __restrict__ obs_rec,
                                                         double* __restrict__ lm_acc,
                                                         double* __restrict__ split_part,
                                                         double* __restrict__ tile_cost) {
  __shared__ double red[kJac ? kObsTile * kLmAcc : 1];
  __shared__ int32_t seg[kJac ? kObsTile : 1];
  __shared__ double lds4[4];
  const int tid = threadIdx.x;
  const int b = tile_begin[blockIdx.x], e = tile_begin[blockIdx.x + 1];
  const int o = b + tid;
  double acc[kLmAcc];
  for (int i = 0; i < kLmAcc; ++i) acc[i] = 0.0;
  double cost = 0.0, fixed = 0.0;
  int32_t l = -1;
  if (o < e) {
    l = t.obs_landmark[o];
    const uint32_t flags = t.lm_flags[l];
    ObsEval ev;
    if (flags & kLmActive) {
      eval_observation<kJac>(t, o, vert, vR, lm, ev);
      const int32_t meta = t.obs_meta[o];
      const int32_t type = obs_type(meta);
      const double s = ev.r[0] * ev.r[0] + ev.r[1] * ev.r[1];
      double rho[3];
      robust_loss(loss.type, loss_scale(loss, type, t.obs_sinv[o]), s, rho);
      const bool pose_coupled = type != kTypeKeyframe;
      const bool has_var = !(flags & kLmConstant) ||
                           (pose_coupled && (obs_observer_var(meta) || t.vertex_var[t.lm_anchor[l]] >= 0));
      if (has_var) cost = 0.5 * rho[0]; else fixed = 0.5 * rho[0];
      if (kJac) {
        // ceres Corrector: J~ = sqrt(rho') (I - alpha r r'/|r|^2) J, r~ = rs r.
        const double sq1 = sqrt(rho[1]);
        double rs = sq1, asq = 0.0;
        if (s > 0.0 && rho[2] > 0.0) {
          const double alpha = 1.0 - sqrt(1.0 + 2.0 * s * rho[2] / rho[1]);
          rs = sq1 / (1.0 - alpha);
          asq = alpha / s;
        }
        double rr[2] = {ev.r[0] * rs, ev.r[1] * rs};
        auto correct = [&](double* J, int cols) {
          for (int c = 0; c < cols; ++c) {
            const double rtj = J[c] * ev.r[0] + J[cols + c] * ev.r[1];
            J[c] = sq1 * (J[c] - asq * ev.r[0] * rtj);
            J[cols + c] = sq1 * (J[cols + c] - asq * ev.r[1] * rtj);
          }
        };
        correct(ev.Jl, 3);
        if (pose_coupled) correct(ev.Jo, 6);
        if (kF32) {
          for (int i = 0; i < 6; ++i) ev.Jl[i] = (double)(float)ev.Jl[i];
          for (int i = 0; i < 12; ++i) ev.Jo[i] = (double)(float)ev.Jo[i];
          Rec32* rec = static_cast<Rec32*>(obs_rec) + o;
          rec->r[0] = rr[0]; rec->r[1] = rr[1];
          for (int i = 0; i < 6; ++i) rec->J[i] = (float)ev.Jl[i];
          for (int i = 0; i < 12; ++i) rec->J[6 + i] = (float)ev.Jo[i];
        } else {
          double* rec = static_cast<double*>(obs_rec) + (size_t)kObsRecord * o;
          rec[0] = rr[0]; rec[1] = rr[1];
          for (int i = 0; i < 6; ++i) rec[2 + i] = ev.Jl[i];
          for (int i = 0; i < 12; ++i) rec[8 + i] = ev.Jo[i];
        }
        const double* J = ev.Jl;
        double V[6], g[3];
        V[0] = J[0] * J[0] + J[3] * J[3];
        V[1] = J[0] * J[1] + J[3] * J[4];
        V[2] = J[0] * J[2] + J[3] * J[5];
        V[3] = J[1] * J[1] + J[4] * J[4];
        V[4] = J[1] * J[2] + J[4] * J[5];
        V[5] = J[2] * J[2] + J[5] * J[5];
        for (int c = 0; c < 3; ++c) g[c] = J[c] * rr[0] + J[3 + c] * rr[1];
        for (int i = 0; i < 6; ++i) acc[i] = V[i];
        for (int i = 0; i < 3; ++i) acc[12 + i] = g[i];
        if (pose_coupled) {
          for (int i = 0; i < 6; ++i) acc[6 + i] = V[i];
          for (int i = 0; i < 3; ++i) acc[15 + i] = g[i];
        }
      }
    } else if (kJac) {
      if (kF32) {
        Rec32* rec = static_cast<Rec32*>(obs_rec) + o;
        rec->r[0] = rec->r[1] = 0.0;
        for (int i = 0; i < 18; ++i) rec->J[i] = 0.0f;
      } else {
        double* rec = static_cast<double*>(obs_rec) + (size_t)kObsRecord * o;
        for (int i = 0; i < kObsRecord; ++i) rec[i] = 0.0;
      }
    }
  }
  if (kJac) {
    for (int i = 0; i < kLmAcc; ++i) red[tid * kLmAcc + i] = acc[i];
    seg[tid] = (o < e) ? l : -1;
    __syncthreads();
    // Segment heads reduce their landmark's run in a fixed order.
    if (o < e && (tid == 0 || seg[tid - 1] != l)) {
      double sum[kLmAcc];
      for (int i = 0; i < kLmAcc; ++i) sum[i] = 0.0;
      for (int k = tid; k < kObsTile && seg[k] == l; ++k)
        for (int i = 0; i < kLmAcc; ++i) sum[i] += red[k * kLmAcc + i];
      // A split landmark owns its tiles whole: the tile's sum goes to the
      // tile's slot, and k_split_sum adds the slots in tile order.
      double* dst = (t.lm_flags[l] & kLmSplit) ? split_part + (size_t)kLmAcc * blockIdx.x
                                               : lm_acc + (size_t)kLmAcc * l;
      for (int i = 0; i < kLmAcc; ++i) dst[i] = sum[i];
    }
  }
  const double c_sum = block_sum256(cost, lds4);
  const double f_sum = block_sum256(fixed, lds4);
  if (tid == 0) {
    tile_cost[2 * blockIdx.x] = c_sum;
    tile_cost[2 * blockIdx.x + 1] = f_sum;
  }
}

// Accumulators of the landmarks split over several tiles: the tile sums of
// each, added in tile order. split_tiles[2k], [2k+1] = first tile, tile count.
__global__ void k_split_sum(int32_t n, const int32_t* __restrict__ ids, const int32_t* __restrict__ split_tiles,
                            const double* __restrict__ split_part, double* __restrict__ lm_acc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * kLmAcc) return;
  const int k = i / kLmAcc, c = i % kLmAcc;
  const int t0 = split_tiles[2 * k], nt = split_tiles[2 * k + 1];
  double s = 0.0;
  for (int j = 0; j < nt; ++j) s += split_part[(size_t)kLmAcc * (t0 + j) + c];
  lm_acc[(size_t)kLmAcc * ids[k] + c] = s;
}

// ------------------------------------------------------------ vertex rows

// Per variable vertex: unscaled gradient g = J'r and diag(J'J) of its six
// tangent columns, over its incidences (observations it makes and the
// landmarks anchored in it).
template <bool kF32>
__global__ __launch_bounds__(kRowThreads) void k_pose_rows(DevTopology t, const double* __restrict__ vert,
                                                            const double* __restrict__ vR,
                                                            const double* __restrict__ lm,
                                                            const void* __restrict__ obs_rec,
                                                            const double* __restrict__ lm_acc,
                                                            double* __restrict__ pose_grad,
                                                            double* __restrict__ pose_diag,
                                                            double* __restrict__ pose_jtj) {
  constexpr int kVals = 12 + 21;  // gradient, diagonal, upper triangle of sum Jo^T Jo
  __shared__ double red[kRowThreads / 64][kVals];
  const int i = blockIdx.x;
  const int ib = t.inc_ptr[i], ie = t.inc_ptr[i + 1];
  double g[6] = {0, 0, 0, 0, 0, 0}, d[6] = {0, 0, 0, 0, 0, 0}, jj[21];
  for (int k = 0; k < 21; ++k) jj[k] = 0.0;
  for (int k = ib + threadIdx.x; k < ie; k += kRowThreads) {
    const int32_t x = t.inc[k];
    if (x >= 0) {
      const RecView<kF32> rec(obs_rec, x);
      double Jo[12];
      for (int k2 = 0; k2 < 12; ++k2) Jo[k2] = rec[8 + k2];
      for (int c = 0; c < 6; ++c) {
        g[c] += Jo[c] * rec[0] + Jo[6 + c] * rec[1];
        d[c] += Jo[c] * Jo[c] + Jo[6 + c] * Jo[6 + c];
      }
      int u = 0;
      for (int r = 0; r < 6; ++r)
        for (int c = r; c < 6; ++c, ++u) jj[u] += Jo[r] * Jo[c] + Jo[6 + r] * Jo[6 + c];
    } else {
      const int32_t l = -1 - x;
      const int32_t a = t.lm_anchor[l];
      double M[18], VLM[9];
      anchor_map(lm + 3 * l, vR + 12 * a, vR[12 * a + 9], M);
      const double* acc = lm_acc + (size_t)kLmAcc * l;
      sym3_full(acc + 6, VLM);
      const double* gLM = acc + 15;
      for (int c = 0; c < 6; ++c) {
        const double m0 = M[c], m1 = M[6 + c], m2 = M[12 + c];
        g[c] += m0 * gLM[0] + m1 * gLM[1] + m2 * gLM[2];
        d[c] += m0 * (VLM[0] * m0 + VLM[1] * m1 + VLM[2] * m2) + m1 * (VLM[3] * m0 + VLM[4] * m1 + VLM[5] * m2) +
                m2 * (VLM[6] * m0 + VLM[7] * m1 + VLM[8] * m2);
      }
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int c = 0; c < 6; ++c) {
    g[c] = wave_sum(g[c]);
    d[c] = wave_sum(d[c]);
  }
  for (int k = 0; k < 21; ++k) jj[k] = wave_sum(jj[k]);
  if (lane == 0) {
    for (int c = 0; c < 6; ++c) { red[wave][c] = g[c]; red[wave][6 + c] = d[c]; }
    for (int k = 0; k < 21; ++k) red[wave][12 + k] = jj[k];
  }
  __syncthreads();
  if (threadIdx.x < 12) {
    double s = 0.0;
    for (int w = 0; w < kRowThreads / 64; ++w) s += red[w][threadIdx.x];
    if (threadIdx.x < 6) pose_grad[6 * i + threadIdx.x] = s;
    else pose_diag[6 * i + threadIdx.x - 6] = s;
  } else if (threadIdx.x >= 64 && threadIdx.x < 100) {
    const int e = threadIdx.x - 64, r = min(e / 6, e % 6), c = max(e / 6, e % 6);
    const int u = r * 6 - r * (r - 1) / 2 + (c - r);  // index of (r, c), r <= c, in the upper triangle
    double s = 0.0;
    for (int w = 0; w < kRowThreads / 64; ++w) s += red[w][12 + u];
    pose_jtj[36 * i + e] = s;
  }
}

// ceres Jacobi scaling: c = 1 / (1 + sqrt(diag(J'J))) of every tangent column.
__global__ void k_jacobi_scaling(int32_t Pvar, int32_t L, int32_t use, const double* __restrict__ pose_diag,
                                 const double* __restrict__ lm_acc, double* __restrict__ pose_scale,
                                 double* __restrict__ lm_scale) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 6 * Pvar) pose_scale[i] = use ? 1.0 / (1.0 + sqrt(pose_diag[i])) : 1.0;
  if (i < 3 * L) {
    const int l = i / 3, c = i % 3;
    const int di = c == 0 ? 0 : (c == 1 ? 3 : 5);
    lm_scale[i] = use ? 1.0 / (1.0 + sqrt(lm_acc[(size_t)kLmAcc * l + di])) : 1.0;
  }
}

// ------------------------------------------------------------ landmark pivot

// Z_l = C (C V C + D^2)^-1 C, D^2 = clamp(diag(C V C)) / mu (ceres
// LevenbergMarquardtStrategy diagonal), followed by a copy of the landmark
// gradient g_l, so that the per-incidence Schur factors read Z_l and g_l of
// their (scattered) landmark from one 72-byte record. Zero for non-variable
// landmarks.
__global__ void k_landmark_prep(DevTopology t, const double* __restrict__ lm_acc,
                                const double* __restrict__ lm_scale, double inv_mu, double min_diag,
                                double max_diag, double* __restrict__ lm_z, int32_t* __restrict__ fail) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= t.L) return;
  double* z = lm_z + kLmZ * l;
  const uint32_t flags = t.lm_flags[l];
  if (!(flags & kLmActive) || (flags & kLmConstant)) {
    for (int i = 0; i < kLmZ; ++i) z[i] = 0.0;
    return;
  }
  const double* V = lm_acc + (size_t)kLmAcc * l;
  const double* c = lm_scale + 3 * l;
  double A[6] = {c[0] * c[0] * V[0], c[0] * c[1] * V[1], c[0] * c[2] * V[2],
                 c[1] * c[1] * V[3], c[1] * c[2] * V[4], c[2] * c[2] * V[5]};
  A[0] += fmin(fmax(A[0], min_diag), max_diag) * inv_mu;
  A[3] += fmin(fmax(A[3], min_diag), max_diag) * inv_mu;
  A[5] += fmin(fmax(A[5], min_diag), max_diag) * inv_mu;
  double inv[6];
  if (!spd3_inverse(A, inv)) {
    atomicAdd(fail, 1);
    for (int i = 0; i < kLmZ; ++i) z[i] = 0.0;
    return;
  }
  z[0] = c[0] * c[0] * inv[0]; z[1] = c[0] * c[1] * inv[1]; z[2] = c[0] * c[2] * inv[2];
  z[3] = c[1] * c[1] * inv[3]; z[4] = c[1] * c[2] * inv[4]; z[5] = c[2] * c[2] * inv[5];
  z[6] = V[12]; z[7] = V[13]; z[8] = V[14];
}

// ------------------------------------------------------------ Schur complement

// W' = Jl^T (Jo C) (3x6): landmark side unscaled, pose side scaled.
template <class Rec>
__device__ inline void wprime_obs(const Rec& rec, const double* cs, double* W) {
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 6; ++c) W[a * 6 + c] = (rec[2 + a] * rec[8 + c] + rec[5 + a] * rec[14 + c]) * cs[c];
}

// Factor blocks of the reduced camera system
//   S = F'F + D^2 - sum_l W_l' Z_l W_l,  rhs = F'r - sum_l W_l' Z_l g_l,
// with Z_l = C_l (C_l V_l C_l + D_l^2)^-1 C_l from k_landmark_prep. An
// incidence x of variable vertex i is an observation by i whose residual
// depends on the pose of i, or the anchor of a landmark l whose anchor
// vertex i is variable (A_l = M_l C_i, the anchor map of the landmark):
//   observation: W_x = Jl^T Jo C_i,  r0_x = C_i Jo^T r
//   anchor:      W_x = V_LM,l A_l,   r0_x = A_l^T g_LM,l
// and per LM iteration (Z_l depends on the damping)
//   Q_x = [anchor: A_l^T] - W_x^T Z_l,  rhs_x = r0_x - W_x^T Z_l g_l
// (ceres' Schur eliminator, schur_eliminator_impl.h, restated on the
// anchor-lifted blocks). S_ij sums, over the landmarks l seen from rows i and
// j >= i, Q_x W_y + [x observation, y anchor] W_x^T A_l over l's incidences
// x in row i and y in row j. Both terms are linear in the incidences of one
// row, so the incidences of a landmark in one row are summed first into a
// factor block (W, r0); an anchor is alone in its row and its block stores
// A_l in place of W (the Schur kernel forms W = V_LM,l A_l next to it).
// Factor blocks are stored landmark-major (rows ascending), per
// linearization.
template <bool kF32>
__global__ __launch_bounds__(256) void k_schur_factors(
    DevTopology t, int32_t n_g, const int32_t* __restrict__ g_row, const int32_t* __restrict__ g_inc_ptr,
    const int32_t* __restrict__ g_inc, const double* __restrict__ vR, const double* __restrict__ lm,
    const void* __restrict__ obs_rec, const double* __restrict__ lm_acc, const double* __restrict__ pose_scale,
    double* __restrict__ gw, double* __restrict__ gr0) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_g) return;
  const int i = g_row[g];
  const double* cs = pose_scale + 6 * i;
  double W[18], r0[6];
  for (int k = 0; k < 18; ++k) W[k] = 0.0;
  for (int c = 0; c < 6; ++c) r0[c] = 0.0;
  for (int m = g_inc_ptr[g]; m < g_inc_ptr[g + 1]; ++m) {
    const int32_t x = g_inc[m];
    double Wx[18], rx[6];
    if (x >= 0) {
      const RecView<kF32> rec(obs_rec, x);
      wprime_obs(rec, cs, Wx);
      for (int c = 0; c < 6; ++c) rx[c] = cs[c] * (rec[8 + c] * rec[0] + rec[14 + c] * rec[1]);
    } else {
      const int32_t l = -1 - x;
      const int v = t.var_vertex[i];
      const double* acc_l = lm_acc + (size_t)kLmAcc * l;
      double M[18], A[18];
      anchor_map(lm + 3 * l, vR + 12 * v, vR[12 * v + 9], M);
      for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 6; ++c) A[a * 6 + c] = M[a * 6 + c] * cs[c];
      const double* gLM = acc_l + 15;
      for (int c = 0; c < 6; ++c) rx[c] = A[c] * gLM[0] + A[6 + c] * gLM[1] + A[12 + c] * gLM[2];
      // The anchor's block keeps A_l; the Schur kernel forms W_x = V_LM,l A_l.
      for (int k = 0; k < 18; ++k) Wx[k] = A[k];
    }
    for (int k = 0; k < 18; ++k) W[k] += Wx[k];
    for (int c = 0; c < 6; ++c) r0[c] += rx[c];
  }
  double* w = gw + (size_t)18 * g;
  for (int k = 0; k < 18; ++k) w[k] = W[k];
  double* rr = gr0 + (size_t)6 * g;
  for (int c = 0; c < 6; ++c) rr[c] = r0[c];
}

// LDS of one batch of a tile: the staged factor blocks (W and Q^T, both 3x6
// row-major, and the rhs contribution r0 - W^T Z g on diagonal tiles), the
// batch's pair list and the segment's batch words. Slot entries are
// slot-minor (entry e of slot s at [e][s]), so the staging writes of
// consecutive threads hit consecutive banks.
struct TileLds {
  double f[2][18][kBatchSlots];  // [0] W, [1] Q^T = [anchor: A_l] - Z_l W
  double r[6][kBatchSlots];
  uint32_t pairs[kBatchPairs / 2];
  int4 bt[kSegmentBatches];
  uint8_t anchor[kBatchSlots];
};

struct TileArgs {
  const int4* batches;
  const int2* slots;
  const uint16_t* batch_off;
  const uint32_t* pairs;  // the pair list as 32-bit words (batches start on a word)
  const double* gw;
  const double* gr0;
  const double* lm_acc;
  const double* lm_z;
};

// A thread's index loads of one batch (two batches ahead): its slot's word,
// its part's pair range and its share of the pair list.
struct TileIdx {
  int2 slot;
  int32_t pb, pe;
  uint32_t pw[kPairWords];
};

// A thread's data loads of one batch (one batch ahead): its slot's factor
// block (W, or A_l for an anchor's block) and r0, its landmark's pivot (Z,
// g) and, for an anchor, V_LM; and the index words.
struct TileStage {
  double w[18], z[9], v[6], r0[6];
  int32_t pb, pe;
  uint32_t pw[kPairWords];
  int2 slot;
};

__device__ inline void tile_idx(TileIdx& q, const TileArgs& ta, const int4& bt, int b, int tid, int blk, int part,
                                int np) {
  q.slot = make_int2(-1, 0);
  if (tid < bt.y) q.slot = ta.slots[bt.x + tid];
  q.pb = q.pe = 0;
  if (np) {
    const uint16_t* off = ta.batch_off + (size_t)b * (kTileBlocks + 1);
    q.pb = off[blk] + part;
    q.pe = off[blk + 1];
  }
  const int nw = (bt.w + 1) >> 1;
  const uint32_t* pw = ta.pairs + (bt.z >> 1);
#pragma unroll
  for (int k = 0; k < kPairWords; ++k) {
    const int j = tid + k * kTileBlocks;
    q.pw[k] = j < nw ? pw[j] : 0u;
  }
}

__device__ inline void tile_data(TileStage& st, const TileIdx& q, const TileArgs& ta, bool diag) {
  st.slot = q.slot;
  st.pb = q.pb;
  st.pe = q.pe;
#pragma unroll
  for (int k = 0; k < kPairWords; ++k) st.pw[k] = q.pw[k];
  if (q.slot.x < 0) return;
  const uint32_t lw = (uint32_t)q.slot.y;
  const int l = (int)(lw & ~kSlotAnchor);
  const double* wp = ta.gw + (size_t)18 * q.slot.x;
#pragma unroll
  for (int e = 0; e < 18; ++e) st.w[e] = wp[e];
  const double* zp = ta.lm_z + (size_t)kLmZ * l;
#pragma unroll
  for (int e = 0; e < 9; ++e) st.z[e] = zp[e];
  if (lw & kSlotAnchor) {
    const double* vp = ta.lm_acc + (size_t)kLmAcc * l + 6;
#pragma unroll
    for (int e = 0; e < 6; ++e) st.v[e] = vp[e];
  }
  if (diag) {
    const double* rp = ta.gr0 + (size_t)6 * q.slot.x;
#pragma unroll
    for (int e = 0; e < 6; ++e) st.r0[e] = rp[e];
  }
}

// One work item = one segment of a tile (I, J). Per batch: every thread
// stages its slot's factor block in LDS (Q^T formed from W, Z and A) and its
// share of the pair list, then each thread adds the products of its part of
// its block's pairs in list order,
//   Q_x W_y              (y not the anchor, or x the anchor too)
//   W_x^T (Q^T_y)        (y the anchor of an observation x: Q_x W_y + W_x^T A_l)
// both as sum_a Lt(a, r) Rt(a, c) over the rows of two 3x6 blocks; a part of
// a diagonal block of a diagonal tile also adds the rhs of its pairs' block.
// Loads are software-pipelined: while a batch's products run, the next
// batch's data loads and the batch after that's index loads are in flight.
// Each part's sums go to its own slot of the segment's partial. Work-groups
// take items_per_wg consecutive items of the dispatch order; with xcd_map
// each XCD walks a contiguous range of work-groups (work-group b runs on XCD
// b mod 8). Neither changes a partial.
__global__ __launch_bounds__(kTileBlocks) __attribute__((amdgpu_waves_per_eu(2))) void k_schur_tiles(
    int32_t n_segs, int32_t items_per_wg, int32_t xcd_map, const int32_t* __restrict__ seg_order,
    const int4* __restrict__ segs, const uint32_t* __restrict__ seg_thread, TileArgs ta,
    double* __restrict__ partial) {
  __shared__ TileLds sh;
  const int wgs = (n_segs + items_per_wg - 1) / items_per_wg;
  const int per_xcd = gridDim.x / 8;
  const int wg = xcd_map ? (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8) : (int)blockIdx.x;
  if (wg >= wgs) return;
  const int tid = threadIdx.x;
  for (int it = 0; it < items_per_wg; ++it) {
    const int si = wg * items_per_wg + it;
    if (si >= n_segs) break;
    const int sid = seg_order[si];
    const int4 sg = segs[sid];
    const bool diag = sg.z == sg.w;
    const int nb = sg.y - sg.x;
    const uint32_t tw = seg_thread[(size_t)sid * kTileBlocks + tid];
    const int np = seg_thread_parts(tw), blk = seg_thread_block(tw), part = seg_thread_part(tw);
    const bool rhs_part = diag && np && blk / kTileBand == blk % kTileBand;
    __syncthreads();
    if (tid < nb) sh.bt[tid] = ta.batches[sg.x + tid];
    __syncthreads();
    double acc[36], racc[6];
#pragma unroll
    for (int e = 0; e < 36; ++e) acc[e] = 0.0;
#pragma unroll
    for (int c = 0; c < 6; ++c) racc[c] = 0.0;
    TileIdx q;
    TileStage st;
    tile_idx(q, ta, sh.bt[0], sg.x, tid, blk, part, np);
    tile_data(st, q, ta, diag);
    if (nb > 1) tile_idx(q, ta, sh.bt[1], sg.x + 1, tid, blk, part, np);
    for (int k = 0; k < nb; ++k) {
      __syncthreads();
      if (st.slot.x >= 0) {
        // Column by column: W (= V_LM A_l for an anchor, whose block holds
        // A_l), Q^T = [anchor: A_l] - Z W and r0 - W^T Z g.
        const double* z = st.z;
        const double Z[9] = {z[0], z[1], z[2], z[1], z[3], z[4], z[2], z[4], z[5]};
        const bool anc = ((uint32_t)st.slot.y & kSlotAnchor) != 0;
        const double* v = st.v;
        const double V[9] = {v[0], v[1], v[2], v[1], v[3], v[4], v[2], v[4], v[5]};
        double zg[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) zg[a] = Z[a * 3] * z[6] + Z[a * 3 + 1] * z[7] + Z[a * 3 + 2] * z[8];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
          const double b0 = st.w[c], b1 = st.w[6 + c], b2 = st.w[12 + c];
          double w0 = b0, w1 = b1, w2 = b2;
          if (anc) {
            w0 = V[0] * b0 + V[1] * b1 + V[2] * b2;
            w1 = V[3] * b0 + V[4] * b1 + V[5] * b2;
            w2 = V[6] * b0 + V[7] * b1 + V[8] * b2;
          }
          const double q0 = Z[0] * w0 + Z[1] * w1 + Z[2] * w2;
          const double q1 = Z[3] * w0 + Z[4] * w1 + Z[5] * w2;
          const double q2 = Z[6] * w0 + Z[7] * w1 + Z[8] * w2;
          sh.f[0][c][tid] = w0;
          sh.f[0][6 + c][tid] = w1;
          sh.f[0][12 + c][tid] = w2;
          sh.f[1][c][tid] = anc ? b0 - q0 : -q0;
          sh.f[1][6 + c][tid] = anc ? b1 - q1 : -q1;
          sh.f[1][12 + c][tid] = anc ? b2 - q2 : -q2;
          if (diag) sh.r[c][tid] = st.r0[c] - (w0 * zg[0] + w1 * zg[1] + w2 * zg[2]);
        }
        sh.anchor[tid] = anc ? 1 : 0;
      }
#pragma unroll
      for (int j = 0; j < kPairWords; ++j) sh.pairs[tid + j * kTileBlocks] = st.pw[j];
      const int pb = st.pb, pe = st.pe;
      __syncthreads();
      if (k + 1 < nb) tile_data(st, q, ta, diag);
      if (k + 2 < nb) tile_idx(q, ta, sh.bt[k + 2], sg.x + k + 2, tid, blk, part, np);
      const uint16_t* pl = reinterpret_cast<const uint16_t*>(sh.pairs);
      for (int p = pb; p < pe; p += np) {
        const uint32_t pr = pl[p];
        const int x = (int)(pr & 255u), y = (int)(pr >> 8);
        const int sel = (sh.anchor[y] && !sh.anchor[x]) ? 1 : 0;
        const double* Lt = &sh.f[1 - sel][0][x];
        const double* Rt = &sh.f[sel][0][y];
#pragma unroll 1
        for (int a = 0; a < 3; ++a) {
          double l6[6], r6[6];
#pragma unroll
          for (int c = 0; c < 6; ++c) {
            l6[c] = Lt[(a * 6 + c) * kBatchSlots];
            r6[c] = Rt[(a * 6 + c) * kBatchSlots];
          }
#pragma unroll
          for (int r = 0; r < 6; ++r)
#pragma unroll
            for (int c = 0; c < 6; ++c) acc[r * 6 + c] += l6[r] * r6[c];
        }
        if (rhs_part)
#pragma unroll
          for (int c = 0; c < 6; ++c) racc[c] += sh.r[c][x];
      }
    }
    if (np) {
      double* out = partial + (size_t)sid * kTilePartial + (size_t)tid * kPartEntries;
#pragma unroll
      for (int e = 0; e < 36; ++e) out[e] = acc[e];
#pragma unroll
      for (int c = 0; c < 6; ++c) out[36 + c] = racc[c];
    }
  }
}

// S_ij = sum of the block's part slots (segments in segment order, parts in
// part order, listed per upper block), plus on S_ii the direct observation
// term C_i (sum Jo^T Jo) C_i and the LM diagonal D^2 / radius; written to
// S_ij and S_ji = S_ij^T. One thread per entry; the first six threads of a
// diagonal block also sum the rhs of its row.
__global__ __launch_bounds__(252) void k_schur_finalize(int32_t nub, const int4* __restrict__ ublk,
                                                        const int32_t* __restrict__ uslot_ptr,
                                                        const int32_t* __restrict__ uslot,
                                                        const double* __restrict__ partial,
                                                        const double* __restrict__ pose_scale,
                                                        const double* __restrict__ pose_diag,
                                                        const double* __restrict__ pose_jtj, double inv_mu,
                                                        double min_diag, double max_diag, double* __restrict__ S,
                                                        double* __restrict__ rhs) {
  const int u = blockIdx.x * 7 + threadIdx.x / 36, e = threadIdx.x % 36;
  if (u >= nub) return;
  const int4 ub = ublk[u];
  const int kb = uslot_ptr[u], ke = uslot_ptr[u + 1];
  const bool rhs_entry = ub.z == ub.w && e < 6;
  double v = 0.0, g = 0.0;
  int k = kb;
  // Four slots' loads in flight at a time; the sums keep slot order.
  for (; k + 4 <= ke; k += 4) {
    const double* p0 = partial + (size_t)uslot[k] * kPartEntries;
    const double* p1 = partial + (size_t)uslot[k + 1] * kPartEntries;
    const double* p2 = partial + (size_t)uslot[k + 2] * kPartEntries;
    const double* p3 = partial + (size_t)uslot[k + 3] * kPartEntries;
    const double a0 = p0[e], a1 = p1[e], a2 = p2[e], a3 = p3[e];
    v += a0;
    v += a1;
    v += a2;
    v += a3;
    if (rhs_entry) {
      const double b0 = p0[36 + e], b1 = p1[36 + e], b2 = p2[36 + e], b3 = p3[36 + e];
      g += b0;
      g += b1;
      g += b2;
      g += b3;
    }
  }
  for (; k < ke; ++k) {
    const double* p = partial + (size_t)uslot[k] * kPartEntries;
    v += p[e];
    if (rhs_entry) g += p[36 + e];
  }
  const int r = e / 6, c = e % 6;
  if (ub.z == ub.w) {
    const int i = ub.z;
    v += pose_scale[6 * i + r] * pose_scale[6 * i + c] * pose_jtj[36 * i + e];
    if (r == c) {
      const double sc = pose_scale[6 * i + c];
      v += fmin(fmax(sc * sc * pose_diag[6 * i + c], min_diag), max_diag) * inv_mu;
    }
    if (rhs_entry) rhs[6 * i + e] = g;
  }
  S[(size_t)36 * ub.x + e] = v;
  if (ub.y != ub.x) S[(size_t)36 * ub.y + c * 6 + r] = v;
}

// ------------------------------------------------------------ landmark admission

struct QualityArgs {
  int32_t min_observers, pad;
  double min_angle_deg;  // smallest accepted max disparity angle of the incidence rays
  double max_dist, min_dist;
  double max_reproj2;    // squared; <= 0: check off
};

constexpr int kQualityRays = 256;  // incidence rays kept in LDS per landmark

// Camera centre and landmark of observation o: ray = c_G - p_G (unit, with
// its length) and the landmark's depth in the camera. The frame chain is
// T_G_C = T_G_M T_M_I T_C_I^-1 and p_C = T_C_I T_M_I^-1 T_G_M^-1 p_G
// (landmark-quality-metrics.cc:118-127).
__device__ inline void quality_ray(const DevTopology& t, const double* __restrict__ vert,
                                   const double* __restrict__ vR, int o, const double* pG, double* ray,
                                   double* dist, double* pC) {
  const int v = t.obs_vertex[o];
  const DevCamera& cam = t.cams[obs_camera(t.obs_meta[o])];
  const double* Rv = vR + 12 * v;           // R_I_M
  const double* pv = vert + 7 * v + 4;      // p_M_I
  const double* TM = t.mission_T + 12 * t.vertex_mission[v];
  double cI[3], cM[3], cG[3];
  mat3t_vec(cam.R_C_I, cam.p_C_I, cI);
  for (int a = 0; a < 3; ++a) cI[a] = -cI[a];
  mat3t_vec(Rv, cI, cM);
  for (int a = 0; a < 3; ++a) cM[a] += pv[a];
  mat3_vec(TM, cM, cG);
  for (int a = 0; a < 3; ++a) cG[a] += TM[9 + a];
  double d[3], pM[3], pI[3];
  for (int a = 0; a < 3; ++a) d[a] = pG[a] - TM[9 + a];
  mat3t_vec(TM, d, pM);
  for (int a = 0; a < 3; ++a) pM[a] -= pv[a];
  mat3_vec(Rv, pM, pI);
  mat3_vec(cam.R_C_I, pI, pC);
  for (int a = 0; a < 3; ++a) pC[a] += cam.p_C_I[a];
  for (int a = 0; a < 3; ++a) ray[a] = cG[a] - pG[a];
  *dist = sqrt(ray[0] * ray[0] + ray[1] * ray[1] + ray[2] * ray[2]);
}

// isVisualLandmarkWellConstrained of every landmark, one 64-thread
// work-group per landmark: lanes take the observations (signed distance to
// the closest observer, optional reprojection check, unit rays kept in LDS)
// and then the pairs of rays (smallest |cos| of the bundle).
__global__ __launch_bounds__(64) void k_landmark_quality(DevTopology t, const double* __restrict__ vert,
                                                         const double* __restrict__ vR,
                                                         const double* __restrict__ lm, QualityArgs q,
                                                         uint8_t* __restrict__ good) {
  __shared__ double rays[kQualityRays][3];
  __shared__ int valid[kQualityRays];
  const int l = blockIdx.x, lane = threadIdx.x;
  const int ob = t.lm_begin[l], n = t.lm_begin[l + 1] - ob;
  if (n < q.min_observers) {
    if (lane == 0) good[l] = 0;
    return;
  }
  // Landmark in G from its store vertex and mission.
  const int a = t.lm_anchor[l];
  const double* TA = t.mission_T + 12 * t.vertex_mission[a];
  double pM[3], pG[3];
  mat3t_vec(vR + 12 * a, lm + 3 * l, pM);
  for (int k = 0; k < 3; ++k) pM[k] += vert[7 * a + 4 + k];
  mat3_vec(TA, pM, pG);
  for (int k = 0; k < 3; ++k) pG[k] += TA[9 + k];
  double smin = 1.7976931348623157e308;  // DBL_MAX
  int bad = 0;
  for (int oi = lane; oi < n; oi += 64) {
    const int o = ob + oi;
    double ray[3], dist, pC[3];
    quality_ray(t, vert, vR, o, pG, ray, &dist, pC);
    smin = fmin(smin, pC[2] < 0.0 ? -dist : dist);
    if (q.max_reproj2 > 0.0) {
      const DevCamera& cam = t.cams[obs_camera(t.obs_meta[o])];
      double kp[2], jp[6];
      const int st = project_camera(cam.model, cam.xi, cam.dist_type, cam.intr, cam.dist, cam.width, cam.height, pC, kp, jp);
      double e2 = 1.7976931348623157e308;
      if (st == kVisible || st == kOutsideImage) {
        const double du = kp[0] - t.obs_kp[2 * o], dv = kp[1] - t.obs_kp[2 * o + 1];
        e2 = du * du + dv * dv;
      }
      if (e2 > q.max_reproj2) bad = 1;
    }
    if (oi < kQualityRays) {
      const double inv = dist > 0.0 ? 1.0 / dist : 0.0;
      for (int k = 0; k < 3; ++k) rays[oi][k] = ray[k] * inv;
      valid[oi] = dist > 0.0;
    }
  }
  smin = wave_min(smin);
  bad = __any(bad);
  __syncthreads();
  // Smallest |cos| over the pairs i < j of valid rays (pairs in LDS, or
  // recomputed beyond kQualityRays).
  double mincos = 1.0;
  if (!bad && !(smin > q.max_dist || smin < q.min_dist))
    for (int i = 0; i + 1 < n; ++i) {
      double ri[3];
      bool vi;
      if (i < kQualityRays) {
        for (int k = 0; k < 3; ++k) ri[k] = rays[i][k];
        vi = valid[i];
      } else {
        double d, pC[3];
        quality_ray(t, vert, vR, ob + i, pG, ri, &d, pC);
        vi = d > 0.0;
        for (int k = 0; k < 3; ++k) ri[k] = vi ? ri[k] / d : 0.0;
      }
      if (!vi) continue;
      for (int j = i + 1 + lane; j < n; j += 64) {
        double rj[3];
        bool vj;
        if (j < kQualityRays) {
          for (int k = 0; k < 3; ++k) rj[k] = rays[j][k];
          vj = valid[j];
        } else {
          double d, pC[3];
          quality_ray(t, vert, vR, ob + j, pG, rj, &d, pC);
          vj = d > 0.0;
          for (int k = 0; k < 3; ++k) rj[k] = vj ? rj[k] / d : 0.0;
        }
        if (vj) mincos = fmin(mincos, fabs(ri[0] * rj[0] + ri[1] * rj[1] + ri[2] * rj[2]));
      }
    }
  mincos = -wave_max(-mincos);
  if (lane == 0) {
    // Fewer than two valid rays gives angle 0.
    const double angle_deg = acos(fmin(mincos, 1.0)) * (180.0 / 3.14159265358979323846);
    good[l] = (!bad && !(smin > q.max_dist || smin < q.min_dist) && !(angle_deg < q.min_angle_deg)) ? 1 : 0;
  }
}

// ------------------------------------------------------------ PCG

// 6x6 SPD inverse of every diagonal block of S (block-Jacobi preconditioner).
__global__ void k_block_inverse(int32_t Pvar, const int32_t* __restrict__ diag_pos, const double* __restrict__ S,
                                double* __restrict__ Minv, int32_t* __restrict__ fail) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Pvar) return;
  const double* A = S + (size_t)36 * diag_pos[i];
  double L[36];
  for (int k = 0; k < 36; ++k) L[k] = 0.0;
  bool ok = true;
  for (int j = 0; j < 6; ++j) {
    double s = A[j * 6 + j];
    for (int k = 0; k < j; ++k) s -= L[j * 6 + k] * L[j * 6 + k];
    if (!(s > 0.0)) { ok = false; s = 1.0; }
    const double ljj = sqrt(s);
    L[j * 6 + j] = ljj;
    for (int r = j + 1; r < 6; ++r) {
      double v = A[r * 6 + j];
      for (int k = 0; k < j; ++k) v -= L[r * 6 + k] * L[j * 6 + k];
      L[r * 6 + j] = v / ljj;
    }
  }
  if (!ok) atomicAdd(fail, 1);
  // Columns of A^-1 by forward / backward substitution.
  double* out = Minv + (size_t)36 * i;
  for (int c = 0; c < 6; ++c) {
    double y[6];
    for (int r = 0; r < 6; ++r) {
      double s = (r == c) ? 1.0 : 0.0;
      for (int k = 0; k < r; ++k) s -= L[r * 6 + k] * y[k];
      y[r] = s / L[r * 6 + r];
    }
    for (int r = 5; r >= 0; --r) {
      double s = y[r];
      for (int k = r + 1; k < 6; ++k) s -= L[k * 6 + r] * y[k];
      y[r] = s / L[r * 6 + r];
    }
    for (int r = 0; r < 6; ++r) out[r * 6 + c] = y[r];
  }
}

// ------------------------------------------------------------ step

// Landmark step and model cost change. With y the solution of the scaled
// system, w = C_p y_p (per vertex) and u = C_l y_l = Z_l (g_l - sum W' y_p):
// J y = Jl u + Jo w_o + Jl M w_a per observation and
// model_cost_change = sum (J y)'r - |J y|^2 / 2 (ceres ComputeTrustRegionStep).
// One work-group per linearization tile of whole landmarks: each observation
// record is read once; the landmark sums are taken by the segment heads in a
// fixed order. Tiles of landmarks split over several tiles are left to
// k_back_substitute_split (their part entries are zero here). u of inactive
// or observation-less landmarks is not written (kept zero).
__device__ inline void landmark_step(const DevTopology& t, int l, const double* __restrict__ vR,
                                     const double* __restrict__ lm, const double* __restrict__ lm_acc,
                                     const double* __restrict__ lm_z, const double* __restrict__ w,
                                     double* tv, double* u, double* Mwa) {
  const uint32_t flags = t.lm_flags[l];
  const double* acc = lm_acc + (size_t)kLmAcc * l;
  Mwa[0] = Mwa[1] = Mwa[2] = 0.0;
  if (flags & kLmAnchorVar) {
    const int a = t.lm_anchor[l];
    double M[18];
    anchor_map(lm + 3 * l, vR + 12 * a, vR[12 * a + 9], M);
    const double* wa = w + 6 * t.vertex_var[a];
    for (int r = 0; r < 3; ++r) {
      double s = 0.0;
      for (int c = 0; c < 6; ++c) s += M[r * 6 + c] * wa[c];
      Mwa[r] = s;
    }
  }
  u[0] = u[1] = u[2] = 0.0;
  if (!(flags & kLmConstant)) {
    if (flags & kLmAnchorVar) {
      double VLM[9];
      sym3_full(acc + 6, VLM);
      for (int a = 0; a < 3; ++a) tv[a] -= VLM[a * 3] * Mwa[0] + VLM[a * 3 + 1] * Mwa[1] + VLM[a * 3 + 2] * Mwa[2];
    }
    double Z[9];
    sym3_full(lm_z + kLmZ * l, Z);
    for (int a = 0; a < 3; ++a) u[a] = Z[a * 3] * tv[0] + Z[a * 3 + 1] * tv[1] + Z[a * 3 + 2] * tv[2];
  }
}

template <bool kF32>
__global__ __launch_bounds__(kObsTile) void k_back_substitute(DevTopology t, const int32_t* __restrict__ tile_begin,
                                                              const double* __restrict__ vR,
                                                              const double* __restrict__ lm,
                                                              const void* __restrict__ obs_rec,
                                                              const double* __restrict__ lm_acc,
                                                              const double* __restrict__ lm_z,
                                                              const double* __restrict__ w, double* __restrict__ u_out,
                                                              double* __restrict__ part) {
  __shared__ double red[kObsTile * 3];
  __shared__ double ush[kObsTile * 6];  // u and M w_a at each segment head
  __shared__ int32_t seg[kObsTile];
  __shared__ double lds4[4];
  const int tid = threadIdx.x;
  const int b = tile_begin[blockIdx.x], e = tile_begin[blockIdx.x + 1];
  if (t.lm_flags[t.obs_landmark[b]] & kLmSplit) {  // whole tile belongs to a split landmark
    if (tid == 0) part[2 * blockIdx.x] = part[2 * blockIdx.x + 1] = 0.0;
    return;
  }
  const int o = b + tid;
  int32_t l = -1;
  uint32_t flags = 0;
  double r[2] = {0, 0}, jl[6], jw[2] = {0, 0}, c[3] = {0, 0, 0};
  bool coupled = false, ovar = false;
  if (o < e) {
    l = t.obs_landmark[o];
    flags = t.lm_flags[l];
  }
  const bool active = o < e && (flags & kLmActive);
  if (active) {
    const RecView<kF32> rec(obs_rec, o);
    const int32_t meta = t.obs_meta[o];
    coupled = obs_type(meta) != kTypeKeyframe;
    ovar = coupled && obs_observer_var(meta);
    r[0] = rec[0];
    r[1] = rec[1];
    for (int k = 0; k < 6; ++k) jl[k] = rec[2 + k];
    if (ovar) {
      const double* wo = w + 6 * t.vertex_var[t.obs_vertex[o]];
      for (int rr = 0; rr < 2; ++rr) {
        double s = 0.0;
        for (int k = 0; k < 6; ++k) s += rec[8 + rr * 6 + k] * wo[k];
        jw[rr] = s;
      }
      for (int a = 0; a < 3; ++a) c[a] = -(jl[a] * jw[0] + jl[3 + a] * jw[1]);
    }
  }
  for (int k = 0; k < 3; ++k) red[tid * 3 + k] = c[k];
  seg[tid] = active ? l : -1;
  __syncthreads();
  double uu = 0.0;
  if (active && (tid == 0 || seg[tid - 1] != l)) {
    const double* acc = lm_acc + (size_t)kLmAcc * l;
    double tv[3] = {acc[12], acc[13], acc[14]};
    for (int k = tid; k < kObsTile && seg[k] == l; ++k)
      for (int a = 0; a < 3; ++a) tv[a] += red[k * 3 + a];
    double u[3], Mwa[3];
    landmark_step(t, l, vR, lm, lm_acc, lm_z, w, tv, u, Mwa);
    uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    for (int a = 0; a < 3; ++a) {
      u_out[3 * l + a] = u[a];
      ush[tid * 6 + a] = u[a];
      ush[tid * 6 + 3 + a] = Mwa[a];
    }
  }
  __syncthreads();
  double mcc = 0.0;
  if (active) {
    const int h = t.lm_begin[l] - b;
    double lu[3];
    for (int a = 0; a < 3; ++a) lu[a] = ush[h * 6 + a] + (coupled ? ush[h * 6 + 3 + a] : 0.0);
    double jy[2];
    for (int rr = 0; rr < 2; ++rr) jy[rr] = jl[rr * 3] * lu[0] + jl[rr * 3 + 1] * lu[1] + jl[rr * 3 + 2] * lu[2] + jw[rr];
    mcc = jy[0] * r[0] + jy[1] * r[1] - 0.5 * (jy[0] * jy[0] + jy[1] * jy[1]);
  }
  const double s1 = block_sum256(mcc, lds4);
  const double s2 = block_sum256(uu, lds4);
  if (tid == 0) {
    part[2 * blockIdx.x] = s1;
    part[2 * blockIdx.x + 1] = s2;
  }
}

// The same for the landmarks split over several tiles, one thread each.
template <bool kF32>
__global__ __launch_bounds__(256) void k_back_substitute_split(DevTopology t, int32_t n_split,
                                                               const int32_t* __restrict__ split_ids,
                                                               const double* __restrict__ vR,
                                                               const double* __restrict__ lm,
                                                               const void* __restrict__ obs_rec,
                                                               const double* __restrict__ lm_acc,
                                                               const double* __restrict__ lm_z,
                                                               const double* __restrict__ w,
                                                               double* __restrict__ u_out, double* __restrict__ part) {
  __shared__ double lds4[4];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  double mcc = 0.0, uu = 0.0;
  if (k < n_split && (t.lm_flags[split_ids[k]] & kLmActive)) {
    const int l = split_ids[k];
    const int lb = t.lm_begin[l], le = t.lm_begin[l + 1];
    const double* acc = lm_acc + (size_t)kLmAcc * l;
    double tv[3] = {acc[12], acc[13], acc[14]};
    for (int o = lb; o < le; ++o) {
      const int32_t meta = t.obs_meta[o];
      if (obs_type(meta) == kTypeKeyframe || !obs_observer_var(meta)) continue;
      const RecView<kF32> rec(obs_rec, o);
      const double* wo = w + 6 * t.vertex_var[t.obs_vertex[o]];
      double jw[2];
      for (int rr = 0; rr < 2; ++rr) {
        double s = 0.0;
        for (int c = 0; c < 6; ++c) s += rec[8 + rr * 6 + c] * wo[c];
        jw[rr] = s;
      }
      for (int a = 0; a < 3; ++a) tv[a] -= rec[2 + a] * jw[0] + rec[5 + a] * jw[1];
    }
    double u[3], Mwa[3];
    landmark_step(t, l, vR, lm, lm_acc, lm_z, w, tv, u, Mwa);
    uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    for (int a = 0; a < 3; ++a) u_out[3 * l + a] = u[a];
    for (int o = lb; o < le; ++o) {
      const int32_t meta = t.obs_meta[o];
      const RecView<kF32> rec(obs_rec, o);
      const bool coupled = obs_type(meta) != kTypeKeyframe;
      double lu[3];
      for (int a = 0; a < 3; ++a) lu[a] = u[a] + (coupled ? Mwa[a] : 0.0);
      double jy[2];
      for (int rr = 0; rr < 2; ++rr) jy[rr] = rec[2 + rr * 3] * lu[0] + rec[3 + rr * 3] * lu[1] + rec[4 + rr * 3] * lu[2];
      if (coupled && obs_observer_var(meta)) {
        const double* wo = w + 6 * t.vertex_var[t.obs_vertex[o]];
        for (int rr = 0; rr < 2; ++rr)
          for (int c = 0; c < 6; ++c) jy[rr] += rec[8 + rr * 6 + c] * wo[c];
      }
      mcc += jy[0] * rec[0] + jy[1] * rec[1] - 0.5 * (jy[0] * jy[0] + jy[1] * jy[1]);
    }
  }
  const double s1 = block_sum256(mcc, lds4);
  const double s2 = block_sum256(uu, lds4);
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = s1; part[2 * blockIdx.x + 1] = s2; }
}

// w = C_p y_p and |w|^2 partials.
__global__ __launch_bounds__(256) void k_unscale_poses(int32_t Pvar, const double* __restrict__ y,
                                                       const double* __restrict__ pose_scale,
                                                       double* __restrict__ w, double* __restrict__ part) {
  __shared__ double lds4[4];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0;
  if (i < 6 * Pvar) {
    const double v = y[i] * pose_scale[i];
    w[i] = v;
    s = v * v;
  }
  const double t = block_sum256(s, lds4);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// Candidate state x (+) delta with delta = -w (vertices, JPL Plus) and
// -u (landmarks); partials of |x|^2 over variable blocks.
// Candidate vertex x [+] (-w) and its rotation record (as k_vertex_prep).
__global__ __launch_bounds__(256) void k_candidate_vertices(DevTopology t, const double* __restrict__ vert,
                                                            const double* __restrict__ w,
                                                            double* __restrict__ cand, double* __restrict__ cand_vR,
                                                            double* __restrict__ part) {
  __shared__ double lds4[4];
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  double xx = 0.0;
  if (v < t.P) {
    const double* x = vert + 7 * v;
    double c[7];
    const int32_t j = t.vertex_var[v];
    if (j >= 0) {
      const double* wv = w + 6 * j;
      const double d[3] = {-wv[0], -wv[1], -wv[2]};
      quat_plus_jpl(x, d, c);
      c[4] = x[4] - wv[3];
      c[5] = x[5] - wv[4];
      c[6] = x[6] - wv[5];
      for (int k = 0; k < 7; ++k) xx += x[k] * x[k];
    } else {
      for (int k = 0; k < 7; ++k) c[k] = x[k];
    }
    double* out = cand + 7 * v;
    for (int k = 0; k < 7; ++k) out[k] = c[k];
    vertex_rotation(c, cand_vR + 12 * v);
  }
  const double s = block_sum256(xx, lds4);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_candidate_landmarks(DevTopology t, const double* __restrict__ lm,
                                                             const double* __restrict__ u,
                                                             double* __restrict__ cand, double* __restrict__ part) {
  __shared__ double lds4[4];
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  double xx = 0.0;
  if (l < t.L) {
    const uint32_t f = t.lm_flags[l];
    const bool var = (f & kLmActive) && !(f & kLmConstant);
    for (int k = 0; k < 3; ++k) {
      const double x = lm[3 * l + k];
      cand[3 * l + k] = var ? x - u[3 * l + k] : x;
      if (var) xx += x * x;
    }
  }
  const double s = block_sum256(xx, lds4);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// max |x - Plus(x, -g)| over variable blocks (ceres gradient_max_norm).
__global__ __launch_bounds__(256) void k_gradient_norm(DevTopology t, const double* __restrict__ vert,
                                                       const double* __restrict__ pose_grad,
                                                       const double* __restrict__ lm_acc,
                                                       double* __restrict__ part) {
  __shared__ double red[4];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  double m = 0.0;
  if (k < t.Pvar) {
    const double* x = vert + 7 * t.var_vertex[k];
    const double* g = pose_grad + 6 * k;
    const double d[3] = {-g[0], -g[1], -g[2]};
    double q[4];
    quat_plus_jpl(x, d, q);
    for (int c = 0; c < 4; ++c) m = fmax(m, fabs(x[c] - q[c]));
    for (int c = 0; c < 3; ++c) m = fmax(m, fabs(g[3 + c]));
  } else if (k - t.Pvar < t.L) {
    const int l = k - t.Pvar;
    const uint32_t f = t.lm_flags[l];
    if ((f & kLmActive) && !(f & kLmConstant)) {
      const double* g = lm_acc + (size_t)kLmAcc * l + 12;
      for (int c = 0; c < 3; ++c) m = fmax(m, fabs(g[c]));
    }
  }
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// Fixed-order final reduction of partial arrays: out[k] = sum (or max) of
// parts[k] (stride `stride`, n entries). One block per output.
__device__ inline void reduce_parts(int32_t n, int32_t stride, int32_t offset, int32_t is_max,
                                    const double* __restrict__ parts, double* __restrict__ out) {
  __shared__ double lds4[4];
  double s = 0.0;
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const double v = parts[(size_t)k * stride + offset];
    s = is_max ? fmax(s, v) : s + v;
  }
  if (is_max) {
    s = wave_max(s);
    __shared__ double red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) *out = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  } else {
    s = block_sum256(s, lds4);
    if (threadIdx.x == 0) *out = s;
  }
}

__global__ void k_reduce(int32_t n, int32_t stride, int32_t offset, int32_t is_max, const double* __restrict__ parts,
                         double* __restrict__ out) {
  reduce_parts(n, stride, offset, is_max, parts, out);
}

// Several independent reductions in one launch, work-group b computing
// reduction b exactly as k_reduce does.
__global__ void k_reduce_many(ReduceBatch batch) {
  const ReduceDesc& d = batch.d[blockIdx.x];
  reduce_parts(d.n, d.stride, d.offset, d.is_max, d.parts, d.out);
}

// ------------------------------------------------------------ Evaluate API

// Every Jacobian block of VisualReprojectionError::Evaluate
// (visual-error-term-inl.h:184-367), in the caller's order, ambient or
// tangent. Observation index is in the original (caller) order via perm.
__global__ __launch_bounds__(256) void k_evaluate_full(DevTopology t, const double* __restrict__ vert,
                                                       const double* __restrict__ vR,
                                                       const double* __restrict__ lm,
                                                       const double* __restrict__ baseframes,
                                                       const double* __restrict__ cam_q, int32_t tangent,
                                                       const int32_t* __restrict__ perm, double* __restrict__ res,
                                                       double* __restrict__ jout, int32_t jstride,
                                                       uint32_t block_mask) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= t.N) return;
  const int32_t meta = t.obs_meta[o];
  const int32_t type = obs_type(meta);
  const int32_t ci = obs_camera(meta);
  const DevCamera& cam = t.cams[ci];
  const int32_t l = t.obs_landmark[o];
  const int32_t v = t.obs_vertex[o];
  const int32_t a = t.lm_anchor[l];
  const double* pB = lm + 3 * l;
  const double* Rv = vR + 12 * v;
  const double* Ra = vR + 12 * a;
  // Frame chain with all intermediate rotations.
  double pI[3], pG[3] = {0, 0, 0};
  double D_pI_pB[9], D_pI_pa[9], D_pI_pGLM[9], D_pI_pGM[9], D_pI_pv[9];
  double D_pI_qa[9], D_pI_qGLM[9], D_pI_qGM[9], D_pI_qv[9];
  for (int k = 0; k < 9; ++k) {
    D_pI_pB[k] = D_pI_pa[k] = D_pI_pGLM[k] = D_pI_pGM[k] = D_pI_pv[k] = 0.0;
    D_pI_qa[k] = D_pI_qGLM[k] = D_pI_qGM[k] = D_pI_qv[k] = 0.0;
  }
  auto skew_right = [](const double* A, const double* p, double* out) {  // A [p]x
    for (int r = 0; r < 3; ++r) {
      out[r * 3 + 0] = A[r * 3 + 1] * p[2] - A[r * 3 + 2] * p[1];
      out[r * 3 + 1] = A[r * 3 + 2] * p[0] - A[r * 3 + 0] * p[2];
      out[r * 3 + 2] = A[r * 3 + 0] * p[1] - A[r * 3 + 1] * p[0];
    }
  };
  if (type == kTypeKeyframe) {
    for (int k = 0; k < 3; ++k) pI[k] = pB[k];
    for (int k = 0; k < 9; ++k) D_pI_pB[k] = (k % 4 == 0) ? 1.0 : 0.0;
  } else {
    double pM[3];
    mat3t_vec(Ra, pB, pM);
    const double* pa = vert + 7 * a + 4;
    for (int k = 0; k < 3; ++k) pM[k] += pa[k];
    double A[9];  // d p_I / d p_M
    for (int k = 0; k < 9; ++k) A[k] = Rv[k];
    double P_M_LM[9];  // d p_M / d p_LM
    for (int k = 0; k < 9; ++k) P_M_LM[k] = (k % 4 == 0) ? 1.0 : 0.0;
    if (type == kTypeGlobal) {
      const double* TL = t.mission_T + 12 * t.vertex_mission[a];
      const double* TM = t.mission_T + 12 * t.vertex_mission[v];
      double pLM[3] = {pM[0], pM[1], pM[2]};
      mat3_vec(TL, pLM, pG);
      for (int k = 0; k < 3; ++k) pG[k] += TL[9 + k];
      double d[3] = {pG[0] - TM[9], pG[1] - TM[10], pG[2] - TM[11]};
      mat3t_vec(TM, d, pM);
      // d p_M / d p_LM = R_M_G R_G_LM
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
          P_M_LM[r * 3 + c] = TM[r] * TL[c] + TM[3 + r] * TL[3 + c] + TM[6 + r] * TL[6 + c];
      double AR[9];  // d p_I / d p_G = R_I_M R_M_G
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) AR[r * 3 + c] = A[r * 3] * TM[c * 3] + A[r * 3 + 1] * TM[c * 3 + 1] + A[r * 3 + 2] * TM[c * 3 + 2];
      for (int k = 0; k < 9; ++k) { D_pI_pGLM[k] = AR[k]; D_pI_pGM[k] = -AR[k]; }
      // Reference forms (visual-error-term-inl.h:208-209) use [p_G]x for both baseframes.
      skew_right(D_pI_pGLM, pG, D_pI_qGLM);
      skew_right(D_pI_pGM, pG, D_pI_qGM);
    }
    const double* pv = vert + 7 * v + 4;
    const double dm[3] = {pM[0] - pv[0], pM[1] - pv[1], pM[2] - pv[2]};
    mat3_vec(Rv, dm, pI);
    for (int k = 0; k < 9; ++k) D_pI_pv[k] = -A[k];
    double T1[9];
    mat3_mul(A, P_M_LM, T1);  // d p_I / d p_LM
    for (int k = 0; k < 9; ++k) D_pI_pa[k] = T1[k];
    mat3_mul_bt(T1, Ra, D_pI_pB);  // * R_LM_B
    skew_right(D_pI_pB, pB, D_pI_qa);
    for (int k = 0; k < 9; ++k) D_pI_qa[k] = -D_pI_qa[k];
    double I3[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    skew_right(I3, pI, D_pI_qv);  // [p_I]x
  }
  double pC[3];
  mat3_vec(cam.R_C_I, pI, pC);
  for (int k = 0; k < 3; ++k) pC[k] += cam.p_C_I[k];
  double kp[2], jp[6];
  const int32_t status = project_camera(cam.model, cam.xi, cam.dist_type, cam.intr, cam.dist, cam.width, cam.height, pC, kp, jp);
  const bool failed = projection_failed(status, pC, kp);
  const double sinv = t.obs_sinv[o];
  const int32_t oo = perm[o];
  if (res) {
    res[2 * oo] = failed ? 0.0 : (kp[0] - t.obs_kp[2 * o]) * sinv;
    res[2 * oo + 1] = failed ? 0.0 : (kp[1] - t.obs_kp[2 * o + 1]) * sinv;
  }
  if (!jout) return;
  // K = d r / d p_I = jp sinv R_C_I; Kc = jp sinv (d r / d p_C).
  double Kc[6], K[6];
  for (int k = 0; k < 6; ++k) Kc[k] = failed ? 0.0 : jp[k] * sinv;
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 3; ++c)
      K[r * 3 + c] = Kc[r * 3] * cam.R_C_I[c] + Kc[r * 3 + 1] * cam.R_C_I[3 + c] + Kc[r * 3 + 2] * cam.R_C_I[6 + c];
  // Block offsets inside the output record of this observation.
  // Block 9 (requested only with the xi flag) is d r / d xi of a
  // unified-projection camera.
  const int amb[10] = {3, 7, 7, 7, 7, 4, 3, 4, 4, 1};
  const int tng[10] = {3, 6, 6, 6, 6, 3, 3, 4, 4, 1};
  double* rec = jout + (size_t)oo * jstride;
  int offs[10];
  int acc_off = 0;
  for (int b = 0; b < 10; ++b) { offs[b] = acc_off; acc_off += 2 * (tangent ? tng[b] : amb[b]); }
  auto zero_block = [&](int b) {
    const int n = 2 * (tangent ? tng[b] : amb[b]);
    for (int k = 0; k < n; ++k) rec[offs[b] + k] = 0.0;
  };
  // Write J = Kx * D (2x3) into columns [c0, c0+3) of a block of width w.
  auto put3 = [&](int b, int c0, const double* Kx, const double* D, double scale) {
    const int w = tangent ? tng[b] : amb[b];
    for (int r = 0; r < 2; ++r)
      for (int c = 0; c < 3; ++c)
        rec[offs[b] + r * w + c0 + c] =
            scale * (Kx[r * 3] * D[c] + Kx[r * 3 + 1] * D[3 + c] + Kx[r * 3 + 2] * D[6 + c]);
  };
  // Rotation block: tangent = J_q |q|^2 ; ambient = J_q * 4 Jq(q)^T.
  auto put_rot = [&](int b, const double* Kx, const double* D, const double* q) {
    double Jr[6];
    for (int r = 0; r < 2; ++r)
      for (int c = 0; c < 3; ++c) Jr[r * 3 + c] = Kx[r * 3] * D[c] + Kx[r * 3 + 1] * D[3 + c] + Kx[r * 3 + 2] * D[6 + c];
    const int w = tangent ? tng[b] : amb[b];
    if (tangent) {
      const double n2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
      for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 3; ++c) rec[offs[b] + r * w + c] = Jr[r * 3 + c] * n2;
    } else {
      // 4 * 0.5 * [[w,-z,y],[z,w,-x],[-y,x,w],[-x,-y,-z]]^T (quaternion-param-jpl.cc:8-32)
      const double x = q[0], y = q[1], z = q[2], w4 = q[3];
      const double JqT[12] = {w4, z, -y, -x, -z, w4, x, -y, y, -x, w4, -z};  // 3x4
      for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 4; ++c)
          rec[offs[b] + r * w + c] =
              2.0 * (Jr[r * 3] * JqT[c] + Jr[r * 3 + 1] * JqT[4 + c] + Jr[r * 3 + 2] * JqT[8 + c]);
    }
  };
  const double* qa = vert + 7 * a;
  const double* qv = vert + 7 * v;
  const double* qGLM = baseframes + 7 * t.vertex_mission[a];
  const double* qGM = baseframes + 7 * t.vertex_mission[v];
  const double* qCI = cam_q + 7 * ci;
  const int rot0 = tangent ? 0 : 0;
  (void)rot0;
  for (int b = 0; b < 10; ++b) {
    if (!((block_mask >> b) & 1u)) continue;
    zero_block(b);
    if (failed) continue;
    const bool used = (b == 0) || (b == 5) || (b == 6) || (b == 7) || (b == 8) ||
                      (b == 9 && cam.model == kCameraUnifiedProjection) ||
                      ((b == 1 || b == 4) && type != kTypeKeyframe) || ((b == 2 || b == 3) && type == kTypeGlobal);
    if (!used) continue;
    const int pc0 = tangent ? 3 : 4;  // column of the position part of a pose block
    switch (b) {
      case 0: put3(0, 0, K, D_pI_pB, 1.0); break;
      case 1: put_rot(1, K, D_pI_qa, qa); put3(1, pc0, K, D_pI_pa, 1.0); break;
      case 2: put_rot(2, K, D_pI_qGLM, qGLM); put3(2, pc0, K, D_pI_pGLM, 1.0); break;
      case 3: put_rot(3, K, D_pI_qGM, qGM); put3(3, pc0, K, D_pI_pGM, 1.0); break;
      case 4: put_rot(4, K, D_pI_qv, qv); put3(4, pc0, K, D_pI_pv, 1.0); break;
      case 5: {
        // d p_C / d q_CI = [R_C_I p_I]x (visual-error-term-inl.h:186)
        double rp[3];
        mat3_vec(cam.R_C_I, pI, rp);
        double I3[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, Sk[9];
        skew_right(I3, rp, Sk);
        put_rot(5, Kc, Sk, qCI);
        break;
      }
      case 6: {
        double I3[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        put3(6, 0, Kc, I3, 1.0);
        break;
      }
      case 7: {
        // d kp / d (fu fv cu cv) at the distorted normalized point.
        const double u = (kp[0] - cam.intr[2]) / cam.intr[0], vv = (kp[1] - cam.intr[3]) / cam.intr[1];
        double* J = rec + offs[7];
        J[0] = u * sinv; J[1] = 0.0; J[2] = sinv; J[3] = 0.0;
        J[4] = 0.0; J[5] = vv * sinv; J[6] = 0.0; J[7] = sinv;
        break;
      }
      case 8: {
        // d kp / d distortion at the undistorted normalized point (the
        // model's projection before distortion), rows x f.
        double* J = rec + offs[8];
        double den = pC[2];
        if (cam.model == kCameraUnifiedProjection)
          den += cam.xi * sqrt(pC[0] * pC[0] + pC[1] * pC[1] + pC[2] * pC[2]);
        const double x = pC[0] / den, y = pC[1] / den;
        double d[8];
        distortion_parameter_jacobian(cam.dist_type, cam.dist, x, y, d);
        for (int k = 0; k < 4; ++k) {
          J[k] = d[k] * cam.intr[0] * sinv;
          J[4 + k] = d[4 + k] * cam.intr[1] * sinv;
        }
        break;
      }
      case 9: {
        // d kp / d xi (camera-unified-projection.cc:176-189).
        double kx[2], jx[6], jxi[2];
        project_unified(cam.xi, cam.dist_type, cam.intr, cam.dist, cam.width, cam.height, pC, kx, jx, jxi);
        rec[offs[9]] = jxi[0] * sinv;
        rec[offs[9] + 1] = jxi[1] * sinv;
        break;
      }
    }
  }
}

// ------------------------------------------------------------ launchers

void launch_vertex_prep(hipStream_t s, int32_t P, const double* vert, double* vR) {
  if (P > 0) k_vertex_prep<<<(P + 255) / 256, 256, 0, s>>>(P, vert, vR);
}
void launch_linearize(hipStream_t s, bool jac, int32_t tiles, const DevTopology& t, const double* vert,
                      const double* vR, const double* lm, DevLoss loss, const int32_t* tile_begin, double* obs_rec,
                      double* lm_acc, double* split_part, double* tile_cost, bool f32) {
  if (tiles <= 0) return;
  if (jac && f32)
    k_linearize<true, true><<<tiles, kObsTile, 0, s>>>(t, vert, vR, lm, loss, tile_begin, obs_rec, lm_acc,
                                                       split_part, tile_cost);
  else if (jac)
    k_linearize<true><<<tiles, kObsTile, 0, s>>>(t, vert, vR, lm, loss, tile_begin, obs_rec, lm_acc, split_part,
                                                 tile_cost);
  else
    k_linearize<false><<<tiles, kObsTile, 0, s>>>(t, vert, vR, lm, loss, tile_begin, obs_rec, lm_acc, split_part,
                                                  tile_cost);
}
void launch_split_sum(hipStream_t s, int32_t n, const int32_t* ids, const int32_t* split_tiles,
                      const double* split_part, double* lm_acc) {
  if (n > 0) k_split_sum<<<(n * kLmAcc + 255) / 256, 256, 0, s>>>(n, ids, split_tiles, split_part, lm_acc);
}
void launch_pose_rows(hipStream_t s, const DevTopology& t, const double* vert, const double* vR, const double* lm,
                      const double* obs_rec, const double* lm_acc, double* pose_grad, double* pose_diag,
                      double* pose_jtj, bool f32) {
  if (t.Pvar <= 0) return;
  if (f32)
    k_pose_rows<true><<<t.Pvar, kRowThreads, 0, s>>>(t, vert, vR, lm, obs_rec, lm_acc, pose_grad, pose_diag,
                                                     pose_jtj);
  else
    k_pose_rows<false><<<t.Pvar, kRowThreads, 0, s>>>(t, vert, vR, lm, obs_rec, lm_acc, pose_grad, pose_diag,
                                                      pose_jtj);
}
void launch_jacobi_scaling(hipStream_t s, int32_t Pvar, int32_t L, int32_t use, const double* pose_diag,
                           const double* lm_acc, double* pose_scale, double* lm_scale) {
  const int n = std::max(6 * Pvar, 3 * L);
  if (n > 0) k_jacobi_scaling<<<(n + 255) / 256, 256, 0, s>>>(Pvar, L, use, pose_diag, lm_acc, pose_scale, lm_scale);
}
void launch_landmark_prep(hipStream_t s, const DevTopology& t, const double* lm_acc, const double* lm_scale,
                          double inv_mu, double min_diag, double max_diag, double* lm_z, int32_t* fail) {
  if (t.L > 0)
    k_landmark_prep<<<(t.L + 255) / 256, 256, 0, s>>>(t, lm_acc, lm_scale, inv_mu, min_diag, max_diag, lm_z, fail);
}
void launch_schur_factors(hipStream_t s, const DevTopology& t, int32_t n_g, const int32_t* g_row,
                          const int32_t* g_inc_ptr, const int32_t* g_inc, const double* vR, const double* lm,
                          const double* obs_rec, const double* lm_acc, const double* pose_scale, double* gw,
                          double* gr0, bool f32) {
  if (t.Pvar <= 0 || n_g <= 0) return;
  if (f32)
    k_schur_factors<true><<<(n_g + 255) / 256, 256, 0, s>>>(t, n_g, g_row, g_inc_ptr, g_inc, vR, lm, obs_rec,
                                                            lm_acc, pose_scale, gw, gr0);
  else
    k_schur_factors<false><<<(n_g + 255) / 256, 256, 0, s>>>(t, n_g, g_row, g_inc_ptr, g_inc, vR, lm, obs_rec,
                                                             lm_acc, pose_scale, gw, gr0);
}
bool schur_items_per_wg_supported(int32_t n) { return n == 1 || n == 2 || n == 4 || n == 8; }
void launch_schur_tiles(hipStream_t s, const SchurTilesDev& d, const double* gw, const double* gr0,
                        const double* lm_acc, const double* lm_z, double* partial, int32_t items_per_wg,
                        int32_t xcd_map) {
  if (d.n_segs <= 0) return;
  const int wgs = (d.n_segs + items_per_wg - 1) / items_per_wg;
  const int per_xcd = (wgs + 7) / 8;
  const TileArgs ta{d.batches, d.slots, d.batch_off, reinterpret_cast<const uint32_t*>(d.pairs), gw, gr0, lm_acc,
                    lm_z};
  k_schur_tiles<<<8 * per_xcd, kTileBlocks, 0, s>>>(d.n_segs, items_per_wg, xcd_map, d.seg_order, d.segs,
                                                     d.seg_thread, ta, partial);
}
void launch_schur_blocks(hipStream_t s, const DevTopology& t, const SchurTilesDev& d, int32_t nub, const int4* ublk,
                         const double* gw, const double* gr0, const double* lm_acc, const double* lm_z,
                         const double* pose_scale, const double* pose_diag, const double* pose_jtj, double inv_mu,
                         double min_diag, double max_diag, double* partial, double* S, double* rhs,
                         int32_t items_per_wg, int32_t xcd_map) {
  if (t.Pvar <= 0 || nub <= 0) return;
  launch_schur_tiles(s, d, gw, gr0, lm_acc, lm_z, partial, items_per_wg, xcd_map);
  k_schur_finalize<<<(nub + 6) / 7, 252, 0, s>>>(nub, ublk, d.uslot_ptr, d.uslot, partial, pose_scale, pose_diag,
                                                 pose_jtj, inv_mu, min_diag, max_diag, S, rhs);
}
void launch_landmark_quality(hipStream_t s, const DevTopology& t, const double* vert, const double* vR,
                             const double* lm, int32_t min_observers, double min_angle_deg, double max_dist,
                             double min_dist, double max_reproj_px, uint8_t* good) {
  if (t.L <= 0) return;
  QualityArgs q;
  q.min_observers = min_observers;
  q.pad = 0;
  q.min_angle_deg = min_angle_deg;
  q.max_dist = max_dist;
  q.min_dist = min_dist;
  q.max_reproj2 = max_reproj_px > 0.0 ? max_reproj_px * max_reproj_px : 0.0;
  k_landmark_quality<<<t.L, 64, 0, s>>>(t, vert, vR, lm, q, good);
}
void launch_block_inverse(hipStream_t s, int32_t Pvar, const int32_t* diag_pos, const double* S, double* Minv,
                          int32_t* fail) {
  if (Pvar > 0) k_block_inverse<<<(Pvar + 63) / 64, 64, 0, s>>>(Pvar, diag_pos, S, Minv, fail);
}

// Part entries: 2 per tile, then 2 per block of 256 split landmarks; returns
// their count.
int32_t launch_back_substitute(hipStream_t s, const DevTopology& t, int32_t tiles, const int32_t* tile_begin,
                               int32_t n_split, const int32_t* split_ids, const double* vR, const double* lm,
                               const double* obs_rec, const double* lm_acc, const double* lm_z, const double* w,
                               double* u, double* part, bool f32) {
  if (tiles > 0) {
    if (f32)
      k_back_substitute<true><<<tiles, kObsTile, 0, s>>>(t, tile_begin, vR, lm, obs_rec, lm_acc, lm_z, w, u, part);
    else
      k_back_substitute<false><<<tiles, kObsTile, 0, s>>>(t, tile_begin, vR, lm, obs_rec, lm_acc, lm_z, w, u, part);
  }
  const int sb = (n_split + 255) / 256;
  double* part_split = part + 2 * (size_t)std::max(tiles, 0);
  if (sb > 0) {
    if (f32)
      k_back_substitute_split<true><<<sb, 256, 0, s>>>(t, n_split, split_ids, vR, lm, obs_rec, lm_acc, lm_z, w, u,
                                                       part_split);
    else
      k_back_substitute_split<false><<<sb, 256, 0, s>>>(t, n_split, split_ids, vR, lm, obs_rec, lm_acc, lm_z, w,
                                                        u, part_split);
  }
  return std::max(tiles, 0) + sb;
}
void launch_unscale_poses(hipStream_t s, int32_t Pvar, const double* y, const double* pose_scale, double* w,
                          double* part) {
  if (Pvar > 0) k_unscale_poses<<<(6 * Pvar + 255) / 256, 256, 0, s>>>(Pvar, y, pose_scale, w, part);
}
void launch_candidate(hipStream_t s, const DevTopology& t, const double* vert, const double* w, double* cand_v,
                      double* cand_vR, double* part_v, const double* lm, const double* u, double* cand_l,
                      double* part_l) {
  if (t.P > 0) k_candidate_vertices<<<(t.P + 255) / 256, 256, 0, s>>>(t, vert, w, cand_v, cand_vR, part_v);
  if (t.L > 0) k_candidate_landmarks<<<(t.L + 255) / 256, 256, 0, s>>>(t, lm, u, cand_l, part_l);
}
void launch_gradient_norm(hipStream_t s, const DevTopology& t, const double* vert, const double* pose_grad,
                          const double* lm_acc, double* part) {
  const int n = t.Pvar + t.L;
  if (n > 0) k_gradient_norm<<<(n + 255) / 256, 256, 0, s>>>(t, vert, pose_grad, lm_acc, part);
}
void launch_reduce(hipStream_t s, int32_t n, int32_t stride, int32_t offset, bool is_max, const double* parts,
                   double* out) {
  k_reduce<<<1, 256, 0, s>>>(n, stride, offset, is_max ? 1 : 0, parts, out);
}
bool launch_reduce_many(hipStream_t s, const ReduceBatch& batch) {
  if (batch.count < 1 || batch.count > kMaxReduceBatch) return false;  // caller reports it; nothing launched
  k_reduce_many<<<batch.count, 256, 0, s>>>(batch);
  return true;
}
// ------------------------------------------------------------ calibration

// Normal equations of the camera calibration blocks when vertices and
// landmarks are constant. Per camera, over the 15 tangent columns
// [q_CI (3) | C_p_CI (3) | fu fv cu cv | distortion (4) | xi], of the
// robust-corrected residuals (ceres Corrector, as k_linearize):
//   entries 0..119  J'J upper triangle (row-major over a <= b),
//   entries 120..134 J'r, entry 135 the cost 0.5 rho.
// One work-group per tile of 256 observations (landmark-sorted order) writes
// its partial sums per camera; k_calib_sum adds the tiles in order.
// `rec` is the tangent k_evaluate_full record of every observation (caller
// order, stride kCalibStride) and `res` its residuals (caller order).
constexpr int kCalibEntries = 136;
constexpr int kCalibStride = 84;  // 2 x (3 + 6 + 6 + 6 + 6 + 3 + 3 + 4 + 4 + 1)

__global__ __launch_bounds__(256) void k_calib_partials(DevTopology t, const int32_t* __restrict__ perm,
                                                        const double* __restrict__ res,
                                                        const double* __restrict__ rec, DevLoss loss,
                                                        int32_t cost_only, double* __restrict__ partial) {
  __shared__ double lds4[4];
  const int o = blockIdx.x * 256 + threadIdx.x;
  int cam = -1;
  double J[30], r[2] = {0.0, 0.0}, cost = 0.0;
  for (int k = 0; k < 30; ++k) J[k] = 0.0;
  if (o < t.N) {
    const int32_t meta = t.obs_meta[o];
    const int32_t l = t.obs_landmark[o];
    if (t.lm_flags[l] & kLmActive) {
      cam = obs_camera(meta);
      const int32_t oo = perm[o];
      const double r0 = res[2 * oo], r1 = res[2 * oo + 1];
      const double s2 = r0 * r0 + r1 * r1;
      double rho[3];
      robust_loss(loss.type, loss_scale(loss, obs_type(meta), t.obs_sinv[o]), s2, rho);
      cost = 0.5 * rho[0];
      if (!cost_only) {
        // Tangent blocks 5 (q_CI), 6 (p_CI), 7 (intrinsics), 8 (distortion), 9 (xi).
        const double* R = rec + (size_t)oo * kCalibStride;
        const int off[5] = {54, 60, 66, 74, 82}, wid[5] = {3, 3, 4, 4, 1}, col[5] = {0, 3, 6, 10, 14};
        for (int b = 0; b < 5; ++b)
          for (int c = 0; c < wid[b]; ++c) {
            J[col[b] + c] = R[off[b] + c];
            J[15 + col[b] + c] = R[off[b] + wid[b] + c];
          }
        const double sq1 = sqrt(rho[1]);
        double rs = sq1, asq = 0.0;
        if (s2 > 0.0 && rho[2] > 0.0) {
          const double alpha = 1.0 - sqrt(1.0 + 2.0 * s2 * rho[2] / rho[1]);
          rs = sq1 / (1.0 - alpha);
          asq = alpha / s2;
        }
        for (int c = 0; c < 15; ++c) {
          const double rtj = J[c] * r0 + J[15 + c] * r1;
          J[c] = sq1 * (J[c] - asq * r0 * rtj);
          J[15 + c] = sq1 * (J[15 + c] - asq * r1 * rtj);
        }
        r[0] = r0 * rs;
        r[1] = r1 * rs;
      }
    }
  }
  for (int c = 0; c < t.C; ++c) {
    double* out = partial + ((size_t)blockIdx.x * t.C + c) * kCalibEntries;
    const bool mine = cam == c;
    if (!cost_only) {
      // Fully unrolled so that J stays in registers.
#pragma unroll
      for (int a = 0; a < 15; ++a)
#pragma unroll
        for (int b = a; b < 15; ++b) {
          const double v = block_sum256(mine ? J[a] * J[b] + J[15 + a] * J[15 + b] : 0.0, lds4);
          if (threadIdx.x == 0) out[a * 15 - a * (a - 1) / 2 + (b - a)] = v;
        }
#pragma unroll
      for (int a = 0; a < 15; ++a) {
        const double v = block_sum256(mine ? J[a] * r[0] + J[15 + a] * r[1] : 0.0, lds4);
        if (threadIdx.x == 0) out[120 + a] = v;
      }
    }
    const double v = block_sum256(mine ? cost : 0.0, lds4);
    if (threadIdx.x == 0) out[135] = v;
  }
}

// out[c][e] = sum over tiles of the partials, in tile order.
__global__ void k_calib_sum(int32_t tiles, int32_t C, const double* __restrict__ partial, double* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= C * kCalibEntries) return;
  const int c = k / kCalibEntries, e = k % kCalibEntries;
  double v = 0.0;
  for (int tile = 0; tile < tiles; ++tile) v += partial[((size_t)tile * C + c) * kCalibEntries + e];
  out[k] = v;
}

int32_t calib_entries() { return kCalibEntries; }
int32_t calib_record_stride() { return kCalibStride; }
void launch_calib_normal(hipStream_t s, const DevTopology& t, const int32_t* perm, const double* res,
                         const double* rec, DevLoss loss, bool cost_only, double* partial, double* out) {
  if (t.N <= 0 || t.C <= 0) return;
  const int tiles = (t.N + 255) / 256;
  k_calib_partials<<<tiles, 256, 0, s>>>(t, perm, res, rec, loss, cost_only ? 1 : 0, partial);
  const int n = t.C * kCalibEntries;
  k_calib_sum<<<(n + 255) / 256, 256, 0, s>>>(tiles, t.C, partial, out);
}

void launch_evaluate_full(hipStream_t s, const DevTopology& t, const double* vert, const double* vR, const double* lm,
                          const double* baseframes, const double* cam_q, bool tangent, const int32_t* perm,
                          double* res, double* jout, int32_t jstride, uint32_t mask) {
  if (t.N > 0)
    k_evaluate_full<<<(t.N + 255) / 256, 256, 0, s>>>(t, vert, vR, lm, baseframes, cam_q, tangent ? 1 : 0, perm, res,
                                                      jout, jstride, mask);
}

// ------------------------------------------------------------ outlier rejection

// p_C of observation o at the current state: the frame chain of
// eval_observation without the Jacobians.
__device__ inline void obs_point_camera(const DevTopology& t, int o, const double* __restrict__ vert,
                                        const double* __restrict__ vR, const double* __restrict__ lm,
                                        double* pC) {
  const int32_t meta = t.obs_meta[o];
  const int32_t type = obs_type(meta);
  const DevCamera& cam = t.cams[obs_camera(meta)];
  const int32_t l = t.obs_landmark[o];
  const int32_t v = t.obs_vertex[o];
  const double* pB = lm + 3 * l;
  double pI[3];
  if (type == kTypeKeyframe) {
    pI[0] = pB[0]; pI[1] = pB[1]; pI[2] = pB[2];
  } else {
    const int32_t a = t.lm_anchor[l];
    double pM[3];
    mat3t_vec(vR + 12 * a, pB, pM);
    for (int i = 0; i < 3; ++i) pM[i] += vert[7 * a + 4 + i];
    if (type == kTypeGlobal) {
      const double* TL = t.mission_T + 12 * t.vertex_mission[a];
      const double* TM = t.mission_T + 12 * t.vertex_mission[v];
      double pG[3], d[3];
      mat3_vec(TL, pM, pG);
      for (int i = 0; i < 3; ++i) d[i] = pG[i] + TL[9 + i] - TM[9 + i];
      mat3t_vec(TM, d, pM);
    }
    const double dm[3] = {pM[0] - vert[7 * v + 4], pM[1] - vert[7 * v + 5], pM[2] - vert[7 * v + 6]};
    mat3_vec(vR + 12 * v, dm, pI);
  }
  mat3_vec(cam.R_C_I, pI, pC);
  for (int i = 0; i < 3; ++i) pC[i] += cam.p_C_I[i];
}

// findOutlierLandmarks (outlier-rejection-solver.cc:28-103) over the active
// landmarks, one wave per landmark: outlier when any observation has
// p_C.z <= 0, or (use_reproj) a squared pixel error above max_reproj2, statuses
// other than visible / outside the image counting as infinite
// (landmark-quality-metrics.cc:57-73). An outlier is deactivated in place
// (kLmActive cleared: its observations then contribute nothing to the cost,
// the landmark sums, S or the step) and bad[l] = 1.
__global__ __launch_bounds__(64) void k_find_outliers(DevTopology t, const double* __restrict__ vert,
                                                      const double* __restrict__ vR, const double* __restrict__ lm,
                                                      int32_t use_reproj, double max_reproj2,
                                                      uint32_t* __restrict__ lm_flags, uint8_t* __restrict__ bad) {
  const int l = blockIdx.x, lane = threadIdx.x;
  const uint32_t flags = lm_flags[l];
  int out = 0;
  if (flags & kLmActive) {
    for (int o = t.lm_begin[l] + lane; o < t.lm_begin[l + 1]; o += 64) {
      double pC[3];
      obs_point_camera(t, o, vert, vR, lm, pC);
      if (pC[2] <= 0.0) { out = 1; continue; }
      if (use_reproj) {
        const DevCamera& cam = t.cams[obs_camera(t.obs_meta[o])];
        double kp[2], jp[6];
        const int st = project_camera(cam.model, cam.xi, cam.dist_type, cam.intr, cam.dist, cam.width, cam.height,
                                      pC, kp, jp);
        double e2 = 1.7976931348623157e308;
        if (st == kVisible || st == kOutsideImage) {
          const double du = kp[0] - t.obs_kp[2 * o], dv = kp[1] - t.obs_kp[2 * o + 1];
          e2 = du * du + dv * dv;
        }
        if (e2 > max_reproj2) out = 1;
      }
    }
  }
  out = __any(out);
  if (lane == 0) {
    bad[l] = (uint8_t)out;
    if (out) lm_flags[l] = flags & ~kLmActive;
  }
}
void launch_find_outliers(hipStream_t s, const DevTopology& t, const double* vert, const double* vR, const double* lm,
                          int32_t use_reproj, double max_reproj2, uint32_t* lm_flags, uint8_t* bad) {
  if (t.L > 0) k_find_outliers<<<t.L, 64, 0, s>>>(t, vert, vR, lm, use_reproj, max_reproj2, lm_flags, bad);
}

// ------------------------------------------------------------ measurement

// Copy of n 16-byte words (the bandwidth reference): each thread moves four
// words of one 4 x 256-word chunk per step (four loads in flight), chunks
// assigned grid-stride; the tail (n % 1024) word by word.
__global__ __launch_bounds__(256) void k_stream_copy(int64_t n, const int4* __restrict__ src, int4* __restrict__ dst) {
  const int64_t chunks = n / 1024;
  for (int64_t c = blockIdx.x; c < chunks; c += gridDim.x) {
    const int64_t b = c * 1024 + threadIdx.x;
    const int4 v0 = src[b], v1 = src[b + 256], v2 = src[b + 512], v3 = src[b + 768];
    dst[b] = v0; dst[b + 256] = v1; dst[b + 512] = v2; dst[b + 768] = v3;
  }
  for (int64_t i = chunks * 1024 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
void launch_stream_copy(hipStream_t s, int64_t words, const void* src, void* dst) {
  if (words > 0)
    k_stream_copy<<<2048, 256, 0, s>>>(words, static_cast<const int4*>(src), static_cast<int4*>(dst));
}

}  // namespace mlba
