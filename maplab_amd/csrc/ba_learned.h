// ba_learned.h -- host interface of the learned correction to the
// block-Jacobi preconditioner of the reduced camera system (ba_learned.hip)
// and of the covisibility-graph network that predicts it (ba_gnn.hip).
//
// Preconditioner. With S_ii = L_i L_i' (Cholesky of the diagonal blocks) and
// the normalized blocks Sh_ij = L_i^-1 S_ij L_j^-T, the preconditioner is
//   M^-1 = G' G,   G = (I - K) L^-1,
// K strictly block-lower-triangular with the pattern of S: for the lower
// neighbours P(i) = {j < i : S_ij != 0} of row i,
//   K_ij = theta_i1 T1_ij + theta_i2 T2_ij,   T1_ij = Sh_ij,
//   T2_ij = sum_{k in P(i), k != j} Sh_ik Sh_kj,
// the first two terms of the neighbourhood Neumann series of the factorized
// sparse approximate inverse (Kolotilina & Yeremin), with per-row
// coefficients theta_i. K = 0 is block-Jacobi. G is unit block-lower-
// triangular times L^-1, so M^-1 is symmetric positive definite for every
// theta. The coefficients that minimize the Kaporin condition number
// tr(G S G') / det(G S G')^(1/n) row by row (with the rows' optimal scales)
// minimize log t_i(theta) = log(6 - 2 theta'c_i + theta'Q_i theta), with
//   c_i = (a, b),  Q_i = [[a + b, b + e], [b + e, e + d]],
//   a = <T1, T1>, b = <T1, T2>, e = <T2, T2>, d = <T2, T3>,  T3 = T2 (Sh_PP - I),
// sums over the row's lower pattern. The network predicts theta_i from the
// graph; the four traces are the training signal.
#ifndef MAPLAB_AMD_BA_LEARNED_H_
#define MAPLAB_AMD_BA_LEARNED_H_

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace mlba {

constexpr int kLpNodeFeatures = 9;   // log(diag(U) s^2) (6), log tr S_ii, log(1 + deg), log(1 + lower deg)
constexpr int kLpEdgeFeatures = 3;   // log(1 + shared observations), |Sh_ij|_F, +1 lower / -1 upper
constexpr int kLpTerms = 2;          // theta per row

// Reduced camera system structure the preconditioner works on (block
// positions of S, full symmetric storage, columns sorted in each row).
struct LpSystem {
  int32_t Pvar, nnzb;
  const int32_t* row_ptr;
  const int32_t* col_var;
  const int32_t* diag_pos;
  const int32_t* pos_row;      // nnzb: row of each block position
  const int32_t* tpos;         // nnzb: position of the transposed block
  const double* S;             // 36 nnzb
};

// Set-up. Returns false when a diagonal block is not positive definite
// (fail counter > 0 on the device is the reason; the host reads it).
// linv: 36 Pvar (L_i^-1, lower triangle, row-major); shat: 36 nnzb FP32
// (identity at the diagonal positions); snorm: nnzb FP32 |Sh_ij|_F.
void launch_lp_factor(hipStream_t s, const LpSystem& sys, double* linv, int32_t* fail);
void launch_lp_normalize(hipStream_t s, const LpSystem& sys, const double* linv, float* shat, float* snorm);

// Graph features of the network, normalized with (mean, std) pairs:
// node [Pvar x kLpNodeFeatures], edge [nnzb x kLpEdgeFeatures] (zero rows at
// the diagonal positions). ujtj: 36 Pvar sum Jo'Jo of each vertex (U),
// scale: 6 Pvar Jacobi column scales; shared: nnzb observation pairs.
void launch_lp_features(hipStream_t s, const LpSystem& sys, const double* ujtj, const double* scale,
                        const float* snorm, const int32_t* shared, const float* norm, float* node_feat,
                        float* edge_feat);

// Neighbourhood products of every row, by mode:
//   kLpAssemble: K_ij = theta_i1 Sh_ij + theta_i2 T2_ij into kmat (36 nnzb FP32, lower positions);
//   kLpStoreT2:  T2_ij into t2 (36 nnzb FP32, lower positions) and traces a, b, e into trace (4 Pvar);
//   kLpTraceT3:  d = <T2, T3> of each row into trace[4 i + 3] (reads t2).
// Returns false when a row does not fit the work-group's LDS (Pvar or the
// row's lower degree too large); nothing is launched then.
enum LpMode { kLpAssemble = 0, kLpStoreT2 = 1, kLpTraceT3 = 2 };
bool launch_lp_terms(hipStream_t s, const LpSystem& sys, int32_t max_lower_deg, LpMode mode, const float* shat,
                     const float* theta, float* kmat, float* t2, float* trace);

// PCG on S x = b preconditioned with M^-1 = G'G (use_k) or block-Jacobi
// (K = 0). Device-side convergence: after every iteration a single
// work-group sums the partials in a fixed order and sets the done flag when
// |r| <= tol |b| (or on breakdown); every launch after that returns at once.
// The host polls the state every `poll` iterations.
struct LpPcgBuffers {
  LpSystem sys;
  const double* linv;
  const float* kmat;           // nullptr: block-Jacobi
  const double* b;
  double* x;
  double* r;
  double* t;                   // L^-1 r
  double* u;                   // (I - K) t
  double* z;
  double* p0;
  double* p1;
  double* q;
  double* part;                // Pvar x 2 row partials
  void* state_dev;             // lp_pcg_state_bytes()
  void* state_host;            // pinned, lp_pcg_state_bytes()
};
size_t lp_pcg_state_bytes();
// Returns iterations (>= 0), or -1 on a launch error. *converged = 0 when the
// iteration cap was reached or the solve broke down (p'q <= 0 or not finite).
int launch_lp_pcg(hipStream_t s, const LpPcgBuffers& B, double tol, int32_t max_iters, int32_t poll,
                  int32_t* converged);

// ---------------------------------------------------------------- network

// Parameter vector layout (FP32) of the 3-layer message-passing network with
// hidden width 64 over the covisibility graph:
//   W_in [9 x 64], b_in [64],
//   per layer l: W_m [64 x 64], b_m [64], a_g [3], c_g, W_u [128 x 64], b_u [64],
//   W_out [64 x 2], b_out [2].
constexpr int kGnnHidden = 64;
constexpr int kGnnLayers = 3;
constexpr int kGnnLayerParams = 64 * 64 + 64 + kLpEdgeFeatures + 1 + 128 * 64 + 64;
constexpr int kGnnParams = kLpNodeFeatures * 64 + 64 + kGnnLayers * kGnnLayerParams + 64 * kLpTerms + kLpTerms;
constexpr int kGnnNormFloats = 2 * (kLpNodeFeatures + kLpEdgeFeatures);  // node mean, std, edge mean, std

// A graph batch on the device in CSR by destination: edges of row i are
// [row_ptr[i], row_ptr[i+1]), source col[e]; skip[e] = 1 marks positions that
// are not edges (the diagonal of S); rev[e] = position of the reverse edge.
struct GnnGraph {
  int32_t nodes, slots;
  const int32_t* row_ptr;
  const int32_t* col;
  const int32_t* rev;
  const float* node_feat;      // nodes x 9 (normalized)
  const float* edge_feat;      // slots x 3 (normalized)
};

class GnnWorkspace;  // activations of one graph batch (ba_gnn.hip)
GnnWorkspace* gnn_workspace_create(int32_t nodes, int32_t slots, bool for_training);
void gnn_workspace_destroy(GnnWorkspace* w);

// theta [nodes x 2] = network(graph) with the FP32 parameters `params`
// (kGnnParams). Activations are BF16, dense layers on MFMA with FP32
// accumulation. Returns false on a launch error.
bool gnn_forward(hipStream_t s, GnnWorkspace* w, const GnnGraph& g, const float* params, float* theta);

// Loss inv_count * sum_i has_i log t_i(theta_i) (qc: 5 per node = a, b, e, d,
// has; inv_count = 1 / sum_i has_i) and its gradient with respect to the
// parameters (kGnnParams FP32) after gnn_forward on the same workspace.
// loss_out: device scalar.
bool gnn_loss_backward(hipStream_t s, GnnWorkspace* w, const GnnGraph& g, const float* params, const float* theta,
                       const float* qc, float inv_count, float* loss_out, float* grad);

// One Adam step (beta 0.9 / 0.999, eps 1e-8) on the parameters.
void launch_adam(hipStream_t s, int32_t n, float lr, int32_t step, const float* grad, float* params, float* m,
                 float* v);

}  // namespace mlba

#endif  // MAPLAB_AMD_BA_LEARNED_H_
