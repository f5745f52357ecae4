/*
 * maplab_ba_learned.h -- learned correction of the block-Jacobi preconditioner
 * of the reduced camera system (RCS) and the covisibility-graph network that
 * predicts it. No counterpart in the reference: maplab's Ceres
 * ITERATIVE_SCHUR runs a fixed SCHUR_JACOBI preconditioner
 * (algorithms/map-optimization/include/map-optimization/solver-options.h:
 * 57-68); these entry points let the device solve replace it, with plain
 * block-Jacobi kept as the fallback.
 *
 * With S_ii = L_i L_i' and Sh_ij = L_i^-1 S_ij L_j^-T, the preconditioner is
 * M^-1 = G'G, G = (I - K) L^-1, K strictly block-lower-triangular with the
 * pattern of S: K_ij = theta_i1 Sh_ij + theta_i2 T2_ij over the lower
 * neighbours P(i) of vertex row i, T2_ij = sum_{k in P(i), k != j} Sh_ik Sh_kj.
 * M^-1 is symmetric positive definite for every theta (G is block
 * unit-lower-triangular times L^-1); theta = 0 is block-Jacobi. A 3-layer
 * message-passing network (hidden width 64, BF16 activations on MFMA, FP32
 * parameters) over the camera covisibility graph predicts theta per vertex
 * from node features (diagonal of the vertex's J'J block U, trace of the
 * Schur-reduced block, degrees) and edge features (shared observations,
 * norm of Sh_ij). It is trained to minimize the Kaporin condition number of
 * G S G' row by row: sum_i log(6 - 2 theta_i'c_i + theta_i'Q_i theta_i) with
 * c_i = (a, b), Q_i = [[a + b, b + e], [b + e, e + d]], a = <T1,T1>,
 * b = <T1,T2>, e = <T2,T2>, d = <T2,T3>, T1 = Sh_iP, T3 = T2 (Sh_PP - I).
 *
 * In mlba_solve the preconditioner is chosen by mlba_options.
 * pcg_preconditioner. With MLBA_PRECOND_LEARNED a solve whose PCG does not
 * reach pcg_relative_tolerance within pcg_max_iterations (or breaks down) is
 * repeated with block-Jacobi; the LM step is accepted by the same cost test
 * as every other step. MLBA_PRECOND_LEARNED_DEFLATED applies G'G inside the
 * hat-deflated PCG of MLBA_PRECOND_DEFLATED (without its recycled Ritz
 * columns); its fallback is that deflated PCG with block-Jacobi.
 */
#ifndef MAPLAB_BA_LEARNED_H_
#define MAPLAB_BA_LEARNED_H_

#include <stdint.h>

#include "maplab_ba.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MLBA_PRECOND_DEFLATED 0      /* block-Jacobi with hat and recycled Ritz deflation (default) */
#define MLBA_PRECOND_BLOCK_JACOBI 1  /* plain block-Jacobi */
#define MLBA_PRECOND_LEARNED 2       /* block-Jacobi with the network's correction G'G */
#define MLBA_PRECOND_LEARNED_DEFLATED 3 /* G'G with the hat deflation of MLBA_PRECOND_DEFLATED */

#define MLBA_GNN_NODE_FEATURES 9
#define MLBA_GNN_EDGE_FEATURES 3
#define MLBA_GNN_NORM_FLOATS 24      /* node mean (9), node std (9), edge mean (3), edge std (3) */

/* Number of FP32 parameters of the network. */
int32_t mlba_gnn_num_params(void);

/* The model the handle's learned preconditioner uses: mlba_gnn_num_params()
 * parameters and the feature normalization (x - mean) / std. Copied. */
int32_t mlba_set_preconditioner_model(mlba_handle* handle, const float* params, const float* norm);

typedef struct mlba_preconditioner_stats {
  int64_t solves;               /* PCG solves with the learned preconditioner */
  int64_t iterations;           /* their PCG iterations */
  int64_t fallbacks;            /* solves repeated with block-Jacobi */
  int64_t fallback_iterations;  /* PCG iterations of those repeats */
} mlba_preconditioner_stats;

/* Counters since the handle was created (or the last reset). */
int32_t mlba_get_preconditioner_stats(const mlba_handle* handle, mlba_preconditioner_stats* stats);
int32_t mlba_reset_preconditioner_stats(mlba_handle* handle);

/* Size of the handle's RCS graph: vertex rows and block positions (full
 * symmetric storage, diagonal included). */
int32_t mlba_rcs_size(const mlba_handle* handle, int32_t* rows, int32_t* slots);

/*
 * Training record of the RCS at the current state with trust-region radius
 * `radius` (damping diag / radius, as in an LM iteration): the graph in CSR
 * by row (row_ptr rows+1, col slots, rev slots = position of the transposed
 * block; a slot whose col equals its row is the diagonal, not an edge), raw
 * features (node rows x 9, edge slots x 3, not normalized) and per row
 * (a, b, e, d, has) with has = 1 when the row has lower neighbours. The
 * state is not changed. Every pointer must hold its size.
 */
int32_t mlba_rcs_records(mlba_handle* handle, const mlba_options* options, double radius, int32_t* row_ptr,
                         int32_t* col, int32_t* rev, float* node_features, float* edge_features, float* qc);

/*
 * The RCS itself at the current state with trust-region radius `radius`:
 * S in block CSR by row (row_ptr rows+1, col slots, blocks 36 slots, each
 * 6x6 row-major; full symmetric storage) and its right-hand side (6 rows), in
 * the Jacobi-scaled variables of an LM iteration (no counterpart in the
 * reference: the system Ceres' Schur eliminator forms internally,
 * schur_eliminator_impl.h). The state is not changed.
 */
int32_t mlba_rcs_system(mlba_handle* handle, const mlba_options* options, double radius, int32_t* row_ptr,
                        int32_t* col, double* blocks, double* rhs);

/*
 * PCG iterations to solve the RCS at the current state and radius, with
 * b = its right-hand side, to ||r|| <= tolerance ||b|| with `preconditioner`
 * (MLBA_PRECOND_*; LEARNED and LEARNED_DEFLATED need a model; no fallback
 * and no recycled columns here). iterations: -1 when not converged
 * within max_iterations. The state is not changed.
 */
int32_t mlba_rcs_pcg_iterations(mlba_handle* handle, const mlba_options* options, double radius,
                                int32_t preconditioner, double tolerance, int32_t max_iterations,
                                int32_t* iterations);

/* A batch of RCS graphs for the network (disjoint union, CSR by row as in
 * mlba_rcs_records, node and edge features already normalized). */
typedef struct mlba_gnn_dataset {
  int32_t nodes, slots;
  const int32_t* row_ptr;
  const int32_t* col;
  const int32_t* rev;
  const float* node_features;
  const float* edge_features;
  const float* qc;              /* nodes x 5: a, b, e, d, has */
} mlba_gnn_dataset;

/* Network on the device: theta (nodes x 2, nullable), loss (nullable) and
 * d loss / d params (nullable) at `params`. */
int32_t mlba_gnn_evaluate(const mlba_gnn_dataset* data, const float* params, float* theta, float* loss,
                          float* grad);

/* Adam (lr, beta 0.9 / 0.999) on the device for `epochs` full-batch steps
 * from params_inout; loss_log (nullable) receives the loss before each step. */
int32_t mlba_gnn_train(const mlba_gnn_dataset* data, int32_t epochs, float learning_rate, float* params_inout,
                       float* loss_log);

/* Last error of the dataset-level calls (mlba_gnn_*). */
const char* mlba_gnn_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* MAPLAB_BA_LEARNED_H_ */
