// ba_pcg.h -- host interface of the deflated PCG solve (ba_pcg.hip).
#ifndef MAPLAB_AMD_BA_PCG_H_
#define MAPLAB_AMD_BA_PCG_H_

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace mlba {

// Device buffers of one deflated PCG solve of S y = b over Pvar vertex rows.
// The deflation space is W = [H | R]: H the piecewise-linear hat functions
// over K nodes (6 K columns, two non-zeros per row), R nr dense recycled
// columns (Ritz vectors of earlier solves). SW is stored per 42-row group
// over the group's window: gwh hat columns from node grp_lo[g], then the nr
// dense columns.
struct PcgBuffers {
  int32_t Pvar, K, nr, gwh;
  int32_t zcap;                // Lanczos vectors kept (iterations), 0: none
  int32_t partials;            // sum of the group partials: 0 automatic, 1 per SpMV work-group, 2 once
  const int32_t* row_ptr;      // block rows of S
  const int32_t* col_var;
  const double* S;
  const double* Minv;          // inverse diagonal blocks
  const double* b;
  const int32_t* hat_node;     // Pvar: first hat node of the vertex
  const double* hat_w;         // 2 Pvar: weights of hat_node and hat_node + 1
  const int32_t* node_range;   // 2 K: first and last vertex of each hat's support
  const int32_t* grp_lo;       // groups: first hat node of the group's window
  const double* R;             // nr x 6 Pvar, column-major
  const int32_t* blk_hat;      // per block of S: hat node of its column vertex
  const double* blk_w;         // per block of S: the two hat weights of its column vertex
  const int32_t* blk_rng;      // Pvar x K x 2: first and end block of row i in the support of hat k
  double* SR;                  // groups x (gwh + nr) x 252
  double* E;                   // ne x ne (ne = 6 K + nr)
  double* Einv;                // ne x ne + groups x (gwh + nr) x ne
  double* y;                   // ne: E^-1 W'b
  double* cden;                // nr: R'b
  double* x;
  double* r;
  double* z;
  double* p0;
  double* p1;
  double* q;
  double* part;                // Pvar
  double* part2;               // (groups + 1) x (3 + ne): per-group partials, then their sum
  double* Zh;                  // zcap x 6 Pvar: deflated preconditioned residual of iteration k
  double* coef;                // 2 zcap: r'z and alpha of iteration k
  void* state_dev;             // pcg_state_bytes()
  void* state_host;            // pinned, pcg_state_bytes()
  // Learned preconditioner z = L^-T (I - K)' (I - K) L^-1 r in place of
  // block-Jacobi (Minv) when kmat is set (no recycled columns, nr = zcap = 0).
  const double* linv;          // 36 Pvar: L_i^-1 of S_ii = L_i L_i' (lower triangle, row-major)
  const float* kmat;           // 36 nnzb: K_ij on the strictly lower blocks
  const int32_t* diag_pos;     // Pvar: position of the diagonal block of each row
  const int32_t* tpos;         // nnzb: position of the transposed block
  double* t;                   // 6 Pvar: L^-1 r
  double* u;                   // 6 Pvar: (I - K) L^-1 r
};

size_t pcg_state_bytes();
int32_t pcg_max_hat_nodes();
// Recycled dense columns of the deflation space, at most.
constexpr int kPcgMaxRitz = 16;
int32_t pcg_max_ritz();
int32_t pcg_group_rows();
int32_t pcg_update_groups(int32_t Pvar);
// Whether the last solve's coarse matrix E was positive definite (its
// deflation was applied); reads the host copy of the state.
bool pcg_deflation_applied(const void* state_host);

// Deflated PCG solve. The host polls the state after `first` iterations, then
// every `poll` iterations. With `probe` (three events), k_pcg_spmv and then
// k_pcg_update of one more iteration are each launched `probe_reps` times
// back to back after the solve and timed as two batches (block-Jacobi only).
// Returns the iteration count (max_iters: not converged) or -1 on a launch
// error or an unsupported combination.
int launch_pcg_deflated(hipStream_t s, const PcgBuffers& B, double tol, int32_t max_iters, int32_t first,
                        int32_t poll, const hipEvent_t* probe, int32_t probe_reps);

// Recycling the deflation space after a solve (m Lanczos vectors kept):
// Rn = Zh' Y (r columns; Y is m x r, row-major, already divided by the
// square roots of the iterations' r'z).
void launch_ritz_form(hipStream_t s, int32_t Pvar, int32_t m, int32_t r, const double* Zh, const double* Y,
                      double* Rn);
// SRn = S Rn (r dense columns).
void launch_ritz_spmv(hipStream_t s, int32_t Pvar, int32_t r, const int32_t* row_ptr, const int32_t* col_var,
                      const double* S, const double* Rn, double* SRn);
// Gram matrices of Y = [R (B.nr columns, with S R in B.SR) | Rn (r columns,
// S Rn = SRn)]: AB[0 .. nb^2) = Y'SY, AB[nb^2 .. 2 nb^2) = Y'MY with M the
// diagonal blocks of S (diag_pos), nb = B.nr + r, row-major.
void launch_ritz_gram(hipStream_t s, const PcgBuffers& B, int32_t r, const double* Rn, const double* SRn,
                      const int32_t* diag_pos, double* AB);
// Rnext = [R | Rn] Cm, Cm (nr_old + r) x r_next row-major.
void launch_ritz_merge(hipStream_t s, int32_t Pvar, const double* R, int32_t nr_old, const double* Rn, int32_t r,
                       const double* Cm, int32_t r_next, double* Rnext);

}  // namespace mlba

#endif  // MAPLAB_AMD_BA_PCG_H_
