// ba_pcg.hip -- deflated, block-Jacobi preconditioned conjugate gradients on
// the reduced camera system S y = b (FP64).
//
// Block-Jacobi (inverse 6x6 diagonal blocks of S, Ceres' SCHUR_JACOBI) leaves
// a handful of very small eigenvalues on long trajectories: smooth, global
// modes of the vertex chain. They are removed by deflation (Saad, Yeung,
// Erhel, Guyomarc'h, "A deflated version of the conjugate gradient
// algorithm", SIAM J. Sci. Comput. 21, 2000) with a coarse space
// W = [H | R]: H the piecewise-linear "hat" functions over the vertex order
// (K nodes, 6 K columns), R up to 16 dense columns recycled from earlier
// solves (Ritz vectors of the preconditioned operator for its smallest
// eigenvalues, from each solve's Lanczos coefficients; see the ritz kernels
// below). x0 = W E^-1 W'b, E = W'SW, and every search direction is made
// S-orthogonal to W with mu = E^-1 (SW)'z. The iteration converges to the
// same solution as plain PCG (the tolerance is on the true residual
// recursion): the hat space cuts iterations 3-4x on long random-walk maps,
// the recycled vectors another 2-3x on the solves that follow a rejected
// LM step (same linearization, new damping).
//
// Rows of S are handled in groups of 42 vertices (252 entries); SW is stored
// per group, column-major over the group's window of coarse columns (the hat
// columns its rows couple to, then the dense ones), so that (SW)'z of a group
// is a set of coalesced column dot products.
//
// Per LM step: k_defl_sr (SW), k_defl_e and k_defl_dense (E, R'b),
// k_defl_prep (state, E^-1 by Gauss-Jordan, y = E^-1 W'b), k_defl_init (x0,
// r0, z0 and the first per-group partials). Per PCG iteration: k_pcg_spmv (sums the previous
// per-group partials; p = z - W mu + beta p_prev and
// q = S z + beta S p_prev - (SW) mu) and k_pcg_update (x, r, z and the
// per-group partials r'z, r'r, E^-1 (SW)'z). Every sum of partials runs in
// the same fixed order, so the set-up and the iteration are deterministic.
// With more groups than one LDS chunk of k_pcg_spmv holds, the work-group of
// k_pcg_update (or k_defl_init) that finishes last sums the groups' partials
// once, into part1, and k_pcg_spmv reads that one vector; the only atomic is
// the integer arrival counter that picks this work-group.
//
// The deflation holds for any SPD preconditioner: with the learned one
// (z = L^-T (I - K)' (I - K) L^-1 r, K strictly block-lower-triangular),
// k_defl_init and k_pcg_update stop at t = L^-1 r, and k_pcg_lower,
// k_pcg_upper (the two row sweeps) and k_defl_partials (the same group
// partials) follow; the hat space is kept, the recycled columns are not
// (their Rayleigh-Ritz step is taken with respect to block-Jacobi's M).
#include "ba_pcg.h"

#include <algorithm>

#include "ba_device.h"
#include "ba_reduce.h"

namespace mlba {

constexpr int kMaxHatNodes = 15;
constexpr int kMaxRitz = kPcgMaxRitz;
constexpr int kMaxCoarse = 6 * kMaxHatNodes + kMaxRitz;  // 106 columns
constexpr int kRowsPerGroup = 42;
constexpr int kGroupThreads = 6 * kRowsPerGroup;  // 252 entries (row, component) per group
constexpr int kGroupBlock = 1024;                 // threads of the per-group kernels
constexpr int kPartChunk = 2048;                  // doubles of partials summed per LDS pass

struct PcgState {
  double rz[2];     // r'z of iteration k in rz[k & 1] (written by k_pcg_spmv)
  double bb;        // |b|^2
  double tol_rel2;  // tol^2
  double tol2;      // (tol |b|)^2
  int32_t iters;
  int32_t done;
  int32_t max_iters;
  int32_t defl_ok;  // E was positive definite: the deflation is applied
  uint32_t arrive;  // work-groups that wrote their partials in the current launch
  int32_t pad;
};

struct PcgArgs {
  int32_t Pvar, ne;            // ne = neh + nr coarse columns (0: no deflation)
  int32_t neh, nr;             // hat (6 K) and dense columns
  int32_t K, groups;
  int32_t gwh, n6;             // hat columns of a group window; 6 Pvar
  int32_t zcap, pad3;
  const int32_t* row_ptr;
  const int32_t* col_var;
  const double* S;
  const double* Minv;
  const double* b;
  const int32_t* hat_node;     // Pvar: first hat node of the vertex (0 without deflation)
  const double* hat_w;         // 2 Pvar: weights of hat_node and hat_node + 1 (0 without deflation)
  const int32_t* node_range;   // 2 K: first and last vertex of each hat's support
  const int32_t* grp_lo;       // groups: first hat node of the group's window of SW columns
  int32_t gw, pad2;            // window width in columns (gwh hat + nr dense)
  const double* R;             // nr x 6 Pvar dense columns
  const int32_t* blk_hat;      // per block of S: hat node of its column vertex
  const double* blk_w;         // per block of S: the two hat weights of its column vertex
  const int32_t* blk_rng;      // Pvar x K x 2: blocks of row i in the support of hat k
  double* SR;                  // groups x gw x 252: SW entry (group row t, window column lc)
  double* E;                   // ne x ne
  double* Einv;                // ne x ne
  double* EinvW;               // groups x gw x ne: the rows of E^-1 of each group's window columns
  double* x;
  double* r;
  double* z;
  double* p0;
  double* p1;
  double* q;
  double* part;                // Pvar: p'q of each row
  double* part2;               // groups x (3 + ne): r'z, r'r, b'b and E^-1 (SW)'z of the group's rows
  double* part1;               // 3 + ne: part2 summed over the groups (with reduce1)
  int32_t reduce1, pad4;       // the groups' partials exceed one LDS chunk of k_pcg_spmv
  double* y;                   // ne: E^-1 W'b
  double* cden;                // nr: R'b
  double* Zh;                  // zcap x 6 Pvar: deflated preconditioned residuals
  double* coef;                // 2 zcap: r'z and alpha per iteration
  PcgState* st;
  const double* linv;          // learned preconditioner (with kmat): L_i^-1, 36 Pvar
  const float* kmat;           // K_ij, 36 nnzb
  const int32_t* diag_pos;
  const int32_t* tpos;
  double* t;                   // L^-1 r
  double* u;                   // (I - K) L^-1 r
};

__device__ inline double hat_weight(const PcgArgs& a, int i, int k) {
  const int n0 = a.hat_node[i];
  return n0 == k ? a.hat_w[2 * i] : (n0 + 1 == k ? a.hat_w[2 * i + 1] : 0.0);
}

__device__ inline double* sr_col(const PcgArgs& a, int g, int lc) {
  return a.SR + ((size_t)g * a.gw + lc) * kGroupThreads;
}

// Coarse column of window column lc of a group whose hat window starts at
// column c0 (6 x its first node); -1 past the last hat column.
__device__ inline int win_col(const PcgArgs& a, int c0, int lc) {
  if (lc >= a.gwh) return a.neh + (lc - a.gwh);
  return c0 + lc < a.neh ? c0 + lc : -1;
}

// Block sums of three values (blockDim.x = 64 x waves, at most 16 waves), in
// every thread.
__device__ inline void block_sum3(double& v0, double& v1, double& v2, double* lds) {
  const int nw = blockDim.x >> 6, w = threadIdx.x >> 6;
  v0 = wave_sum(v0);
  v1 = wave_sum(v1);
  v2 = wave_sum(v2);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    lds[w] = v0;
    lds[16 + w] = v1;
    lds[32 + w] = v2;
  }
  __syncthreads();
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int k = 0; k < nw; ++k) {
    s0 += lds[k];
    s1 += lds[16 + k];
    s2 += lds[32 + k];
  }
  v0 = s0;
  v1 = s1;
  v2 = s2;
}

// SW for the hat columns, one block row of S per work-group:
// (SH)_i = sum_j S_ij H_j over the group window's hat columns, written into
// the group window. Thread o owns entry (component o % 6, window column
// o / 6) of hat node k and sums, in order, the row's blocks whose vertex lies
// in the hat's support: a contiguous range of the row (blk_rng), each with
// weight w0 or w1 of its vertex (its first hat node is k or k - 1).
__global__ __launch_bounds__(256) void k_defl_sr(PcgArgs a) {
  const int i = blockIdx.x, g = i / kRowsPerGroup, t0 = 6 * (i - g * kRowsPerGroup);
  const int c0 = 6 * a.grp_lo[g];
  for (int o = threadIdx.x; o < 6 * a.gwh; o += blockDim.x) {
    const int comp = o % 6, lc = o / 6, col = win_col(a, c0, lc);
    double acc = 0.0;
    if (col >= 0) {
      const int k = col / 6, cc = col % 6;
      const int lo = a.blk_rng[2 * (i * a.K + k)], hi = a.blk_rng[2 * (i * a.K + k) + 1];
#pragma unroll 4
      for (int kb = lo; kb < hi; ++kb) {
        const int n0 = a.blk_hat[kb];
        const double w0 = a.blk_w[2 * kb], w1 = a.blk_w[2 * kb + 1];
        acc += (n0 == k ? w0 : w1) * a.S[(size_t)36 * kb + 6 * comp + cc];
      }
    }
    sr_col(a, g, lc)[t0 + comp] = acc;
  }
}

// SW for the dense columns (or S Rn for the recycling), one block row per
// work-group of six waves: wave w takes the dense columns 3 w .. 3 w + 2 for
// all six components (so R_j is read once per block, not once per
// component), lane l sums over the row's blocks l, l + 64, ... the products
// S_ij[comp][:] . R_c[6 j ..], then a fixed-order wave sum per (component,
// column). Writes the group window (dense part) or a plain column-major
// buffer.
constexpr int kDenseThreads = 384;
constexpr int kDenseColsPerWave = (kMaxRitz + kDenseThreads / 64 - 1) / (kDenseThreads / 64);  // 3
template <bool kWindow>
__global__ __launch_bounds__(kDenseThreads) void k_row_dense(PcgArgs a, const int32_t* __restrict__ row_ptr,
                                                             const int32_t* __restrict__ col_var,
                                                             const double* __restrict__ S,
                                                             const double* __restrict__ R, int32_t nr, int32_t n6,
                                                             double* __restrict__ out) {
  const int i = blockIdx.x, wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int c0 = wave * kDenseColsPerWave;
  if (c0 >= nr) return;  // the whole wave: no barrier in this kernel
  const int rb = row_ptr[i], re = row_ptr[i + 1];
  double acc[6][kDenseColsPerWave];
#pragma unroll
  for (int comp = 0; comp < 6; ++comp)
#pragma unroll
    for (int m = 0; m < kDenseColsPerWave; ++m) acc[comp][m] = 0.0;
  for (int kb = rb + lane; kb < re; kb += 64) {
    const double* sb = S + (size_t)36 * kb;
    double sv[36];
#pragma unroll
    for (int e = 0; e < 36; ++e) sv[e] = sb[e];
    const double* rj = R + 6 * col_var[kb];
#pragma unroll
    for (int m = 0; m < kDenseColsPerWave; ++m) {
      if (c0 + m < nr) {
        const double* r = rj + (size_t)(c0 + m) * n6;
        const double r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3], r4 = r[4], r5 = r[5];
#pragma unroll
        for (int comp = 0; comp < 6; ++comp) {
          const double* s6 = sv + 6 * comp;
          acc[comp][m] += ((s6[0] * r0 + s6[1] * r1) + (s6[2] * r2 + s6[3] * r3)) + (s6[4] * r4 + s6[5] * r5);
        }
      }
    }
  }
#pragma unroll
  for (int m = 0; m < kDenseColsPerWave; ++m) {
    const int c = c0 + m;
    if (c >= nr) break;
#pragma unroll
    for (int comp = 0; comp < 6; ++comp) {
      const double v = wave_sum(acc[comp][m]);
      if (lane == 0) {
        if (kWindow) {
          const int g = i / kRowsPerGroup;
          sr_col(a, g, a.gwh + c)[6 * (i - g * kRowsPerGroup) + comp] = v;
        } else {
          out[(size_t)c * n6 + 6 * i + comp] = v;
        }
      }
    }
  }
}

// Entry (e, col) of SW from the group windows.
__device__ inline double sw_entry(const PcgArgs& a, int e, int col) {
  const int g = e / kGroupThreads;
  const int lc = col >= a.neh ? a.gwh + (col - a.neh) : col - 6 * a.grp_lo[g];
  return (lc >= 0 && lc < a.gw) ? sr_col(a, g, lc)[e - g * kGroupThreads] : 0.0;
}

// E = W'(SW) for the hat rows: one wave per (hat node k, column oc) sums the
// six rows (k, comp) over the hat's support (lanes take every 64th support
// row, then a fixed-order wave sum). A dense column's entries are also the
// transposed (dense, hat) entries. The dense-dense block is k_defl_dense's.
__global__ __launch_bounds__(64) void k_defl_e(PcgArgs a) {
  const int k = blockIdx.x, oc = blockIdx.y, lane = threadIdx.x;
  double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int i = a.node_range[2 * k] + lane; i <= a.node_range[2 * k + 1]; i += 64) {
    const double w = hat_weight(a, i, k);
#pragma unroll
    for (int comp = 0; comp < 6; ++comp) s[comp] += w * sw_entry(a, 6 * i + comp, oc);
  }
#pragma unroll
  for (int comp = 0; comp < 6; ++comp) {
    const double v = wave_sum(s[comp]);
    if (lane == 0) {
      a.E[(size_t)(6 * k + comp) * a.ne + oc] = v;
      if (oc >= a.neh) a.E[(size_t)oc * a.ne + 6 * k + comp] = v;
    }
  }
}

// Dense block of E and R'b, one work-group per entry: (a, b < nr) is
// R_a'(SR)_b over all rows, (a, nr) is R_a'b.
__global__ __launch_bounds__(256) void k_defl_dense(PcgArgs a) {
  __shared__ double lds[4];
  const int ra = blockIdx.x, cb = blockIdx.y;
  const double* Ra = a.R + (size_t)ra * a.n6;
  double s = 0.0;
  for (int e = threadIdx.x; e < a.n6; e += blockDim.x)
    s += Ra[e] * (cb < a.nr ? sw_entry(a, e, a.neh + cb) : a.b[e]);
  s = block_sum256(s, lds);
  if (threadIdx.x == 0) {
    if (cb < a.nr)
      a.E[(size_t)(a.neh + ra) * a.ne + a.neh + cb] = s;
    else
      a.cden[ra] = s;
  }
}

constexpr int kPrepThreads = 1024;
constexpr int kTri = kMaxCoarse * (kMaxCoarse + 1) / 2;  // entries of the lower triangle
constexpr int kInvPerThread = (kTri + kPrepThreads - 1) / kPrepThreads;

// Solve set-up (one work-group): the PCG state, c = W'b, E^-1 by the sweep
// operator on the lower triangle (Goodnight, "A tutorial on the SWEEP
// operator", Amer. Statistician 33, 1979: sweeping pivot k of a symmetric A
// maps a_kk -> -1/a_kk, a_ik -> a_ik / a_kk, a_ij -> a_ij - a_ik a_kj / a_kk,
// keeps A symmetric, and sweeping every pivot gives -A^-1; E is SPD, so the
// pivots are the positive Cholesky pivots and no pivoting is needed). Each
// thread keeps its lower-triangle entries in registers, pivot column k goes
// through LDS. Then y = E^-1 c. A non-positive pivot disables deflation for
// this solve (E^-1 = 0, y = 0, defl_ok = 0).
__global__ __launch_bounds__(kPrepThreads) void k_defl_prep(PcgArgs a, double tol, int32_t max_iters) {
  __shared__ double Ms[kMaxCoarse * kMaxCoarse];
  __shared__ double cpart[kPrepThreads];
  __shared__ double cvec[kMaxCoarse];
  __shared__ double colb[2][kMaxCoarse];  // column k of the swept matrix in colb[k & 1]
  const int ne = a.ne, tid = threadIdx.x;
  if (tid == 0) {
    PcgState* st = a.st;
    st->rz[0] = st->rz[1] = 0.0;
    st->bb = 0.0;
    st->tol_rel2 = tol * tol;
    st->tol2 = 0.0;
    st->iters = 0;
    st->done = 0;
    st->max_iters = max_iters;
    st->defl_ok = 0;
    st->arrive = 0;
  }
  if (ne == 0) return;
  // c = W'b: thread (column, g) sums the support rows lo + g, lo + g + G, ...
  // of a hat column; dense columns come from k_defl_dense.
  {
    const int G = kPrepThreads / ne, col = tid % ne, g = tid / ne, k = col / 6, comp = col % 6;
    double acc = 0.0;
    if (g < G) {
      if (col < a.neh) {
        for (int i = a.node_range[2 * k] + g; i <= a.node_range[2 * k + 1]; i += G)
          acc += hat_weight(a, i, k) * a.b[6 * i + comp];
      } else if (g == 0) {
        acc = a.cden[col - a.neh];
      }
    }
    cpart[tid] = acc;
    __syncthreads();
    if (tid < ne) {
      double c = 0.0;
      for (int gg = 0; gg < G; ++gg) c += cpart[gg * ne + tid];
      cvec[tid] = c;
    }
  }
  const int nt = ne * (ne + 1) / 2;
  double v[kInvPerThread];
  int er[kInvPerThread], ec[kInvPerThread];  // entry (row >= column); -1 past the triangle
#pragma unroll
  for (int m = 0; m < kInvPerThread; ++m) {
    const int e = tid + m * kPrepThreads;
    int i = -1, j = -1;
    if (e < nt) {
      i = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
      while (i * (i + 1) / 2 > e) --i;
      while ((i + 1) * (i + 2) / 2 <= e) ++i;
      j = e - i * (i + 1) / 2;
    }
    er[m] = i;
    ec[m] = j;
    v[m] = e < nt ? a.E[(size_t)i * ne + j] : 0.0;
  }
  bool ok = true;
  // One barrier per pivot: step k reads column k from colb[k & 1] and writes
  // column k + 1 of its result into the other buffer, which nobody reads
  // until the next step's barrier (its last readers were in step k - 1).
#pragma unroll
  for (int m = 0; m < kInvPerThread; ++m) {
    if (er[m] == 0) colb[0][ec[m]] = v[m];
    else if (ec[m] == 0) colb[0][er[m]] = v[m];
  }
  // Pivot k reads column k from colk and writes column k + 1 into next; the
  // loop takes two pivots per pass so that each buffer is a fixed LDS address.
  auto pivot = [&](int k, const double* colk, double* next) -> bool {
    const double p = colk[k];
    if (!(p > 0.0)) return false;  // the same value in every thread
    const double ip = 1.0 / p;
#pragma unroll
    for (int m = 0; m < kInvPerThread; ++m) {
      const int i = max(er[m], 0), j = max(ec[m], 0);
      const bool ri = er[m] == k, rj = ec[m] == k;
      double nv = fma(-colk[i] * ip, colk[j], v[m]);
      nv = (ri || rj) ? v[m] * ip : nv;
      v[m] = (ri && rj) ? -ip : nv;
      if (er[m] == k + 1) next[ec[m]] = v[m];       // (k + 1, j <= k + 1)
      else if (ec[m] == k + 1) next[er[m]] = v[m];  // (i > k + 1, k + 1)
    }
    return true;
  };
  for (int k = 0; k < ne; k += 2) {
    __syncthreads();
    if (!pivot(k, colb[0], colb[1])) {
      ok = false;
      break;
    }
    if (k + 1 == ne) break;
    __syncthreads();
    if (!pivot(k + 1, colb[1], colb[0])) {
      ok = false;
      break;
    }
  }
  __syncthreads();
  // E^-1 = -(swept matrix), both triangles.
#pragma unroll
  for (int m = 0; m < kInvPerThread; ++m)
    if (er[m] >= 0) {
      const double x = ok ? -v[m] : 0.0;
      Ms[er[m] * ne + ec[m]] = x;
      Ms[ec[m] * ne + er[m]] = x;
    }
  __syncthreads();
  const int n2 = ne * ne;
  for (int e = tid; e < n2; e += kPrepThreads) a.Einv[e] = Ms[e];
  if (tid == 0) a.st->defl_ok = ok ? 1 : 0;
  // y = E^-1 c: thread (row, part) sums a contiguous part of the columns.
  const int nparts = kPrepThreads / ne, row = tid % ne, part = tid / ne, len = (ne + nparts - 1) / nparts;
  double yv = 0.0;
  if (part < nparts)
    for (int c = part * len; c < min(ne, (part + 1) * len); ++c) yv += Ms[row * ne + c] * cvec[c];
  cpart[tid] = yv;
  __syncthreads();
  if (tid < ne) {
    double y = 0.0;
    for (int pp = 0; pp < nparts; ++pp) y += cpart[pp * ne + tid];
    a.y[tid] = y;
  }
}

// Operands of a group's partials that do not change during the solve, loaded
// into registers at the start of the per-group kernels (kGroupBlock threads)
// so that their latency overlaps the kernel's reductions: wave w holds the
// SW window columns w + 16 m, its lanes the group entries lane + 64 j; thread
// (row, part) of E^-1 holds the entries of its part of the window columns.
constexpr int kGroupWaves = kGroupBlock / 64;
constexpr int kSrCols = (kMaxCoarse + kGroupWaves - 1) / kGroupWaves;
constexpr int kSrRows = (kGroupThreads + 63) / 64;
constexpr int einv_per_thread() {
  int m = 0;
  for (int ne = 1; ne <= kMaxCoarse; ++ne) {
    const int parts = kGroupBlock / ne, len = (ne + parts - 1) / parts;
    m = len > m ? len : m;
  }
  return m;
}
constexpr int kEinvPer = einv_per_thread();

struct GroupLoads {
  double sr[kSrCols][kSrRows];
  double einv[kEinvPer];
};

// The E^-1 entries come from the per-group copy EinvW (addresses that do not
// depend on a prior load) or, with kFromEinv, from E^-1 itself.
template <bool kFromEinv>
__device__ inline void load_group(const PcgArgs& a, int g, GroupLoads& L) {
  const int ne = a.ne, tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
#pragma unroll
  for (int m = 0; m < kSrCols; ++m) {
    const int lc = wave + kGroupWaves * m;
#pragma unroll
    for (int j = 0; j < kSrRows; ++j) {
      const int t = lane + 64 * j;
      L.sr[m][j] = (ne && lc < a.gw && t < kGroupThreads) ? sr_col(a, g, lc)[t] : 0.0;
    }
  }
#pragma unroll
  for (int m = 0; m < kEinvPer; ++m) L.einv[m] = 0.0;
  if (ne == 0) return;
  const int nparts = kGroupBlock / ne, rrow = tid % ne, part = tid / ne, gw = a.gw;
  const int len = (gw + nparts - 1) / nparts;
  const int c0 = kFromEinv ? 6 * a.grp_lo[g] : 0;
  if (part < nparts) {
#pragma unroll
    for (int m = 0; m < kEinvPer; ++m) {
      const int lc = part * len + m;
      if (m < len && lc < gw) {
        if (kFromEinv) {
          const int gc = win_col(a, c0, lc);
          L.einv[m] = gc >= 0 ? a.Einv[(size_t)gc * ne + rrow] : 0.0;  // E^-1 is symmetric
        } else {
          L.einv[m] = a.EinvW[((size_t)g * gw + lc) * ne + rrow];
        }
      }
    }
  }
}

// Per-group partials of group g: out = {rz, rr, bb, E^-1 (SW)'z}; zloc holds
// the group's z (0 past the last row). (SW)'z is non-zero only on the group's
// window of columns, so E^-1 (SW)'z needs only those columns of E^-1 (window
// columns past the last coarse column have zero SW and E^-1 entries).
__device__ inline void group_partials(const PcgArgs& a, int g, const GroupLoads& L, const double* zloc, double rz,
                                      double rr, double bb, double* cs, double* mpart, double* out) {
  const int ne = a.ne, tid = threadIdx.x;
  if (tid == 0) {
    out[0] = rz;
    out[1] = rr;
    out[2] = bb;
  }
  if (ne == 0) return;
  const int wave = tid >> 6, lane = tid & 63, w = a.gw;
#pragma unroll
  for (int m = 0; m < kSrCols; ++m) {
    const int lc = wave + kGroupWaves * m;
    if (lc < w) {
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < kSrRows; ++j) {
        const int t = lane + 64 * j;
        if (t < kGroupThreads) s += L.sr[m][j] * zloc[t];
      }
      s = wave_sum(s);
      if (lane == 0) cs[lc] = s;
    }
  }
  __syncthreads();
  const int nparts = kGroupBlock / ne, part = tid / ne, len = (w + nparts - 1) / nparts;
  double m = 0.0;
  if (part < nparts) {
#pragma unroll
    for (int j = 0; j < kEinvPer; ++j) {
      const int lc = part * len + j;
      if (j < len && lc < w) m += L.einv[j] * cs[lc];
    }
  }
  mpart[tid] = m;
  __syncthreads();
  if (tid < ne) {
    double mu = 0.0;
    for (int pp = 0; pp < nparts; ++pp) mu += mpart[pp * ne + tid];
    out[3 + tid] = mu;
  }
}

// With a.reduce1: the work-group that writes the last group's partials sums
// every group's, in group order (the order k_pcg_spmv sums them in without
// reduce1), into part1 and re-arms the counter. Which work-group arrives last
// does not change the sum.
// The partials are written by the first two waves (thread 0 and threads
// below ne <= 106); only they release their stores at device scope (an L2
// write-back, since each XCD has its own L2), and only the last work-group
// acquires (an L2 invalidate): a full fence in every wave of every
// work-group would invalidate the L2 under the other work-groups' loads.
// The summing work-group stages whole groups' partials through LDS (red,
// kRedChunk doubles), every load of a pass in flight at once.
static_assert(3 + kMaxCoarse <= 128, "the group partials are written by the first two waves");
constexpr int kRedChunk = 4096;
constexpr int kRedPer = kRedChunk / kGroupBlock;

__device__ inline void sum_partials_if_last(const PcgArgs& a, bool* last, double* red) {
  const int tid = threadIdx.x, stride = 3 + a.ne;
  if (tid < 128) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (tid == 0) *last = atomicAdd(&a.st->arrive, 1u) == (uint32_t)a.groups - 1u;
  __syncthreads();
  if (!*last) return;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // every group's partials are visible here
  const int per = kRedChunk / stride;
  double acc = 0.0;
  for (int g0 = 0; g0 < a.groups; g0 += per) {
    const int ng = min(per, a.groups - g0), n = ng * stride;
    const double* src = a.part2 + (size_t)g0 * stride;
    double v[kRedPer];
#pragma unroll
    for (int m = 0; m < kRedPer; ++m) {
      const int q = tid + kGroupBlock * m;
      v[m] = q < n ? src[q] : 0.0;
    }
    __syncthreads();  // the previous pass is summed
#pragma unroll
    for (int m = 0; m < kRedPer; ++m) red[tid + kGroupBlock * m] = v[m];
    __syncthreads();
    if (tid < stride) {
#pragma unroll 8
      for (int gg = 0; gg < ng; ++gg) acc += red[gg * stride + tid];
    }
  }
  if (tid < stride) a.part1[tid] = acc;
  if (tid == 0) a.st->arrive = 0u;
}

// t_e = (L_i^-1 r_i)_comp for entry e = 6 i + comp; rloc holds the group's r.
__device__ inline double lower_solve_entry(const PcgArgs& a, int row, int comp, const double* rloc, int base) {
  const double* Li = a.linv + (size_t)36 * row + 6 * comp;
  double s = 0.0;
  for (int m = 0; m <= comp; ++m) s += Li[m] * rloc[base + m];
  return s;
}

// x0 = W y, r0 = b - (SW) y, p = 0 on one group of rows per work-group; with
// block-Jacobi z0 = M^-1 r0 and the group's partials, with the learned
// preconditioner t0 = L^-1 r0 (k_pcg_lower, k_pcg_upper and
// k_defl_partials follow).
template <bool kLearned>
__global__ __launch_bounds__(kGroupBlock) void k_defl_init(PcgArgs a) {
  __shared__ double rloc[kGroupThreads];
  __shared__ double zloc[kGroupThreads];
  __shared__ double lds[48];
  __shared__ double cs[kMaxCoarse];
  __shared__ double mpart[kGroupBlock];
  __shared__ double ys[kMaxCoarse];
  const int ne = a.ne, tid = threadIdx.x, P = a.Pvar, g = blockIdx.x;
  GroupLoads L;
  if (!kLearned) load_group<true>(a, g, L);
  for (int c = tid; c < ne; c += blockDim.x) ys[c] = a.y[c];
  if (ne) {
    // The group's window rows of E^-1 for k_pcg_update.
    const int c0 = 6 * a.grp_lo[g];
    for (int e = tid; e < a.gw * ne; e += blockDim.x) {
      const int lc = e / ne, col = e - lc * ne, gc = win_col(a, c0, lc);
      a.EinvW[(size_t)g * a.gw * ne + e] = gc >= 0 ? a.Einv[(size_t)gc * ne + col] : 0.0;
    }
  }
  __syncthreads();
  const int row = g * kRowsPerGroup + tid / 6, comp = tid % 6;
  const bool active = tid < kGroupThreads && row < P;
  double rv = 0.0, bv = 0.0;
  if (active) {
    const int e = 6 * row + comp;
    bv = a.b[e];
    double sy = 0.0, xv = 0.0;
    if (ne) {
      const int c0 = 6 * a.grp_lo[g];
      for (int lc = 0; lc < a.gw; ++lc) {
        const int gc = win_col(a, c0, lc);
        if (gc >= 0) sy += sr_col(a, g, lc)[tid] * ys[gc];
      }
      const int n0 = a.hat_node[row];
      const double w0 = a.hat_w[2 * row], w1 = a.hat_w[2 * row + 1];
      xv = w0 * ys[6 * n0 + comp] + (w1 != 0.0 ? w1 * ys[6 * (n0 + 1) + comp] : 0.0);
      for (int c = 0; c < a.nr; ++c) xv += a.R[(size_t)c * a.n6 + e] * ys[a.neh + c];
    }
    rv = bv - sy;
    a.x[e] = xv;
    a.r[e] = rv;
    a.p0[e] = 0.0;
  }
  if (tid < kGroupThreads) rloc[tid] = rv;
  __syncthreads();
  if (kLearned) {
    if (active) a.t[6 * row + comp] = lower_solve_entry(a, row, comp, rloc, tid - comp);
    return;
  }
  double zv = 0.0;
  if (active) {
    const double* M = a.Minv + (size_t)36 * row + 6 * comp;
    const int base = tid - comp;
    for (int c = 0; c < 6; ++c) zv += M[c] * rloc[base + c];
    a.z[6 * row + comp] = zv;
  }
  if (tid < kGroupThreads) zloc[tid] = zv;
  double rz = rv * zv, rr = rv * rv, bb = bv * bv;
  block_sum3(rz, rr, bb, lds);
  group_partials(a, g, L, zloc, rz, rr, bb, cs, mpart, a.part2 + (size_t)g * (3 + ne));
  __shared__ bool last;
  __shared__ double red[kRedChunk];
  if (a.reduce1) sum_partials_if_last(a, &last, red);
}

// Iteration k: stops when |r|^2 <= tol2 or k = max_iters, else
// q = S p_k with p_k = z - W mu + beta p_{k-1}, evaluated as
// S z + beta S p_{k-1} - (SW) mu so that the row's block products (the bulk
// of the loads) do not wait for the sums of the previous update's per-group
// partials (r'z, r'r and mu); the own row's p_k is stored and
// part[i] = p_k,i . q_i. Thread t < 252 handles component t % 6 of every 42nd
// block of the row and window columns t / 6 + 42 m of (SW)_i. kReduced reads
// the summed partials (part1) instead of summing every group's.
template <bool kReduced>
__global__ __launch_bounds__(256) void k_pcg_spmv(PcgArgs a, int32_t k, const double* __restrict__ p_prev,
                                                  double* __restrict__ p_cur) {
  __shared__ double chunk[kReduced ? 1 : kPartChunk];
  __shared__ double sums[3 + kMaxCoarse];
  __shared__ double ysz[kGroupThreads], ysp[kGroupThreads], ysw[kGroupThreads];
  PcgState* st = a.st;
  const int32_t done = st->done, max_iters = st->max_iters;
  const double rz_prev = st->rz[(k + 1) & 1], tol2_prev = st->tol2, tol_rel2 = st->tol_rel2;
  const int ne = a.ne, tid = threadIdx.x, stride = 3 + ne;
  const int i = blockIdx.x, comp = tid % 6, g = tid / 6;
  // Every read of the work-group is issued before its first wait (each
  // dependent round of global loads costs microseconds): the state, the first
  // chunk of group partials, the row's own z, p_{k-1} and hat weights, (SW)_i
  // and the block products.
  constexpr int kChunkPer = kReduced ? 1 : kPartChunk / 256;
  const int per_chunk = kPartChunk / stride, ng0 = min(per_chunk, a.groups);
  double pre[kChunkPer];
  if (kReduced) {
    pre[0] = tid < stride ? a.part1[tid] : 0.0;
  } else {
#pragma unroll
    for (int m = 0; m < kChunkPer; ++m) {
      const int q = tid + 256 * m;
      pre[m] = q < ng0 * stride ? a.part2[q] : 0.0;
    }
  }
  double z_own = 0.0, p_own = 0.0, hw0 = 0.0, hw1 = 0.0, r_own[kMaxRitz];
  int hn0 = 0;
#pragma unroll
  for (int c = 0; c < kMaxRitz; ++c) r_own[c] = 0.0;
  if (tid < 6) {
    z_own = a.z[6 * i + tid];
    p_own = p_prev[6 * i + tid];
    if (ne) {
      hn0 = a.hat_node[i];
      hw0 = a.hat_w[2 * i];
      hw1 = a.hat_w[2 * i + 1];
#pragma unroll
      for (int c = 0; c < kMaxRitz; ++c)
        if (c < a.nr) r_own[c] = a.R[(size_t)c * a.n6 + 6 * i + tid];
    }
  }
  constexpr int kSwCols = (kMaxCoarse + kRowsPerGroup - 1) / kRowsPerGroup;
  double sz = 0.0, sp = 0.0, sw[kSwCols] = {};
  const int gi = i / kRowsPerGroup, c0 = 6 * (ne ? a.grp_lo[gi] : 0);
  if (tid < kGroupThreads) {
    const int rb = a.row_ptr[i], re = a.row_ptr[i + 1];
#pragma unroll 4
    for (int kb = rb + g; kb < re; kb += kRowsPerGroup) {
      const int j = a.col_var[kb];
      const double* srow = a.S + (size_t)36 * kb + 6 * comp;
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        sz += srow[c] * a.z[6 * j + c];
        sp += srow[c] * p_prev[6 * j + c];
      }
    }
    if (ne) {
      const int t = 6 * (i - gi * kRowsPerGroup) + comp;
#pragma unroll
      for (int m = 0; m < kSwCols; ++m) {
        const int lc = g + m * kRowsPerGroup;
        if (lc < a.gw) sw[m] = sr_col(a, gi, lc)[t];
      }
    }
  }
  if (done) return;  // the same value in every thread of the work-group
  if (kReduced) {
    if (tid < stride) sums[tid] = pre[0];
  } else {
    // Fixed-order sum of the group partials, a chunk of whole groups per pass.
#pragma unroll
    for (int m = 0; m < kChunkPer; ++m) chunk[tid + 256 * m] = pre[m];
    __syncthreads();
    double acc = 0.0;
    if (tid < stride) {
#pragma unroll 8
      for (int gg = 0; gg < ng0; ++gg) acc += chunk[gg * stride + tid];
    }
    for (int g0 = ng0; g0 < a.groups; g0 += per_chunk) {
      const int ng = min(per_chunk, a.groups - g0);
      __syncthreads();
      for (int q = tid; q < ng * stride; q += blockDim.x) chunk[q] = a.part2[(size_t)g0 * stride + q];
      __syncthreads();
      if (tid < stride) {
#pragma unroll 8
        for (int gg = 0; gg < ng; ++gg) acc += chunk[gg * stride + tid];
      }
    }
    if (tid < stride) sums[tid] = acc;
  }
  __syncthreads();
  const double* mu_s = sums + 3;
  const double rz = sums[0], rr = sums[1];
  double tol2;
  bool stop;
  if (k == 0) {
    const double bb = sums[2];
    tol2 = tol_rel2 * bb;
    stop = bb <= 0.0 || max_iters <= 0 || rr <= tol2;
    if (blockIdx.x == 0 && tid == 0) {
      st->bb = bb;
      st->tol2 = tol2;
    }
  } else {
    tol2 = tol2_prev;
    stop = rr <= tol2 || k >= max_iters;
  }
  if (stop) {
    if (blockIdx.x == 0 && tid == 0) {
      st->iters = k;
      st->done = 1;
    }
    return;
  }
  const double beta = k > 0 ? rz / rz_prev : 0.0;
  if (blockIdx.x == 0 && tid == 0) {
    st->rz[k & 1] = rz;
    st->iters = k;
    if (k < a.zcap) a.coef[2 * k] = rz;
  }
  if (tid < kGroupThreads) {
    double w = 0.0;
    if (ne) {
#pragma unroll
      for (int m = 0; m < kSwCols; ++m) {
        const int lc = g + m * kRowsPerGroup, gc = lc < a.gw ? win_col(a, c0, lc) : -1;
        if (gc >= 0) w += sw[m] * mu_s[gc];
      }
    }
    ysz[tid] = sz;
    ysp[tid] = sp;
    ysw[tid] = w;
  }
  __syncthreads();
  if (tid < 6) {
    double qz = 0.0, qp = 0.0, qw = 0.0;
#pragma unroll 6
    for (int gg = 0; gg < kRowsPerGroup; ++gg) {
      qz += ysz[6 * gg + tid];
      qp += ysp[6 * gg + tid];
      qw += ysw[6 * gg + tid];
    }
    const double qa = qz + beta * qp - qw;
    double zh = z_own;  // deflated preconditioned residual z - W mu
    if (ne) {
      double wm = hw0 * mu_s[6 * hn0 + tid] + (hw1 != 0.0 ? hw1 * mu_s[6 * (hn0 + 1) + tid] : 0.0);
#pragma unroll
      for (int c = 0; c < kMaxRitz; ++c)
        if (c < a.nr) wm += r_own[c] * mu_s[a.neh + c];
      zh -= wm;
    }
    const double pa = zh + beta * p_own;
    a.q[6 * i + tid] = qa;
    p_cur[6 * i + tid] = pa;
    if (k < a.zcap) a.Zh[(size_t)k * a.n6 + 6 * i + tid] = zh;
    ysz[tid] = pa * qa;
  }
  __syncthreads();
  if (tid == 0) a.part[i] = ((ysz[0] + ysz[1]) + (ysz[2] + ysz[3])) + (ysz[4] + ysz[5]);
}

// alpha = r'z / p'q; x += alpha p; r -= alpha q on one group of rows per
// work-group; with block-Jacobi z = M^-1 r and the group's partials r'z, r'r,
// E^-1 (SW)'z, with the learned preconditioner t = L^-1 r (k_pcg_lower,
// k_pcg_upper and k_defl_partials follow).
template <bool kLearned>
__global__ __launch_bounds__(kGroupBlock) void k_pcg_update(PcgArgs a, int32_t k, const double* __restrict__ p_cur) {
  __shared__ double rloc[kGroupThreads];
  __shared__ double zloc[kGroupThreads];
  __shared__ double lds[48];
  __shared__ double cs[kMaxCoarse];
  __shared__ double mpart[kGroupBlock];
  const PcgState* st = a.st;
  const int32_t done = st->done;
  const double rz = st->rz[k & 1];
  const int ne = a.ne, tid = threadIdx.x, P = a.Pvar, g = blockIdx.x;
  const int row = g * kRowsPerGroup + tid / 6, comp = tid % 6;
  const bool active = tid < kGroupThreads && row < P;
  GroupLoads L;
  if (!kLearned) load_group<false>(a, g, L);
  // The group's vectors and preconditioner rows (M^-1 or L^-1), loaded
  // before the reduction.
  double xv = 0.0, pv = 0.0, rv = 0.0, qv = 0.0, M[6] = {};
  if (active) {
    const int e = 6 * row + comp;
    xv = a.x[e];
    pv = p_cur[e];
    rv = a.r[e];
    qv = a.q[e];
    const double* Mr = (kLearned ? a.linv : a.Minv) + (size_t)36 * row + 6 * comp;
#pragma unroll
    for (int c = 0; c < 6; ++c) M[c] = (!kLearned || c <= comp) ? Mr[c] : 0.0;
  }
  // p'q, summed by every group in the same order.
  double pq = 0.0, unused0 = 0.0, unused1 = 0.0;
  for (int r = tid; r < P; r += blockDim.x) pq += a.part[r];
  if (done) return;  // the same value in every thread
  block_sum3(pq, unused0, unused1, lds);
  const double alpha = rz / pq;
  if (g == 0 && tid == 0 && k < a.zcap) a.coef[2 * k + 1] = alpha;
  if (active) {
    const int e = 6 * row + comp;
    a.x[e] = xv + alpha * pv;
    rv -= alpha * qv;
    a.r[e] = rv;
  }
  if (tid < kGroupThreads) rloc[tid] = rv;
  __syncthreads();
  double zv = 0.0;
  if (active) {
    const int base = tid - comp;
#pragma unroll
    for (int c = 0; c < 6; ++c) zv += M[c] * rloc[base + c];
    (kLearned ? a.t : a.z)[6 * row + comp] = zv;
  }
  if (kLearned) return;
  if (tid < kGroupThreads) zloc[tid] = zv;
  double rzg = rv * zv, rr = rv * rv, zero = 0.0;
  block_sum3(rzg, rr, zero, lds);
  group_partials(a, g, L, zloc, rzg, rr, 0.0, cs, mpart, a.part2 + (size_t)g * (3 + ne));
  __shared__ bool last;
  __shared__ double red[kRedChunk];
  if (a.reduce1) sum_partials_if_last(a, &last, red);
}

// Learned preconditioner, after t = L^-1 r: u_i = t_i - sum_{j < i} K_ij t_j,
// one wave per row (lanes take every 64th lower block, then a fixed-order
// wave sum).
__global__ __launch_bounds__(64) void k_pcg_lower(PcgArgs a) {
  if (a.st->done) return;
  const int i = blockIdx.x, lane = threadIdx.x;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int pos = a.row_ptr[i] + lane; pos < a.diag_pos[i]; pos += 64) {
    const int j = a.col_var[pos];
    const float* K = a.kmat + (size_t)36 * pos;
    const double* tj = a.t + 6 * (size_t)j;
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      double s = 0.0;
#pragma unroll
      for (int c = 0; c < 6; ++c) s += (double)K[6 * r + c] * tj[c];
      acc[r] += s;
    }
  }
#pragma unroll
  for (int r = 0; r < 6; ++r) acc[r] = wave_sum(acc[r]);
  if (lane < 6) {
    double v = acc[0];
#pragma unroll
    for (int r = 1; r < 6; ++r)
      if (lane == r) v = acc[r];
    a.u[6 * i + lane] = a.t[6 * i + lane] - v;
  }
}

// w_j = u_j - sum_{i > j} K_ij' u_i, z_j = L_j^-T w_j; one wave per row (K_ij
// is stored in row i, at the transposed position of block (j, i)).
__global__ __launch_bounds__(64) void k_pcg_upper(PcgArgs a) {
  __shared__ double w[6];
  if (a.st->done) return;
  const int j = blockIdx.x, lane = threadIdx.x;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int pos = a.diag_pos[j] + 1 + lane; pos < a.row_ptr[j + 1]; pos += 64) {
    const int i = a.col_var[pos];
    const float* K = a.kmat + (size_t)36 * a.tpos[pos];
    const double* ui = a.u + 6 * (size_t)i;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      double s = 0.0;
#pragma unroll
      for (int r = 0; r < 6; ++r) s += (double)K[6 * r + c] * ui[r];
      acc[c] += s;
    }
  }
#pragma unroll
  for (int c = 0; c < 6; ++c) acc[c] = wave_sum(acc[c]);
  if (lane < 6) {
    double v = acc[0];
#pragma unroll
    for (int c = 1; c < 6; ++c)
      if (lane == c) v = acc[c];
    w[lane] = a.u[6 * j + lane] - v;
  }
  __syncthreads();
  if (lane < 6) {
    const double* Lj = a.linv + (size_t)36 * j;
    double s = 0.0;
    for (int m = lane; m < 6; ++m) s += Lj[6 * m + lane] * w[m];
    a.z[6 * j + lane] = s;
  }
}

// Learned preconditioner: the group's partials r'z, r'r (b'b with
// `with_b`) and E^-1 (SW)'z from r and z, one group of rows per work-group,
// as k_pcg_update and k_defl_init form them with block-Jacobi.
__global__ __launch_bounds__(kGroupBlock) void k_defl_partials(PcgArgs a, int32_t with_b) {
  __shared__ double zloc[kGroupThreads];
  __shared__ double lds[48];
  __shared__ double cs[kMaxCoarse];
  __shared__ double mpart[kGroupBlock];
  const int32_t done = a.st->done;
  const int ne = a.ne, tid = threadIdx.x, P = a.Pvar, g = blockIdx.x;
  const int row = g * kRowsPerGroup + tid / 6, comp = tid % 6;
  const bool active = tid < kGroupThreads && row < P;
  GroupLoads L;
  load_group<false>(a, g, L);
  double rv = 0.0, zv = 0.0, bv = 0.0;
  if (active) {
    const int e = 6 * row + comp;
    rv = a.r[e];
    zv = a.z[e];
    if (with_b) bv = a.b[e];
  }
  if (done) return;  // the same value in every thread
  if (tid < kGroupThreads) zloc[tid] = zv;
  double rz = rv * zv, rr = rv * rv, bb = bv * bv;
  block_sum3(rz, rr, bb, lds);
  group_partials(a, g, L, zloc, rz, rr, bb, cs, mpart, a.part2 + (size_t)g * (3 + ne));
  __shared__ bool last;
  __shared__ double red[kRedChunk];
  if (a.reduce1) sum_partials_if_last(a, &last, red);
}

// Clears the convergence flag and lifts the iteration limit so that one more
// iteration runs (kernel timing only).
__global__ void k_pcg_reopen(PcgState* st) {
  st->done = 0;
  st->max_iters = st->iters + 2;
  st->tol2 = -1.0;
}

// ------------------------------------------------------------ recycling
// After a solve, Ritz vectors of its Lanczos process (Saad, Yeung, Erhel,
// Guyomarc'h 2000, sec. 4; Stathopoulos & Orginos, "Computing and deflating
// eigenvalues while solving multiple right-hand side linear systems", SIAM
// J. Sci. Comput. 32, 2010): the stored deflated preconditioned residuals
// z_k - W mu_k, scaled by 1 / sqrt(r_k'z_k), are the Lanczos vectors of the
// deflated preconditioned operator, and the CG coefficients give its
// tridiagonal matrix. The host takes the eigenvectors of that tridiagonal
// for its smallest eigenvalues (Y), k_ritz_form forms Rn = Zh' Y, and a
// Rayleigh-Ritz step on span[R, Rn] with respect to (S, M) (k_row_dense for S Rn,
// k_ritz_gram, a small generalized eigenproblem on the host, k_ritz_merge)
// keeps the best r vectors for the next solve.

constexpr int kRitzFormThreads = 256;
constexpr int kMaxLanczos = 64;

// Rn[c] = sum_j Zh[j] Y[j][c], one thread per (entry, column).
__global__ __launch_bounds__(kRitzFormThreads) void k_ritz_form(int32_t n6, int32_t m, int32_t r,
                                                                const double* __restrict__ Zh,
                                                                const double* __restrict__ Y, double* __restrict__ Rn) {
  __shared__ double Ys[kMaxLanczos];
  const int c = blockIdx.y;
  for (int t = threadIdx.x; t < m; t += blockDim.x) Ys[t] = Y[t * r + c];
  __syncthreads();
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n6) return;
  double acc = 0.0;
#pragma unroll 8
  for (int j = 0; j < m; ++j) acc += Zh[(size_t)j * n6 + e] * Ys[j];
  Rn[(size_t)c * n6 + e] = acc;
}

// Gram matrices of Y = [R | Rn], one work-group per entry (a, b):
// A_ab = Y_a'(S Y_b) (S R from the group windows of this solve's SW),
// B_ab = Y_a'(M Y_b) with M the 6x6 diagonal blocks of S.
__global__ __launch_bounds__(256) void k_ritz_gram(PcgArgs a, int32_t r, const double* __restrict__ Rn,
                                                   const double* __restrict__ SRn, const int32_t* __restrict__ diag_pos,
                                                   double* __restrict__ AB) {
  __shared__ double lds[4];
  const int ia = blockIdx.x, ib = blockIdx.y, nb = a.nr + r;
  const double* Ya = ia < a.nr ? a.R + (size_t)ia * a.n6 : Rn + (size_t)(ia - a.nr) * a.n6;
  const double* Yb = ib < a.nr ? a.R + (size_t)ib * a.n6 : Rn + (size_t)(ib - a.nr) * a.n6;
  double sa = 0.0, sb = 0.0;
  for (int e = threadIdx.x; e < a.n6; e += blockDim.x) {
    const double sy = ib < a.nr ? sw_entry(a, e, a.neh + ib) : SRn[(size_t)(ib - a.nr) * a.n6 + e];
    const int i = e / 6, comp = e % 6;
    const double* Mi = a.S + (size_t)36 * diag_pos[i] + 6 * comp;
    const double* yb = Yb + 6 * i;
    const double my = ((Mi[0] * yb[0] + Mi[1] * yb[1]) + (Mi[2] * yb[2] + Mi[3] * yb[3])) + (Mi[4] * yb[4] + Mi[5] * yb[5]);
    sa += Ya[e] * sy;
    sb += Ya[e] * my;
  }
  sa = block_sum256(sa, lds);
  __syncthreads();
  sb = block_sum256(sb, lds);
  if (threadIdx.x == 0) {
    AB[(size_t)ia * nb + ib] = sa;
    AB[(size_t)nb * nb + (size_t)ia * nb + ib] = sb;
  }
}

// Rnext[c] = sum_a R[a] Cm[a][c] + sum_a Rn[a] Cm[nr_old + a][c], one
// thread per (entry, column).
__global__ __launch_bounds__(256) void k_ritz_merge(int32_t n6, const double* __restrict__ R, int32_t nr_old,
                                                    const double* __restrict__ Rn, int32_t r,
                                                    const double* __restrict__ Cm, int32_t r_next,
                                                    double* __restrict__ Rnext) {
  __shared__ double Cs[2 * kMaxRitz];
  const int nb = nr_old + r, c = blockIdx.y;
  for (int t = threadIdx.x; t < nb; t += blockDim.x) Cs[t] = Cm[t * r_next + c];
  __syncthreads();
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n6) return;
  double acc = 0.0;
  for (int ia = 0; ia < nb; ++ia)
    acc += (ia < nr_old ? R[(size_t)ia * n6 + e] : Rn[(size_t)(ia - nr_old) * n6 + e]) * Cs[ia];
  Rnext[(size_t)c * n6 + e] = acc;
}

// ------------------------------------------------------------ launchers

size_t pcg_state_bytes() { return sizeof(PcgState); }
int32_t pcg_max_hat_nodes() { return kMaxHatNodes; }
int32_t pcg_max_ritz() { return kMaxRitz; }
int32_t pcg_max_lanczos() { return kMaxLanczos; }
int32_t pcg_group_rows() { return kRowsPerGroup; }
int32_t pcg_update_groups(int32_t Pvar) { return (Pvar + kRowsPerGroup - 1) / kRowsPerGroup; }
bool pcg_deflation_applied(const void* state_host) {
  return static_cast<const PcgState*>(state_host)->defl_ok != 0;
}

static PcgArgs make_args(const PcgBuffers& B) {
  PcgArgs a;
  a.Pvar = B.Pvar;
  a.K = B.K;
  a.neh = 6 * B.K;
  a.nr = B.K > 0 ? B.nr : 0;
  a.ne = a.neh + a.nr;
  a.groups = pcg_update_groups(B.Pvar);
  a.gwh = B.gwh;
  a.gw = B.gwh + a.nr;
  a.n6 = 6 * B.Pvar;
  a.zcap = std::min(B.zcap, kMaxLanczos);
  a.pad2 = a.pad3 = 0;
  a.row_ptr = B.row_ptr; a.col_var = B.col_var; a.S = B.S; a.Minv = B.Minv; a.b = B.b;
  a.hat_node = B.hat_node; a.hat_w = B.hat_w; a.node_range = B.node_range; a.grp_lo = B.grp_lo; a.R = B.R;
  a.blk_hat = B.blk_hat; a.blk_w = B.blk_w; a.blk_rng = B.blk_rng;
  a.SR = B.SR; a.E = B.E; a.Einv = B.Einv; a.EinvW = B.Einv + (size_t)a.ne * a.ne;
  a.x = B.x; a.r = B.r; a.z = B.z; a.p0 = B.p0; a.p1 = B.p1; a.q = B.q;
  a.part = B.part; a.part2 = B.part2; a.y = B.y; a.cden = B.cden; a.Zh = B.Zh; a.coef = B.coef;
  // part2 holds groups + 1 rows of partials; the last one is part1.
  a.part1 = B.part2 + (size_t)a.groups * (3 + a.ne);
  a.reduce1 = B.partials == 0 ? (a.groups > kPartChunk / (3 + a.ne) ? 1 : 0) : (B.partials == 2 ? 1 : 0);
  a.pad4 = 0;
  a.st = static_cast<PcgState*>(B.state_dev);
  a.linv = B.linv; a.kmat = B.kmat; a.diag_pos = B.diag_pos; a.tpos = B.tpos; a.t = B.t; a.u = B.u;
  return a;
}

// With the learned preconditioner: z = L^-T (I - K)' (I - K) t after t = L^-1 r,
// then the group partials.
static void launch_learned_z(hipStream_t s, const PcgArgs& a, int32_t with_b) {
  k_pcg_lower<<<a.Pvar, 64, 0, s>>>(a);
  k_pcg_upper<<<a.Pvar, 64, 0, s>>>(a);
  k_defl_partials<<<a.groups, kGroupBlock, 0, s>>>(a, with_b);
}

static void launch_update(hipStream_t s, const PcgArgs& a, int32_t k, const double* p_cur) {
  if (a.kmat) {
    k_pcg_update<true><<<a.groups, kGroupBlock, 0, s>>>(a, k, p_cur);
    launch_learned_z(s, a, 0);
  } else {
    k_pcg_update<false><<<a.groups, kGroupBlock, 0, s>>>(a, k, p_cur);
  }
}

static void launch_spmv(hipStream_t s, const PcgArgs& a, int32_t k, const double* p_prev, double* p_cur) {
  if (a.reduce1)
    k_pcg_spmv<true><<<a.Pvar, 256, 0, s>>>(a, k, p_prev, p_cur);
  else
    k_pcg_spmv<false><<<a.Pvar, 256, 0, s>>>(a, k, p_prev, p_cur);
}

// Deflated PCG solve of S y = b. The host polls the state after `first`
// iterations, then every `poll` iterations. With `probe` (three events),
// k_pcg_spmv and then k_pcg_update of one more iteration are each launched
// `probe_reps` times back to back after the solve and timed as two batches
// (k_pcg_spmv is idempotent; x, r and z then hold extra steps). Returns the
// iteration count or -1 on a launch error.
int launch_pcg_deflated(hipStream_t s, const PcgBuffers& B, double tol, int32_t max_iters, int32_t first,
                        int32_t poll, const hipEvent_t* probe, int32_t probe_reps) {
  const int32_t Pvar = B.Pvar;
  if (Pvar <= 0) return 0;
  if (B.K > kMaxHatNodes || B.nr > kMaxRitz || B.nr < 0) return -1;
  const bool learned = B.kmat != nullptr;
  if (learned && (B.nr > 0 || B.zcap > 0 || probe || !B.linv || !B.diag_pos || !B.tpos || !B.t || !B.u)) return -1;
  const PcgArgs a = make_args(B);
  double* p0 = B.p0;
  double* p1 = B.p1;
  const int groups = a.groups;
  if (a.ne > 0) {
    k_defl_sr<<<Pvar, 256, 0, s>>>(a);
    if (a.nr > 0)
      k_row_dense<true><<<Pvar, kDenseThreads, 0, s>>>(a, a.row_ptr, a.col_var, a.S, a.R, a.nr, a.n6, nullptr);
    k_defl_e<<<dim3(a.K, a.ne), 64, 0, s>>>(a);
    if (a.nr > 0) k_defl_dense<<<dim3(a.nr, a.nr + 1), 256, 0, s>>>(a);
  }
  k_defl_prep<<<1, kPrepThreads, 0, s>>>(a, tol, max_iters);
  if (learned) {
    k_defl_init<true><<<groups, kGroupBlock, 0, s>>>(a);
    launch_learned_z(s, a, 1);
  } else {
    k_defl_init<false><<<groups, kGroupBlock, 0, s>>>(a);
  }
  PcgState* hs = static_cast<PcgState*>(B.state_host);
  int k = 0;
  while (true) {
    const int n = std::min(k == 0 ? std::max(first, 1) : poll, max_iters - k);
    for (int j = 0; j < n; ++j, ++k) {
      launch_spmv(s, a, k, (k & 1) ? p1 : p0, (k & 1) ? p0 : p1);
      launch_update(s, a, k, (k & 1) ? p0 : p1);
    }
    // At k = max_iters this launch only records the final state.
    if (k >= max_iters) launch_spmv(s, a, k, p0, p1);
    if (hipMemcpyAsync(hs, a.st, sizeof(PcgState), hipMemcpyDeviceToHost, s) != hipSuccess) return -1;
    if (hipStreamSynchronize(s) != hipSuccess) return -1;
    if (hs->done || k >= max_iters) break;
  }
  if (hipGetLastError() != hipSuccess) return -1;
  const int iters = hs->iters;
  if (probe && iters > 0) {
    // Re-open the state for one timed iteration k = iters.
    k_pcg_reopen<<<1, 1, 0, s>>>(a.st);
    const int kk = iters;
    if (hipEventRecord(probe[0], s) != hipSuccess) return -1;
    for (int j = 0; j < probe_reps; ++j) launch_spmv(s, a, kk, (kk & 1) ? p1 : p0, (kk & 1) ? p0 : p1);
    if (hipEventRecord(probe[1], s) != hipSuccess) return -1;
    for (int j = 0; j < probe_reps; ++j) k_pcg_update<false><<<groups, kGroupBlock, 0, s>>>(a, kk, (kk & 1) ? p0 : p1);
    if (hipEventRecord(probe[2], s) != hipSuccess) return -1;
    if (hipGetLastError() != hipSuccess) return -1;
  }
  return iters;
}

void launch_ritz_form(hipStream_t s, int32_t Pvar, int32_t m, int32_t r, const double* Zh, const double* Y,
                      double* Rn) {
  const int n6 = 6 * Pvar;
  if (n6 <= 0 || m <= 0 || m > kMaxLanczos || r <= 0 || r > kMaxRitz) return;
  k_ritz_form<<<dim3((n6 + kRitzFormThreads - 1) / kRitzFormThreads, r), kRitzFormThreads, 0, s>>>(n6, m, r, Zh, Y,
                                                                                                 Rn);
}

void launch_ritz_spmv(hipStream_t s, int32_t Pvar, int32_t r, const int32_t* row_ptr, const int32_t* col_var,
                      const double* S, const double* Rn, double* SRn) {
  if (Pvar <= 0 || r <= 0 || r > kMaxRitz) return;
  k_row_dense<false><<<Pvar, kDenseThreads, 0, s>>>(PcgArgs{}, row_ptr, col_var, S, Rn, r, 6 * Pvar, SRn);
}

void launch_ritz_gram(hipStream_t s, const PcgBuffers& B, int32_t r, const double* Rn, const double* SRn,
                      const int32_t* diag_pos, double* AB) {
  const PcgArgs a = make_args(B);
  const int nb = a.nr + r;
  if (B.Pvar <= 0 || r <= 0 || r > kMaxRitz || a.K <= 0) return;
  k_ritz_gram<<<dim3(nb, nb), 256, 0, s>>>(a, r, Rn, SRn, diag_pos, AB);
}

void launch_ritz_merge(hipStream_t s, int32_t Pvar, const double* R, int32_t nr_old, const double* Rn, int32_t r,
                       const double* Cm, int32_t r_next, double* Rnext) {
  const int n6 = 6 * Pvar;
  if (n6 <= 0 || r_next <= 0 || r_next > kMaxRitz || nr_old + r > 2 * kMaxRitz) return;
  k_ritz_merge<<<dim3((n6 + 255) / 256, r_next), 256, 0, s>>>(n6, R, nr_old, Rn, r, Cm, r_next, Rnext);
}

}  // namespace mlba
