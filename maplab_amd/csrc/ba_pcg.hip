// ba_pcg.hip -- deflated, block-Jacobi preconditioned conjugate gradients on
// the reduced camera system S y = b (FP64).
//
// Block-Jacobi (inverse 6x6 diagonal blocks of S, Ceres' SCHUR_JACOBI) leaves
// a handful of very small eigenvalues on long trajectories: smooth, global
// modes of the vertex chain. They are removed by deflation (Saad, Yeung,
// Erhel, Guyomarc'h, "A deflated version of the conjugate gradient
// algorithm", SIAM J. Sci. Comput. 21, 2000) with a coarse space W of
// piecewise-linear "hat" functions over the vertex order (K nodes, 6 K
// columns): x0 = W E^-1 W'b, E = W'SW, and every search direction is made
// S-orthogonal to W with mu = E^-1 (SW)'z. The iteration converges to the
// same solution as plain PCG (the tolerance is on the true residual
// recursion), in 3-4x fewer iterations on long random-walk maps.
//
// Rows of S are handled in groups of 42 vertices (252 entries); SW is stored
// per group, column-major over the group's window of coarse columns, so that
// (SW)'z of a group is a set of coalesced column dot products.
//
// Per LM step: k_defl_sr (SW), k_defl_e (E), k_defl_prep (state, E^-1 by
// Gauss-Jordan, y = E^-1 W'b), k_defl_init (x0, r0, z0 and the first
// per-group partials). Per PCG iteration: k_pcg_spmv (sums the previous
// per-group partials; p = z - W mu + beta p_prev and
// q = S z + beta S p_prev - (SW) mu) and k_pcg_update (x, r, z and the
// per-group partials r'z, r'r, E^-1 (SW)'z). Every work-group sums partials in the same fixed order: the
// iteration is deterministic and uses no atomics.
#include "ba_device.h"
#include "ba_reduce.h"

namespace mlba {

constexpr int kMaxCoarse = 90;   // 6 x 15 hat nodes
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
  int32_t pad;
};

struct PcgArgs {
  int32_t Pvar, ne;            // ne = 6 K coarse columns (0: no deflation)
  int32_t K, groups;
  const int32_t* row_ptr;
  const int32_t* col_var;
  const double* S;
  const double* Minv;
  const double* b;
  const int32_t* hat_node;     // Pvar: first hat node of the vertex (0 without deflation)
  const double* hat_w;         // 2 Pvar: weights of hat_node and hat_node + 1 (0 without deflation)
  const int32_t* node_range;   // 2 K: first and last vertex of each hat's support
  const int32_t* grp_lo;       // groups: first hat node of the group's window of SW columns
  int32_t gw, pad2;            // window width in columns (6 x nodes)
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
  double* y;                   // ne: E^-1 W'b
  PcgState* st;
};

__device__ inline double hat_weight(const PcgArgs& a, int i, int k) {
  const int n0 = a.hat_node[i];
  return n0 == k ? a.hat_w[2 * i] : (n0 + 1 == k ? a.hat_w[2 * i + 1] : 0.0);
}

__device__ inline double* sr_col(const PcgArgs& a, int g, int lc) {
  return a.SR + ((size_t)g * a.gw + lc) * kGroupThreads;
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

// SW, one block row of S per work-group: (SW)_i = sum_j S_ij W_j with
// W_j = w0 e_n0 + w1 e_n0+1, written into the group window.
__global__ __launch_bounds__(256) void k_defl_sr(PcgArgs a) {
  __shared__ double acc[6 * kMaxCoarse];
  const int i = blockIdx.x, ne = a.ne;
  for (int k = threadIdx.x; k < 6 * ne; k += blockDim.x) acc[k] = 0.0;
  __syncthreads();
  const int rb = a.row_ptr[i], re = a.row_ptr[i + 1];
  for (int e = rb * 36 + threadIdx.x; e < re * 36; e += blockDim.x) {
    const int k = e / 36, rem = e % 36, row = rem / 6, c = rem % 6;
    const int j = a.col_var[k];
    const int n0 = a.hat_node[j];
    const double v = a.S[e];
    atomicAdd(&acc[row * ne + 6 * n0 + c], a.hat_w[2 * j] * v);
    const double w1 = a.hat_w[2 * j + 1];
    if (w1 != 0.0) atomicAdd(&acc[row * ne + 6 * (n0 + 1) + c], w1 * v);
  }
  __syncthreads();
  const int g = i / kRowsPerGroup, t0 = 6 * (i - g * kRowsPerGroup), c0 = 6 * a.grp_lo[g];
  for (int k = threadIdx.x; k < 6 * a.gw; k += blockDim.x) {
    const int comp = k % 6, lc = k / 6, gc = c0 + lc;
    sr_col(a, g, lc)[t0 + comp] = gc < ne ? acc[comp * ne + gc] : 0.0;
  }
}

// Entry (e, col) of SW from the group windows.
__device__ inline double sw_entry(const PcgArgs& a, int e, int col) {
  const int g = e / kGroupThreads;
  const int lc = col - 6 * a.grp_lo[g];
  return (lc >= 0 && lc < a.gw) ? sr_col(a, g, lc)[e - g * kGroupThreads] : 0.0;
}

// E = W'(SW), one entry per thread.
__global__ void k_defl_e(PcgArgs a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int ne = a.ne;
  if (t >= ne * ne) return;
  const int row = t / ne, col = t % ne, k = row / 6, comp = row % 6;
  double s = 0.0;
  for (int i = a.node_range[2 * k]; i <= a.node_range[2 * k + 1]; ++i)
    s += hat_weight(a, i, k) * sw_entry(a, 6 * i + comp, col);
  a.E[t] = s;
}

constexpr int kPrepThreads = 1024;
constexpr int kInvPerThread = (kMaxCoarse * kMaxCoarse + kPrepThreads - 1) / kPrepThreads;

// Solve set-up (one work-group): the PCG state, c = W'b, E^-1 by in-place
// Gauss-Jordan elimination without pivoting (E is SPD; each thread keeps its
// entries in registers, the pivot row and column go through LDS), and
// y = E^-1 c. A non-positive pivot disables deflation for this solve
// (E^-1 = 0, y = 0).
__global__ __launch_bounds__(kPrepThreads) void k_defl_prep(PcgArgs a, double tol, int32_t max_iters) {
  __shared__ double Ms[kMaxCoarse * kMaxCoarse];
  __shared__ double cpart[kPrepThreads];
  __shared__ double cvec[kMaxCoarse];
  __shared__ double rowk[kMaxCoarse];
  __shared__ double colk[kMaxCoarse];
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
  }
  if (ne == 0) return;
  // c = W'b: thread (column, g) sums the support rows lo + g, lo + g + G, ...
  {
    const int G = kPrepThreads / ne, col = tid % ne, g = tid / ne, k = col / 6, comp = col % 6;
    double acc = 0.0;
    if (g < G)
      for (int i = a.node_range[2 * k] + g; i <= a.node_range[2 * k + 1]; i += G)
        acc += hat_weight(a, i, k) * a.b[6 * i + comp];
    cpart[tid] = acc;
    __syncthreads();
    if (tid < ne) {
      double c = 0.0;
      for (int gg = 0; gg < G; ++gg) c += cpart[gg * ne + tid];
      cvec[tid] = c;
    }
  }
  // In-place Gauss-Jordan: for pivot k, a_rc -= a_rk a_kc / p, a_kc /= p,
  // a_rk = -a_rk / p, a_kk = 1 / p.
  const int n2 = ne * ne;
  double v[kInvPerThread];
  int er[kInvPerThread], ec[kInvPerThread];  // entry (row, column); -1 past the matrix
#pragma unroll
  for (int m = 0; m < kInvPerThread; ++m) {
    const int e = tid + m * kPrepThreads;
    er[m] = e < n2 ? e / ne : -1;
    ec[m] = e < n2 ? e - (e / ne) * ne : -1;
    v[m] = e < n2 ? (ec[m] <= er[m] ? a.E[er[m] * ne + ec[m]] : a.E[ec[m] * ne + er[m]]) : 0.0;
  }
  bool ok = true;
  for (int k = 0; k < ne; ++k) {
    __syncthreads();
#pragma unroll
    for (int m = 0; m < kInvPerThread; ++m) {
      if (er[m] == k) rowk[ec[m]] = v[m];
      if (ec[m] == k) colk[er[m]] = v[m];
    }
    __syncthreads();
    const double p = rowk[k];
    if (!(p > 0.0)) {
      ok = false;
      break;
    }
    const double ip = 1.0 / p;
    // Branch-free update of the owned entries.
#pragma unroll
    for (int m = 0; m < kInvPerThread; ++m) {
      const bool isr = er[m] == k, isc = ec[m] == k;
      const double ck = colk[max(er[m], 0)], rk = rowk[max(ec[m], 0)];
      const double vs = v[m] * ip;
      double nv = fma(-ck * ip, rk, v[m]);
      nv = isc ? -vs : nv;
      nv = isr ? vs : nv;
      v[m] = (isr && isc) ? ip : nv;
    }
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < kInvPerThread; ++m)
    if (er[m] >= 0) Ms[er[m] * ne + ec[m]] = ok ? v[m] : 0.0;
  __syncthreads();
  for (int e = tid; e < n2; e += kPrepThreads) a.Einv[e] = Ms[e];
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
  for (int ne = 6; ne <= kMaxCoarse; ne += 6) {
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
        if (kFromEinv)
          L.einv[m] = c0 + lc < ne ? a.Einv[(size_t)(c0 + lc) * ne + rrow] : 0.0;  // E^-1 is symmetric
        else
          L.einv[m] = a.EinvW[((size_t)g * gw + lc) * ne + rrow];
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

// x0 = W y, r0 = b - (SW) y, z0 = M^-1 r0, p = 0 on one group of rows per
// work-group, and the group's partials.
__global__ __launch_bounds__(kGroupBlock) void k_defl_init(PcgArgs a) {
  __shared__ double rloc[kGroupThreads];
  __shared__ double zloc[kGroupThreads];
  __shared__ double lds[48];
  __shared__ double cs[kMaxCoarse];
  __shared__ double mpart[kGroupBlock];
  __shared__ double ys[kMaxCoarse];
  const int ne = a.ne, tid = threadIdx.x, P = a.Pvar, g = blockIdx.x;
  GroupLoads L;
  load_group<true>(a, g, L);
  for (int c = tid; c < ne; c += blockDim.x) ys[c] = a.y[c];
  if (ne) {
    // The group's window rows of E^-1 for k_pcg_update.
    const int c0 = 6 * a.grp_lo[g];
    for (int e = tid; e < a.gw * ne; e += blockDim.x) {
      const int lc = e / ne, col = e - lc * ne;
      a.EinvW[(size_t)g * a.gw * ne + e] = c0 + lc < ne ? a.Einv[(size_t)(c0 + lc) * ne + col] : 0.0;
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
      for (int lc = 0; lc < a.gw && c0 + lc < ne; ++lc) sy += sr_col(a, g, lc)[tid] * ys[c0 + lc];
      const int n0 = a.hat_node[row];
      const double w0 = a.hat_w[2 * row], w1 = a.hat_w[2 * row + 1];
      xv = w0 * ys[6 * n0 + comp] + (w1 != 0.0 ? w1 * ys[6 * (n0 + 1) + comp] : 0.0);
    }
    rv = bv - sy;
    a.x[e] = xv;
    a.r[e] = rv;
    a.p0[e] = 0.0;
  }
  if (tid < kGroupThreads) rloc[tid] = rv;
  __syncthreads();
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
}

// Iteration k: stops when |r|^2 <= tol2 or k = max_iters, else
// q = S p_k with p_k = z - W mu + beta p_{k-1}, evaluated as
// S z + beta S p_{k-1} - (SW) mu so that the row's block products (the bulk
// of the loads) do not wait for the sums of the previous update's per-group
// partials (r'z, r'r and mu); the own row's p_k is stored and
// part[i] = p_k,i . q_i. Thread t < 252 handles component t % 6 of every 42nd
// block of the row and window columns t / 6 + 42 m of (SW)_i.
__global__ __launch_bounds__(256) void k_pcg_spmv(PcgArgs a, int32_t k, const double* __restrict__ p_prev,
                                                  double* __restrict__ p_cur) {
  __shared__ double chunk[kPartChunk];
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
  constexpr int kChunkPer = kPartChunk / 256;
  const int per_chunk = kPartChunk / stride, ng0 = min(per_chunk, a.groups);
  double pre[kChunkPer];
#pragma unroll
  for (int m = 0; m < kChunkPer; ++m) {
    const int q = tid + 256 * m;
    pre[m] = q < ng0 * stride ? a.part2[q] : 0.0;
  }
  double z_own = 0.0, p_own = 0.0, hw0 = 0.0, hw1 = 0.0;
  int hn0 = 0;
  if (tid < 6) {
    z_own = a.z[6 * i + tid];
    p_own = p_prev[6 * i + tid];
    if (ne) {
      hn0 = a.hat_node[i];
      hw0 = a.hat_w[2 * i];
      hw1 = a.hat_w[2 * i + 1];
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
  // Fixed-order sum of the group partials, a chunk of whole groups per pass.
#pragma unroll
  for (int m = 0; m < kChunkPer; ++m) chunk[tid + 256 * m] = pre[m];
  __syncthreads();
  double acc = 0.0;
  if (tid < stride)
    for (int gg = 0; gg < ng0; ++gg) acc += chunk[gg * stride + tid];
  for (int g0 = ng0; g0 < a.groups; g0 += per_chunk) {
    const int ng = min(per_chunk, a.groups - g0);
    __syncthreads();
    for (int q = tid; q < ng * stride; q += blockDim.x) chunk[q] = a.part2[(size_t)g0 * stride + q];
    __syncthreads();
    if (tid < stride)
      for (int gg = 0; gg < ng; ++gg) acc += chunk[gg * stride + tid];
  }
  if (tid < stride) sums[tid] = acc;
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
  }
  if (tid < kGroupThreads) {
    double w = 0.0;
    if (ne) {
#pragma unroll
      for (int m = 0; m < kSwCols; ++m) {
        const int lc = g + m * kRowsPerGroup;
        if (lc < a.gw && c0 + lc < ne) w += sw[m] * mu_s[c0 + lc];
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
    double pa = z_own + beta * p_own;
    if (ne) pa -= hw0 * mu_s[6 * hn0 + tid] + (hw1 != 0.0 ? hw1 * mu_s[6 * (hn0 + 1) + tid] : 0.0);
    a.q[6 * i + tid] = qa;
    p_cur[6 * i + tid] = pa;
    ysz[tid] = pa * qa;
  }
  __syncthreads();
  if (tid == 0) a.part[i] = ((ysz[0] + ysz[1]) + (ysz[2] + ysz[3])) + (ysz[4] + ysz[5]);
}

// alpha = r'z / p'q; x += alpha p; r -= alpha q; z = M^-1 r on one group of
// rows per work-group, and the group's partials r'z, r'r, E^-1 (SW)'z.
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
  load_group<false>(a, g, L);
  // The group's vectors and preconditioner rows, loaded before the reduction.
  double xv = 0.0, pv = 0.0, rv = 0.0, qv = 0.0, M[6] = {};
  if (active) {
    const int e = 6 * row + comp;
    xv = a.x[e];
    pv = p_cur[e];
    rv = a.r[e];
    qv = a.q[e];
    const double* Mr = a.Minv + (size_t)36 * row + 6 * comp;
#pragma unroll
    for (int c = 0; c < 6; ++c) M[c] = Mr[c];
  }
  // p'q, summed by every group in the same order.
  double pq = 0.0, unused0 = 0.0, unused1 = 0.0;
  for (int r = tid; r < P; r += blockDim.x) pq += a.part[r];
  if (done) return;  // the same value in every thread
  block_sum3(pq, unused0, unused1, lds);
  const double alpha = rz / pq;
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
    a.z[6 * row + comp] = zv;
  }
  if (tid < kGroupThreads) zloc[tid] = zv;
  double rzg = rv * zv, rr = rv * rv, zero = 0.0;
  block_sum3(rzg, rr, zero, lds);
  group_partials(a, g, L, zloc, rzg, rr, 0.0, cs, mpart, a.part2 + (size_t)g * (3 + ne));
}

// Clears the convergence flag and lifts the iteration limit so that one more
// iteration runs (kernel timing only).
__global__ void k_pcg_reopen(PcgState* st) {
  st->done = 0;
  st->max_iters = st->iters + 2;
  st->tol2 = -1.0;
}

// ------------------------------------------------------------ launchers

size_t pcg_state_bytes() { return sizeof(PcgState); }
int32_t pcg_max_coarse() { return kMaxCoarse; }
int32_t pcg_group_rows() { return kRowsPerGroup; }
int32_t pcg_update_groups(int32_t Pvar) { return (Pvar + kRowsPerGroup - 1) / kRowsPerGroup; }

// Deflated PCG solve of S y = b. Buffers (device): x, r, z, p0, p1, q
// (6 Pvar), part (Pvar), part2 (groups x (3 + ne)), SR (groups x gw x 252),
// E (ne x ne), Einv (ne x ne + groups x gw x ne), y (ne). The host polls the state after `first`
// iterations, then every `poll` iterations. With `probe` (three events),
// k_pcg_spmv and then k_pcg_update of one more iteration are each launched
// `probe_reps` times back to back after the solve and timed as two batches
// (k_pcg_spmv is idempotent; x, r and z then hold extra steps). Returns the iteration count or -1 on a launch error.
int launch_pcg_deflated(hipStream_t s, int32_t Pvar, int32_t K, const int32_t* row_ptr, const int32_t* col_var,
                        const double* S, const double* Minv, const double* b, const int32_t* hat_node,
                        const double* hat_w, const int32_t* node_range, const int32_t* grp_lo, int32_t gw,
                        double* SR, double* E, double* Einv, double* y,
                        double* x, double* r, double* z, double* p0, double* p1, double* q,
                        double* part, double* part2, void* state_dev, void* state_host, double tol,
                        int32_t max_iters, int32_t first, int32_t poll, const hipEvent_t* probe,
                        int32_t probe_reps) {
  if (Pvar <= 0) return 0;
  PcgArgs a;
  a.Pvar = Pvar;
  a.K = K;
  a.ne = 6 * K;
  a.groups = pcg_update_groups(Pvar);
  a.row_ptr = row_ptr; a.col_var = col_var; a.S = S; a.Minv = Minv; a.b = b;
  a.hat_node = hat_node; a.hat_w = hat_w; a.node_range = node_range; a.grp_lo = grp_lo; a.gw = gw; a.pad2 = 0;
  a.SR = SR; a.E = E; a.Einv = Einv; a.EinvW = Einv + (size_t)a.ne * a.ne; a.x = x; a.r = r; a.z = z; a.p0 = p0; a.p1 = p1; a.q = q;
  a.part = part; a.part2 = part2; a.y = y;
  a.st = static_cast<PcgState*>(state_dev);
  const int groups = a.groups;
  if (a.ne > 0) {
    k_defl_sr<<<Pvar, 256, 0, s>>>(a);
    k_defl_e<<<(a.ne * a.ne + 255) / 256, 256, 0, s>>>(a);
  }
  k_defl_prep<<<1, kPrepThreads, 0, s>>>(a, tol, max_iters);
  k_defl_init<<<groups, kGroupBlock, 0, s>>>(a);
  PcgState* hs = static_cast<PcgState*>(state_host);
  int k = 0;
  while (true) {
    const int n = std::min(k == 0 ? std::max(first, 1) : poll, max_iters - k);
    for (int j = 0; j < n; ++j, ++k) {
      k_pcg_spmv<<<Pvar, 256, 0, s>>>(a, k, (k & 1) ? p1 : p0, (k & 1) ? p0 : p1);
      k_pcg_update<<<groups, kGroupBlock, 0, s>>>(a, k, (k & 1) ? p0 : p1);
    }
    // At k = max_iters this launch only records the final state.
    if (k >= max_iters) k_pcg_spmv<<<Pvar, 256, 0, s>>>(a, k, p0, p1);
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
    for (int j = 0; j < probe_reps; ++j) k_pcg_spmv<<<Pvar, 256, 0, s>>>(a, kk, (kk & 1) ? p1 : p0, (kk & 1) ? p0 : p1);
    if (hipEventRecord(probe[1], s) != hipSuccess) return -1;
    for (int j = 0; j < probe_reps; ++j) k_pcg_update<<<groups, kGroupBlock, 0, s>>>(a, kk, (kk & 1) ? p0 : p1);
    if (hipEventRecord(probe[2], s) != hipSuccess) return -1;
    if (hipGetLastError() != hipSuccess) return -1;
  }
  return iters;
}

}  // namespace mlba
