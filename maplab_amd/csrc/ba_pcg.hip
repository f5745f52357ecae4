// ba_pcg.hip -- deflated, block-Jacobi preconditioned conjugate gradients on
// the reduced camera system S y = b (FP64).
//
// Block-Jacobi (inverse 6x6 diagonal blocks of S, Ceres' SCHUR_JACOBI) leaves
// a handful of very small eigenvalues on long trajectories: smooth, global
// modes of the vertex chain. They are removed by deflation (Saad, Yeung,
// Erhel, Guyomarc'h, "A deflated version of the conjugate gradient
// algorithm", SIAM J. Sci. Comput. 21, 2000) with a coarse space W = R of
// piecewise-linear "hat" functions over the vertex order (K nodes, 6 K
// columns): x0 = W E^-1 W'b, E = W'SW, and every search direction is made
// S-orthogonal to W with mu = E^-1 (SW)'z. The iteration converges to the
// same solution as plain PCG (the tolerance is on the true residual
// recursion), in 3-4x fewer iterations on the benchmark map.
//
// Per LM step: k_defl_sr (SW), k_defl_e (E), k_defl_inverse (E^-1, Gauss-
// Jordan in LDS), k_defl_init; per PCG iteration: k_pcg_spmv (S p with the
// direction update formed on the fly) and k_pcg_update (vector update; the
// last work-group to finish completes the reductions, no inter-group waits).
#include "ba_device.h"
#include "ba_reduce.h"

namespace mlba {

constexpr int kMaxCoarse = 90;   // 6 x 15 hat nodes; E^-1 by Gauss-Jordan in LDS
constexpr int kRowsPerGroup = 42;
constexpr int kGroupThreads = 6 * kRowsPerGroup;  // 252: one thread per (row, component)

struct PcgState {
  double rz;        // r'z of the current iterate
  double bb;        // |b|^2
  double rr;        // |r|^2
  double tol2;      // (tol |b|)^2
  double beta;      // direction update factor for the next iteration
  int32_t iters;
  int32_t done;
  int32_t max_iters;
  uint32_t counter; // work-groups of k_pcg_update that finished this iteration
};

struct PcgArgs {
  int32_t Pvar, ne;            // ne = 6 K coarse columns (0: no deflation)
  int32_t K, pad;
  const int32_t* row_ptr;
  const int32_t* col_var;
  const double* S;
  const double* Minv;
  const double* b;
  const int32_t* hat_node;     // Pvar: first hat node of the vertex (0 without deflation)
  const double* hat_w;         // 2 Pvar: weights of hat_node and hat_node + 1 (0 without deflation)
  const int32_t* node_range;   // 2 K: first and last vertex of each hat's support
  const int32_t* sr_lo;        // Pvar: first hat node of the window of row i of SW
  int32_t sw, pad2;            // window width of SW rows, in columns (6 x nodes)
  double* SR;                  // 6 Pvar x sw: row e = 6 i + c holds columns [6 sr_lo[i], 6 sr_lo[i] + sw) of SW
  double* E;                   // ne x ne
  double* Einv;                // ne x ne
  double* x;
  double* r;
  double* z;
  double* p0;
  double* p1;
  double* q;
  double* mu;                  // ne
  double* part;                // Pvar: p'q of each row
  double* part2;               // groups x (2 + ne)
  PcgState* st;
  int32_t* fail;
};

__device__ inline double hat_weight(const PcgArgs& a, int i, int k) {
  const int n0 = a.hat_node[i];
  return n0 == k ? a.hat_w[2 * i] : (n0 + 1 == k ? a.hat_w[2 * i + 1] : 0.0);
}

// SW, one block row of S per work-group: (SW)_i = sum_j S_ij W_j, W_j = w0 e_n0 + w1 e_n0+1.
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
  const int c0 = 6 * a.sr_lo[i], sw = a.sw;
  for (int k = threadIdx.x; k < 6 * sw; k += blockDim.x) {
    const int row = k / sw, gc = c0 + k % sw;
    a.SR[(size_t)(6 * i) * sw + k] = gc < ne ? acc[row * ne + gc] : 0.0;
  }
}

// Entry (e, col) of SW from the windowed storage.
__device__ inline double sw_entry(const PcgArgs& a, int e, int col) {
  const int lc = col - 6 * a.sr_lo[e / 6];
  return (lc >= 0 && lc < a.sw) ? a.SR[(size_t)e * a.sw + lc] : 0.0;
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

// E^-1 by Gauss-Jordan elimination without pivoting (E is SPD), in LDS.
// A non-positive pivot disables deflation for this solve (E^-1 = 0).
__global__ __launch_bounds__(1024) void k_defl_inverse(PcgArgs a) {
  __shared__ double M[kMaxCoarse * 2 * kMaxCoarse];
  __shared__ double fcol[kMaxCoarse];
  __shared__ int bad;
  const int ne = a.ne, w = 2 * ne;
  for (int t = threadIdx.x; t < ne * w; t += blockDim.x) {
    const int r = t / w, c = t % w;
    // symmetrize: lower triangle of E mirrored
    M[t] = c < ne ? (c <= r ? a.E[r * ne + c] : a.E[c * ne + r]) : (c - ne == r ? 1.0 : 0.0);
  }
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  for (int j = 0; j < ne; ++j) {
    const double piv = M[j * w + j];
    if (!(piv > 0.0)) {
      if (threadIdx.x == 0) bad = 1;
      break;
    }
    for (int r = threadIdx.x; r < ne; r += blockDim.x) fcol[r] = M[r * w + j];
    __syncthreads();
    for (int c = threadIdx.x; c < w; c += blockDim.x) M[j * w + c] /= piv;
    __syncthreads();
    for (int t = threadIdx.x; t < ne * w; t += blockDim.x) {
      const int r = t / w, c = t % w;
      if (r != j) M[t] -= fcol[r] * M[j * w + c];
    }
    __syncthreads();
  }
  __syncthreads();
  for (int t = threadIdx.x; t < ne * ne; t += blockDim.x) {
    const int r = t / ne, c = t % ne;
    a.Einv[t] = bad ? 0.0 : M[r * w + ne + c];
  }
}

// x0 = W E^-1 W'b, r0 = b - (SW) E^-1 W'b, z0 = M^-1 r0, mu0 = E^-1 (SW)'z0,
// p = 0 and the scalars (one work-group).
__global__ __launch_bounds__(1024) void k_defl_init(PcgArgs a, double tol, int32_t max_iters) {
  __shared__ double y[kMaxCoarse];
  __shared__ double cvec[kMaxCoarse];
  __shared__ double lds16[16];
  const int ne = a.ne, P = a.Pvar, tid = threadIdx.x;
  // c = W'b
  for (int rr = tid; rr < ne; rr += blockDim.x) {
    const int k = rr / 6, comp = rr % 6;
    double s = 0.0;
    for (int i = a.node_range[2 * k]; i <= a.node_range[2 * k + 1]; ++i) s += hat_weight(a, i, k) * a.b[6 * i + comp];
    cvec[rr] = s;
  }
  __syncthreads();
  for (int rr = tid; rr < ne; rr += blockDim.x) {
    double s = 0.0;
    for (int c = 0; c < ne; ++c) s += a.Einv[rr * ne + c] * cvec[c];
    y[rr] = s;
  }
  __syncthreads();
  double bb = 0.0, rz = 0.0, rr2 = 0.0;
  for (int i = tid; i < P; i += blockDim.x) {
    double rv[6];
    const int n0 = a.hat_node[i];
    const double w0 = a.hat_w[2 * i], w1 = a.hat_w[2 * i + 1];
    for (int comp = 0; comp < 6; ++comp) {
      const double bv = a.b[6 * i + comp];
      double sy = 0.0;
      const double* srow = a.SR + (size_t)(6 * i + comp) * a.sw;
      const int c0 = 6 * a.sr_lo[i];
      for (int c = 0; c < a.sw && c0 + c < ne; ++c) sy += srow[c] * y[c0 + c];
      rv[comp] = bv - sy;
      a.r[6 * i + comp] = rv[comp];
      a.x[6 * i + comp] = ne ? w0 * y[6 * n0 + comp] + (w1 != 0.0 ? w1 * y[6 * (n0 + 1) + comp] : 0.0) : 0.0;
      a.p0[6 * i + comp] = 0.0;
      bb += bv * bv;
      rr2 += rv[comp] * rv[comp];
    }
    const double* M = a.Minv + (size_t)36 * i;
    for (int comp = 0; comp < 6; ++comp) {
      double s = 0.0;
      for (int c = 0; c < 6; ++c) s += M[comp * 6 + c] * rv[c];
      a.z[6 * i + comp] = s;
      rz += rv[comp] * s;
    }
  }
  __syncthreads();
  // t = (SW)'z0 with LDS atomics over the windowed rows, then mu0 = E^-1 t.
  if (ne > 0) {
    for (int k = tid; k < ne; k += blockDim.x) cvec[k] = 0.0;
    __syncthreads();
    const int sw = a.sw;
    for (size_t el = tid; el < (size_t)6 * P * sw; el += blockDim.x) {
      const int e = (int)(el / sw), lc = (int)(el % sw);
      const int gc = 6 * a.sr_lo[e / 6] + lc;
      const double v = a.SR[el];
      if (gc < ne && v != 0.0) atomicAdd(&cvec[gc], v * a.z[e]);
    }
    __syncthreads();
    for (int rrow = tid; rrow < ne; rrow += blockDim.x) {
      double m = 0.0;
      for (int c = 0; c < ne; ++c) m += a.Einv[rrow * ne + c] * cvec[c];
      a.mu[rrow] = m;
    }
  }
  // block sums (1024 threads = 16 waves)
  auto bsum = [&](double v) {
    v = wave_sum(v);
    __syncthreads();
    if ((tid & 63) == 0) lds16[tid >> 6] = v;
    __syncthreads();
    double s = 0.0;
    for (int k = 0; k < 16; ++k) s += lds16[k];
    return s;
  };
  bb = bsum(bb);
  rz = bsum(rz);
  rr2 = bsum(rr2);
  if (tid == 0) {
    PcgState* st = a.st;
    st->rz = rz;
    st->bb = bb;
    st->rr = rr2;
    st->tol2 = tol * tol * bb;
    st->beta = 0.0;
    st->iters = 0;
    st->max_iters = max_iters;
    st->counter = 0;
    st->done = (bb <= 0.0 || max_iters <= 0 || rr2 <= st->tol2) ? 1 : 0;
  }
}

// q = S p_k with p_k = z - W mu + beta p_{k-1} formed on the fly (the own
// row is stored as the new direction); part[i] = p_k,i . q_i.
// Thread t < 252 handles component t % 6 of every 42nd block of the row.
__global__ __launch_bounds__(256) void k_pcg_spmv(PcgArgs a, const double* __restrict__ p_prev,
                                                  double* __restrict__ p_cur) {
  __shared__ double mu_s[kMaxCoarse];
  __shared__ double ysum[kGroupThreads];
  const PcgState* st = a.st;
  if (st->done) return;
  const int ne = a.ne, tid = threadIdx.x;
  for (int k = tid; k < ne; k += blockDim.x) mu_s[k] = a.mu[k];
  __syncthreads();
  const double beta = st->beta;
  const int i = blockIdx.x;
  const int rb = a.row_ptr[i], re = a.row_ptr[i + 1];
  double y = 0.0;
  if (tid < kGroupThreads) {
    const int comp = tid % 6, g = tid / 6;
    for (int k = rb + g; k < re; k += kRowsPerGroup) {
      const int j = a.col_var[k];
      const double* srow = a.S + (size_t)36 * k + 6 * comp;
      const int n0 = a.hat_node[j];
      const double w0 = a.hat_w[2 * j], w1 = a.hat_w[2 * j + 1];
      double s = 0.0;
      for (int c = 0; c < 6; ++c) {
        double v = a.z[6 * j + c] + beta * p_prev[6 * j + c];
        if (ne) v -= w0 * mu_s[6 * n0 + c] + (w1 != 0.0 ? w1 * mu_s[6 * (n0 + 1) + c] : 0.0);
        s += srow[c] * v;
      }
      y += s;
    }
    ysum[tid] = y;
  }
  __syncthreads();
  if (tid < 6) {
    double qa = 0.0;
    for (int g = 0; g < kRowsPerGroup; ++g) qa += ysum[6 * g + tid];
    double pa = a.z[6 * i + tid] + beta * p_prev[6 * i + tid];
    if (ne) {
      const int n0 = a.hat_node[i];
      const double w0 = a.hat_w[2 * i], w1 = a.hat_w[2 * i + 1];
      pa -= w0 * mu_s[6 * n0 + tid] + (w1 != 0.0 ? w1 * mu_s[6 * (n0 + 1) + tid] : 0.0);
    }
    a.q[6 * i + tid] = qa;
    p_cur[6 * i + tid] = pa;
    ysum[tid] = pa * qa;
  }
  __syncthreads();
  if (tid == 0) a.part[i] = ((ysum[0] + ysum[1]) + (ysum[2] + ysum[3])) + (ysum[4] + ysum[5]);
}

// alpha = r'z / p'q; x += alpha p; r -= alpha q; z = M^-1 r on 42 rows per
// work-group; partial r'z, r'r and (SW)'z. The last work-group to finish
// (atomic ticket) sums the partials in group order and sets mu = E^-1 (SW)'z,
// beta, the iteration count and the convergence flag.
__global__ __launch_bounds__(256) void k_pcg_update(PcgArgs a, const double* __restrict__ p_cur) {
  __shared__ double rloc[kGroupThreads];
  __shared__ double zloc[kGroupThreads];
  __shared__ double lds4[4];
  __shared__ double cs[kMaxCoarse];
  __shared__ double bc[2];
  __shared__ int is_last;
  PcgState* st = a.st;
  if (st->done) return;
  const int ne = a.ne, tid = threadIdx.x, P = a.Pvar;
  // p'q, summed by every group in the same order.
  double pq = 0.0;
  for (int k = tid; k < P; k += blockDim.x) pq += a.part[k];
  pq = block_sum256(pq, lds4);
  if (tid == 0) bc[0] = pq;
  __syncthreads();
  const double rz_old = st->rz;
  const double alpha = rz_old / bc[0];
  const int row = blockIdx.x * kRowsPerGroup + tid / 6, comp = tid % 6;
  const bool active = tid < kGroupThreads && row < P;
  double rv = 0.0;
  if (active) {
    const int e = 6 * row + comp;
    a.x[e] += alpha * p_cur[e];
    rv = a.r[e] - alpha * a.q[e];
    a.r[e] = rv;
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
  const double rz_l = block_sum256(rv * zv, lds4);
  const double rr_l = block_sum256(rv * rv, lds4);
  double* out = a.part2 + (size_t)blockIdx.x * (2 + ne);
  if (tid == 0) { out[0] = rz_l; out[1] = rr_l; }
  if (ne > 0) {
    // (SW)'z over the windowed rows of this group, all threads, LDS atomics.
    for (int k = tid; k < ne; k += blockDim.x) cs[k] = 0.0;
    __syncthreads();
    const int row0 = blockIdx.x * kRowsPerGroup, sw = a.sw;
    const int nent = 6 * min(kRowsPerGroup, P - row0);
    const double* base = a.SR + (size_t)(6 * row0) * sw;
    for (int el = tid; el < nent * sw; el += blockDim.x) {
      const int t = el / sw, lc = el % sw;
      const int gc = 6 * a.sr_lo[row0 + t / 6] + lc;
      const double v = base[el];
      if (gc < ne && v != 0.0) atomicAdd(&cs[gc], v * zloc[t]);
    }
    __syncthreads();
    for (int col = tid; col < ne; col += blockDim.x) out[2 + col] = cs[col];
  }
  __threadfence();
  __syncthreads();
  if (tid == 0) is_last = atomicAdd(&st->counter, 1u) == gridDim.x - 1;
  __syncthreads();
  if (!is_last) return;
  __threadfence();
  const int G = gridDim.x;
  double rz = 0.0, rr = 0.0;
  if (tid == 0)
    for (int g = 0; g < G; ++g) {
      rz += a.part2[(size_t)g * (2 + ne)];
      rr += a.part2[(size_t)g * (2 + ne) + 1];
    }
  for (int col = tid; col < ne; col += blockDim.x) {
    double s = 0.0;
    for (int g = 0; g < G; ++g) s += a.part2[(size_t)g * (2 + ne) + 2 + col];
    cs[col] = s;
  }
  __syncthreads();
  for (int rrow = tid; rrow < ne; rrow += blockDim.x) {
    double m = 0.0;
    for (int c = 0; c < ne; ++c) m += a.Einv[rrow * ne + c] * cs[c];
    a.mu[rrow] = m;
  }
  if (tid == 0) {
    st->rz = rz;
    st->rr = rr;
    st->beta = rz / rz_old;
    st->iters += 1;
    st->counter = 0;
    if (rr <= st->tol2 || st->iters >= st->max_iters) st->done = 1;
  }
}

// ------------------------------------------------------------ launchers

size_t pcg_state_bytes() { return sizeof(PcgState); }
int32_t pcg_max_coarse() { return kMaxCoarse; }
int32_t pcg_update_groups(int32_t Pvar) { return (Pvar + kRowsPerGroup - 1) / kRowsPerGroup; }

// Deflated PCG solve of S y = b. Buffers (device): x, r, z, p0, p1, q, mu
// (6 Pvar / ne), part (Pvar), part2 (groups x (2 + ne)), SR (6 Pvar x ne),
// E / Einv (ne x ne). Returns the iteration count or -1 on a launch error.
int launch_pcg_deflated(hipStream_t s, int32_t Pvar, int32_t K, const int32_t* row_ptr, const int32_t* col_var,
                        const double* S, const double* Minv, const double* b, const int32_t* hat_node,
                        const double* hat_w, const int32_t* node_range, const int32_t* sr_lo, int32_t sw,
                        double* SR, double* E, double* Einv,
                        double* x, double* r, double* z, double* p0, double* p1, double* q, double* mu,
                        double* part, double* part2, void* state_dev, void* state_host, int32_t* fail, double tol,
                        int32_t max_iters, int32_t poll) {
  if (Pvar <= 0) return 0;
  PcgArgs a;
  a.Pvar = Pvar;
  a.K = K;
  a.ne = 6 * K;
  a.pad = 0;
  a.row_ptr = row_ptr; a.col_var = col_var; a.S = S; a.Minv = Minv; a.b = b;
  a.hat_node = hat_node; a.hat_w = hat_w; a.node_range = node_range; a.sr_lo = sr_lo; a.sw = sw; a.pad2 = 0;
  a.SR = SR; a.E = E; a.Einv = Einv; a.x = x; a.r = r; a.z = z; a.p0 = p0; a.p1 = p1; a.q = q; a.mu = mu;
  a.part = part; a.part2 = part2;
  a.st = static_cast<PcgState*>(state_dev);
  a.fail = fail;
  if (a.ne > 0) {
    k_defl_sr<<<Pvar, 256, 0, s>>>(a);
    k_defl_e<<<(a.ne * a.ne + 255) / 256, 256, 0, s>>>(a);
    k_defl_inverse<<<1, 1024, 0, s>>>(a);
  }
  k_defl_init<<<1, 1024, 0, s>>>(a, tol, max_iters);
  const int groups = pcg_update_groups(Pvar);
  PcgState* hs = static_cast<PcgState*>(state_host);
  int launched = 0;
  while (launched < max_iters) {
    const int n = std::min(poll, max_iters - launched);
    for (int k = 0; k < n; ++k) {
      double* p_prev = ((launched + k) & 1) ? p1 : p0;
      double* p_cur = ((launched + k) & 1) ? p0 : p1;
      k_pcg_spmv<<<Pvar, 256, 0, s>>>(a, p_prev, p_cur);
      k_pcg_update<<<groups, 256, 0, s>>>(a, p_cur);
    }
    launched += n;
    if (hipMemcpyAsync(hs, a.st, sizeof(PcgState), hipMemcpyDeviceToHost, s) != hipSuccess) return -1;
    if (hipStreamSynchronize(s) != hipSuccess) return -1;
    if (hs->done) break;
  }
  if (hipGetLastError() != hipSuccess) return -1;
  return hs->iters;
}

}  // namespace mlba
