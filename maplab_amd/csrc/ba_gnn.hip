// ba_gnn.hip -- the covisibility-graph network that predicts the per-row
// coefficients theta_i of the learned preconditioner (ba_learned.h), its
// backward pass and the Adam step, as hand-written gfx950 kernels. No
// counterpart in the reference (Ceres' preconditioners are fixed).
//
// Network (hidden width 64, BF16 activations, FP32 parameters and
// accumulation):
//   h0   = bf16(relu(x W_in + b_in))                              (FP32 FMA, 9 inputs)
//   for l = 1..3:
//     y_l   = bf16(h_{l-1} bf16(W_m) + b_m)                       (MFMA)
//     g_e   = sigmoid(e_e . a_l + c_l)                            (edge gate)
//     agg_i = bf16(sum_{e into i} g_e y_{l,src(e)} / deg_i)       (edge order)
//     h_l   = bf16(relu([h_{l-1} | agg_l] bf16(W_u) + b_u))        (MFMA, K = 128)
//   theta = h_3 W_out + b_out                                      (FP32)
// Dense products use v_mfma_f32_32x32x8_bf16 (one wave per 32-row tile);
// the weight-gradient products A'B reduce over rows in chunks of 256 with
// fixed-order chunk sums, so a gradient repeats bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ba_learned.h"
#include "ba_reduce.h"

namespace mlba {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int kH = kGnnHidden;
constexpr int kChunk = 256;  // rows per partial of the A'B reductions

// Round to nearest even (finite inputs).
__host__ __device__ inline uint16_t to_bf16(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  const uint32_t lsb = (u >> 16) & 1u;
  u += 0x7FFFu + lsb;
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float from_bf16(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// Parameter offsets.
struct Layout {
  int w_in, b_in, w_out, b_out;
  int w_m[kGnnLayers], b_m[kGnnLayers], a_g[kGnnLayers], c_g[kGnnLayers], w_u[kGnnLayers], b_u[kGnnLayers];
};
Layout layout() {
  Layout L{};
  int o = 0;
  L.w_in = o; o += kLpNodeFeatures * kH;
  L.b_in = o; o += kH;
  for (int l = 0; l < kGnnLayers; ++l) {
    L.w_m[l] = o; o += kH * kH;
    L.b_m[l] = o; o += kH;
    L.a_g[l] = o; o += kLpEdgeFeatures;
    L.c_g[l] = o; o += 1;
    L.w_u[l] = o; o += 2 * kH * kH;
    L.b_u[l] = o; o += kH;
  }
  L.w_out = o; o += kH * kLpTerms;
  L.b_out = o; o += kLpTerms;
  return L;
}

// ------------------------------------------------------------ MFMA products

// C[M x N] = A[M x K] B[K x N] + bias (relu optional), A, B BF16 row-major,
// N = 32 * nt (nt <= 4), K % 8 == 0. One wave per 32-row tile, four waves per
// work-group. Rows >= m_real are written as zeros. Output BF16 (out16, ldo,
// column offset col0) or FP32 (out32, ldo, col0; accumulate adds to it).
template <int kNt>
__global__ __launch_bounds__(256) void k_mfma_gemm(int32_t m_pad, int32_t m_real, int32_t K,
                                                   const uint16_t* __restrict__ A, int32_t lda,
                                                   const uint16_t* __restrict__ B, const float* __restrict__ bias,
                                                   int32_t relu, uint16_t* __restrict__ out16,
                                                   float* __restrict__ out32, int32_t ldo, int32_t col0,
                                                   int32_t accumulate) {
  const int lane = threadIdx.x & 63;
  const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int row0 = tile * 32;
  if (row0 >= m_pad) return;
  constexpr int N = 32 * kNt;
  floatx16 acc[kNt];
  for (int t = 0; t < kNt; ++t)
    for (int v = 0; v < 16; ++v) acc[t][v] = 0.0f;
  const int ar = row0 + (lane & 31), kq = 4 * (lane >> 5);
  for (int k0 = 0; k0 < K; k0 += 8) {
    bf16x4 a;
    const uint16_t* ap = A + (size_t)ar * lda + k0 + kq;
    for (int i = 0; i < 4; ++i) a[i] = (short)ap[i];
    for (int t = 0; t < kNt; ++t) {
      bf16x4 b;
      const uint16_t* bp = B + (size_t)(k0 + kq) * N + 32 * t + (lane & 31);
      for (int i = 0; i < 4; ++i) b[i] = (short)bp[(size_t)i * N];
      acc[t] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc[t], 0, 0, 0);
    }
  }
  // D layout: column lane % 32, row 8 (v / 4) + 4 (lane / 32) + v % 4.
  for (int t = 0; t < kNt; ++t) {
    const int col = 32 * t + (lane & 31);
    const float bb = bias ? bias[col] : 0.0f;
    for (int v = 0; v < 16; ++v) {
      const int row = row0 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
      float x = acc[t][v] + bb;
      if (relu) x = fmaxf(x, 0.0f);
      if (row >= m_real) x = 0.0f;
      const size_t at = (size_t)row * ldo + col0 + col;
      if (out16) {
        out16[at] = to_bf16(x);
      } else {
        out32[at] = accumulate ? out32[at] + (row < m_real ? acc[t][v] : 0.0f) : x;
      }
    }
  }
}

// part[chunk][K x N] = sum over the chunk's rows m of A[m][0..K)' B[m][0..N),
// A BF16 (lda), B BF16 (ldb), K = 32 kt, N = 32 nt. Four waves split the
// (kt x nt) output tiles; all iterate over the chunk's rows 8 at a time.
__global__ __launch_bounds__(256) void k_mfma_gemm_tn(int32_t m_pad, int32_t kt, int32_t nt,
                                                      const uint16_t* __restrict__ A, int32_t lda,
                                                      const uint16_t* __restrict__ B, int32_t ldb,
                                                      float* __restrict__ part) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m0 = blockIdx.x * kChunk;
  const int mend = min(m_pad, m0 + kChunk);
  const int K = 32 * kt, N = 32 * nt;
  float* out = part + (size_t)blockIdx.x * K * N;
  for (int tile = wave; tile < kt * nt; tile += 4) {
    const int f0 = 32 * (tile / nt), n0 = 32 * (tile % nt);
    floatx16 acc;
    for (int v = 0; v < 16; ++v) acc[v] = 0.0f;
    for (int m = m0; m < mend; m += 8) {
      bf16x4 a, b;
      const int mm = m + 4 * (lane >> 5);
      for (int i = 0; i < 4; ++i) {
        a[i] = (short)A[(size_t)(mm + i) * lda + f0 + (lane & 31)];
        b[i] = (short)B[(size_t)(mm + i) * ldb + n0 + (lane & 31)];
      }
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
    for (int v = 0; v < 16; ++v) {
      const int f = f0 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
      out[(size_t)f * N + n0 + (lane & 31)] = acc[v];
    }
  }
}

// out[e] = sum over chunks c (in order) of part[c][e], e < n; scale applied.
__global__ void k_sum_chunks(int32_t chunks, int32_t n, const float* __restrict__ part, float* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  float s = 0.0f;
  for (int c = 0; c < chunks; ++c) s += part[(size_t)c * n + e];
  out[e] = s;
}

// Column sums of a [rows x cols] matrix (BF16 or FP32) per chunk of 256
// rows: part[chunk][cols].
__global__ void k_col_partials(int32_t rows, int32_t cols, const uint16_t* __restrict__ a16,
                               const float* __restrict__ a32, int32_t lda, float* __restrict__ part) {
  const int c = threadIdx.x;
  if (c >= cols) return;
  const int m0 = blockIdx.x * kChunk, mend = min(rows, m0 + kChunk);
  float s = 0.0f;
  for (int m = m0; m < mend; ++m) s += a16 ? from_bf16(a16[(size_t)m * lda + c]) : a32[(size_t)m * lda + c];
  part[(size_t)blockIdx.x * cols + c] = s;
}

// ------------------------------------------------------------ element-wise layers

// h0 = bf16(relu(x W_in + b_in)) into HA[:, 0:64] (row stride 128).
__global__ void k_gnn_input(int32_t nodes, const float* __restrict__ x, const float* __restrict__ w,
                            const float* __restrict__ b, uint16_t* __restrict__ ha) {
  const int i = blockIdx.x, c = threadIdx.x;
  if (i >= nodes) return;
  float s = b[c];
  for (int k = 0; k < kLpNodeFeatures; ++k) s += x[(size_t)kLpNodeFeatures * i + k] * w[k * kH + c];
  ha[(size_t)i * 2 * kH + c] = to_bf16(fmaxf(s, 0.0f));
}

__device__ inline float gate(const float* __restrict__ ef, int e, const float* __restrict__ a, float c) {
  float s = c;
  for (int k = 0; k < kLpEdgeFeatures; ++k) s += a[k] * ef[(size_t)kLpEdgeFeatures * e + k];
  return 1.0f / (1.0f + expf(-s));
}

// Edges into each node (slots other than the node itself).
__global__ void k_gnn_degree(GnnGraph g, int32_t* __restrict__ deg) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.nodes) return;
  int d = 0;
  for (int e = g.row_ptr[i]; e < g.row_ptr[i + 1]; ++e) d += g.col[e] != i;
  deg[i] = d;
}

// agg_i = bf16(sum_e g_e y_src / deg_i) into HA[:, 64:128]; one wave per node.
__global__ __launch_bounds__(64) void k_gnn_aggregate(GnnGraph g, const int32_t* __restrict__ deg,
                                                      const uint16_t* __restrict__ y, const float* __restrict__ a,
                                                      const float* __restrict__ cg, uint16_t* __restrict__ ha) {
  const int i = blockIdx.x, f = threadIdx.x;
  const float c = cg[0];
  float s = 0.0f;
  for (int e = g.row_ptr[i]; e < g.row_ptr[i + 1]; ++e) {
    const int j = g.col[e];
    if (j == i) continue;
    s += gate(g.edge_feat, e, a, c) * from_bf16(y[(size_t)j * kH + f]);
  }
  ha[(size_t)i * 2 * kH + kH + f] = to_bf16(deg[i] ? s / (float)deg[i] : 0.0f);
}

// theta = h3 W_out + b_out (FP32); one thread per (node, term).
__global__ void k_gnn_output(int32_t nodes, const uint16_t* __restrict__ h3, const float* __restrict__ w,
                             const float* __restrict__ b, float* __restrict__ theta) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nodes * kLpTerms) return;
  const int i = t / kLpTerms, o = t % kLpTerms;
  float s = b[o];
  for (int k = 0; k < kH; ++k) s += from_bf16(h3[(size_t)i * kH + k]) * w[k * kLpTerms + o];
  theta[t] = s;
}

// Loss terms and d loss / d theta. qc: a, b, e, d, has per node.
// lpart[i] = has_i log t_i (FP32 of an FP64 evaluation), dtheta FP32.
__global__ void k_gnn_loss(int32_t nodes, const float* __restrict__ theta, const float* __restrict__ qc,
                           float inv_count, float* __restrict__ lpart, float* __restrict__ dtheta) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nodes) return;
  const float* r = qc + 5 * (size_t)i;
  const double a = r[0], b = r[1], e = r[2], d = r[3];
  const double t1 = theta[2 * i], t2 = theta[2 * i + 1];
  if (r[4] == 0.0f) {
    lpart[i] = 0.0f;
    dtheta[2 * i] = dtheta[2 * i + 1] = 0.0f;
    return;
  }
  // t = 6 - 2 theta'c + theta'Q theta, c = (a, b), Q = [[a + b, b + e], [b + e, e + d]].
  const double q11 = a + b, q12 = b + e, q22 = e + d;
  const double t = 6.0 - 2.0 * (t1 * a + t2 * b) + t1 * (q11 * t1 + q12 * t2) + t2 * (q12 * t1 + q22 * t2);
  const double tt = fmax(t, 1e-12);
  lpart[i] = (float)(inv_count * log(tt));
  dtheta[2 * i] = (float)(inv_count * (2.0 * (q11 * t1 + q12 * t2) - 2.0 * a) / tt);
  dtheta[2 * i + 1] = (float)(inv_count * (2.0 * (q12 * t1 + q22 * t2) - 2.0 * b) / tt);
}

// dz3 = bf16((dtheta W_out') * [h3 > 0]); one thread per (node, feature).
__global__ void k_gnn_output_back(int32_t nodes, int32_t m_pad, const float* __restrict__ dtheta,
                                  const float* __restrict__ w, const uint16_t* __restrict__ h3,
                                  uint16_t* __restrict__ dz) {
  const int i = blockIdx.x, c = threadIdx.x;
  if (i >= m_pad) return;
  float s = 0.0f;
  if (i < nodes && from_bf16(h3[(size_t)i * kH + c]) > 0.0f)
    for (int o = 0; o < kLpTerms; ++o) s += dtheta[kLpTerms * i + o] * w[c * kLpTerms + o];
  dz[(size_t)i * kH + c] = to_bf16(s);
}

// Per chunk of 256 nodes: partial sums of h3' dtheta (64 x 2) and dtheta (2).
__global__ void k_gnn_output_wgrad(int32_t nodes, const uint16_t* __restrict__ h3, const float* __restrict__ dtheta,
                                   float* __restrict__ part) {
  const int t = threadIdx.x;  // 0..129
  const int m0 = blockIdx.x * kChunk, mend = min(nodes, m0 + kChunk);
  float s = 0.0f;
  if (t < kH * kLpTerms) {
    const int k = t / kLpTerms, o = t % kLpTerms;
    for (int m = m0; m < mend; ++m) s += from_bf16(h3[(size_t)m * kH + k]) * dtheta[kLpTerms * m + o];
  } else if (t < kH * kLpTerms + kLpTerms) {
    const int o = t - kH * kLpTerms;
    for (int m = m0; m < mend; ++m) s += dtheta[kLpTerms * m + o];
  } else {
    return;
  }
  part[(size_t)blockIdx.x * (kH * kLpTerms + kLpTerms) + t] = s;
}

// Backward of the aggregation of node j (one wave): dy_j = bf16(sum over the
// edges (dst i, src j) of g / deg_i dagg_i), and per node the partial
// gradients of the gate parameters over the edges into j:
// gpart[j] = (d a_g (3), d c_g) with dg_e = <dagg_j, y_src> / deg_j.
__global__ __launch_bounds__(64) void k_gnn_aggregate_back(GnnGraph g, int32_t m_pad, const int32_t* __restrict__ deg,
                                                           const uint16_t* __restrict__ y,
                                                           const float* __restrict__ dha,
                                                           const float* __restrict__ a, const float* __restrict__ cg,
                                                           uint16_t* __restrict__ dy, float* __restrict__ gpart) {
  const int j = blockIdx.x, f = threadIdx.x;
  if (j >= g.nodes) {
    if (j < m_pad) dy[(size_t)j * kH + f] = 0;
    return;
  }
  const float c = cg[0];
  const int degj = deg[j];
  const float dagg_j = dha[(size_t)j * 2 * kH + kH + f];
  float s = 0.0f;
  float ga[kLpEdgeFeatures + 1] = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int e = g.row_ptr[j]; e < g.row_ptr[j + 1]; ++e) {
    const int i = g.col[e];
    if (i == j) continue;
    // Into i from j: the reverse slot, in row i.
    const int er = g.rev[e];
    s += gate(g.edge_feat, er, a, c) / (float)deg[i] * dha[(size_t)i * 2 * kH + kH + f];
    // Edge e into j from i: d gate.
    const float dg = wave_sum_f(dagg_j * from_bf16(y[(size_t)i * kH + f])) / (float)degj;
    const float ge = gate(g.edge_feat, e, a, c);
    const float ds = dg * ge * (1.0f - ge);
    for (int k = 0; k < kLpEdgeFeatures; ++k) ga[k] += ds * g.edge_feat[(size_t)kLpEdgeFeatures * e + k];
    ga[kLpEdgeFeatures] += ds;
  }
  dy[(size_t)j * kH + f] = to_bf16(s);
  if (f < kLpEdgeFeatures + 1) gpart[(size_t)(kLpEdgeFeatures + 1) * j + f] = ga[f];
}

// dz = bf16(dh * [h > 0]) for h in HA[:, 0:64], dh in dHA[:, 0:64].
__global__ void k_gnn_relu_back(int32_t nodes, int32_t m_pad, const uint16_t* __restrict__ ha,
                                const float* __restrict__ dha, uint16_t* __restrict__ dz) {
  const int i = blockIdx.x, c = threadIdx.x;
  if (i >= m_pad) return;
  float s = 0.0f;
  if (i < nodes && from_bf16(ha[(size_t)i * 2 * kH + c]) > 0.0f) s = dha[(size_t)i * 2 * kH + c];
  dz[(size_t)i * kH + c] = to_bf16(s);
}

// Input layer gradient per chunk: x' dz0 (9 x 64) and column sums of dz0 (64),
// dz0 = dh0 * [h0 > 0] in FP32.
__global__ void k_gnn_input_back(int32_t nodes, const float* __restrict__ x, const uint16_t* __restrict__ ha,
                                 const float* __restrict__ dha, float* __restrict__ part) {
  const int c = threadIdx.x;  // 0..63
  const int m0 = blockIdx.x * kChunk, mend = min(nodes, m0 + kChunk);
  float s[kLpNodeFeatures + 1];
  for (int k = 0; k <= kLpNodeFeatures; ++k) s[k] = 0.0f;
  for (int m = m0; m < mend; ++m) {
    if (!(from_bf16(ha[(size_t)m * 2 * kH + c]) > 0.0f)) continue;
    const float d = dha[(size_t)m * 2 * kH + c];
    for (int k = 0; k < kLpNodeFeatures; ++k) s[k] += x[(size_t)kLpNodeFeatures * m + k] * d;
    s[kLpNodeFeatures] += d;
  }
  float* out = part + (size_t)blockIdx.x * (kLpNodeFeatures + 1) * kH;
  for (int k = 0; k <= kLpNodeFeatures; ++k) out[k * kH + c] = s[k];
}

// BF16 copies of the dense-layer weights: W_m, W_u and their transposes.
__global__ void k_gnn_weights(const float* __restrict__ params, Layout L, uint16_t* __restrict__ wm,
                              uint16_t* __restrict__ wu, uint16_t* __restrict__ wmt, uint16_t* __restrict__ wut) {
  const int l = blockIdx.y;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < kH * kH) {
    const int r = e / kH, c = e % kH;
    const uint16_t v = to_bf16(params[L.w_m[l] + e]);
    wm[(size_t)l * kH * kH + e] = v;
    wmt[(size_t)l * kH * kH + c * kH + r] = v;
  }
  if (e < 2 * kH * kH) {
    const int r = e / kH, c = e % kH;  // W_u [128 x 64]
    const uint16_t v = to_bf16(params[L.w_u[l] + e]);
    wu[(size_t)l * 2 * kH * kH + e] = v;
    wut[(size_t)l * 2 * kH * kH + c * 2 * kH + r] = v;  // [64 x 128]
  }
}

__global__ void k_adam(int32_t n, float lr, float c1, float c2, const float* __restrict__ g, float* __restrict__ p,
                       float* __restrict__ m, float* __restrict__ v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = 0.9f * m[i] + 0.1f * gi;
  const float vi = 0.999f * v[i] + 0.001f * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
}

}  // namespace

// ------------------------------------------------------------ workspace

class GnnWorkspace {
 public:
  int32_t nodes = 0, m_pad = 0, slots = 0;
  bool training = false;
  uint16_t* ha[kGnnLayers] = {};   // [m_pad x 128]: h_{l-1} | agg_l
  uint16_t* h3 = nullptr;          // [m_pad x 64]
  uint16_t* y[kGnnLayers] = {};    // [m_pad x 64]
  uint16_t* wm = nullptr;          // 3 x [64 x 64]
  uint16_t* wu = nullptr;          // 3 x [128 x 64]
  uint16_t* wmt = nullptr;         // 3 x [64 x 64]
  uint16_t* wut = nullptr;         // 3 x [64 x 128]
  // training
  float* dtheta = nullptr;         // m_pad x 2
  float* lpart = nullptr;          // m_pad
  uint16_t* dz = nullptr;          // m_pad x 64
  uint16_t* dy = nullptr;          // m_pad x 64
  float* dha = nullptr;            // m_pad x 128
  float* gpart = nullptr;          // m_pad x 4
  float* part = nullptr;           // chunks x 128 x 64
  int32_t* deg = nullptr;          // m_pad: edges into each node
  std::vector<void*> allocs;
  bool ok = true;

  template <typename T>
  T* get(size_t n) {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) { ok = false; return nullptr; }
    (void)hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(T));
    allocs.push_back(p);
    return static_cast<T*>(p);
  }
  ~GnnWorkspace() {
    for (void* p : allocs) (void)hipFree(p);
  }
};

GnnWorkspace* gnn_workspace_create(int32_t nodes, int32_t slots, bool for_training) {
  GnnWorkspace* w = new GnnWorkspace();
  w->nodes = nodes;
  w->slots = slots;
  w->m_pad = std::max(128, ((nodes + 127) / 128) * 128);
  w->training = for_training;
  const size_t M = w->m_pad;
  for (int l = 0; l < kGnnLayers; ++l) {
    w->ha[l] = w->get<uint16_t>(M * 2 * kH);
    w->y[l] = w->get<uint16_t>(M * kH);
  }
  w->h3 = w->get<uint16_t>(M * kH);
  w->deg = w->get<int32_t>(M);
  w->wm = w->get<uint16_t>(kGnnLayers * kH * kH);
  w->wu = w->get<uint16_t>(kGnnLayers * 2 * kH * kH);
  w->wmt = w->get<uint16_t>(kGnnLayers * kH * kH);
  w->wut = w->get<uint16_t>(kGnnLayers * 2 * kH * kH);
  if (for_training) {
    w->dtheta = w->get<float>(M * kLpTerms);
    w->lpart = w->get<float>(M);
    w->dz = w->get<uint16_t>(M * kH);
    w->dy = w->get<uint16_t>(M * kH);
    w->dha = w->get<float>(M * 2 * kH);
    w->gpart = w->get<float>(M * (kLpEdgeFeatures + 1));
    w->part = w->get<float>((M / kChunk + 1) * 2 * kH * kH);
  }
  if (!w->ok) { delete w; return nullptr; }
  return w;
}

void gnn_workspace_destroy(GnnWorkspace* w) { delete w; }

namespace {

void gemm(hipStream_t s, const GnnWorkspace* w, int K, int nt, const uint16_t* A, int lda, const uint16_t* B,
          const float* bias, bool relu, uint16_t* out16, float* out32, int ldo, int col0, bool accumulate) {
  const int tiles = w->m_pad / 32;
  const dim3 grid((tiles + 3) / 4);
  if (nt == 2)
    k_mfma_gemm<2><<<grid, 256, 0, s>>>(w->m_pad, w->nodes, K, A, lda, B, bias, relu, out16, out32, ldo, col0,
                                        accumulate);
  else
    k_mfma_gemm<4><<<grid, 256, 0, s>>>(w->m_pad, w->nodes, K, A, lda, B, bias, relu, out16, out32, ldo, col0,
                                        accumulate);
}

// out (K x N) = A' B summed over all rows (fixed chunk order).
void gemm_tn(hipStream_t s, GnnWorkspace* w, int kt, int nt, const uint16_t* A, int lda, const uint16_t* B, int ldb,
             float* out) {
  const int chunks = (w->m_pad + kChunk - 1) / kChunk;
  k_mfma_gemm_tn<<<chunks, 256, 0, s>>>(w->m_pad, kt, nt, A, lda, B, ldb, w->part);
  const int n = 32 * kt * 32 * nt;
  k_sum_chunks<<<(n + 255) / 256, 256, 0, s>>>(chunks, n, w->part, out);
}

void col_sums(hipStream_t s, GnnWorkspace* w, int rows, int cols, const uint16_t* a16, const float* a32, int lda,
              float* out) {
  const int chunks = (rows + kChunk - 1) / kChunk;
  if (chunks <= 0) { (void)hipMemsetAsync(out, 0, cols * sizeof(float), s); return; }
  k_col_partials<<<chunks, 128, 0, s>>>(rows, cols, a16, a32, lda, w->part);
  k_sum_chunks<<<(cols + 255) / 256, 256, 0, s>>>(chunks, cols, w->part, out);
}

}  // namespace

bool gnn_forward(hipStream_t s, GnnWorkspace* w, const GnnGraph& g, const float* params, float* theta) {
  if (!w || g.nodes != w->nodes) return false;
  if (g.nodes == 0) return true;
  const Layout L = layout();
  k_gnn_weights<<<dim3((2 * kH * kH + 255) / 256, kGnnLayers), 256, 0, s>>>(params, L, w->wm, w->wu, w->wmt, w->wut);
  k_gnn_degree<<<(g.nodes + 255) / 256, 256, 0, s>>>(g, w->deg);
  k_gnn_input<<<g.nodes, kH, 0, s>>>(g.nodes, g.node_feat, params + L.w_in, params + L.b_in, w->ha[0]);
  for (int l = 0; l < kGnnLayers; ++l) {
    gemm(s, w, kH, 2, w->ha[l], 2 * kH, w->wm + (size_t)l * kH * kH, params + L.b_m[l], false, w->y[l], nullptr, kH,
         0, false);
    k_gnn_aggregate<<<g.nodes, kH, 0, s>>>(g, w->deg, w->y[l], params + L.a_g[l], params + L.c_g[l], w->ha[l]);
    uint16_t* out = (l + 1 < kGnnLayers) ? w->ha[l + 1] : w->h3;
    const int ldo = (l + 1 < kGnnLayers) ? 2 * kH : kH;
    gemm(s, w, 2 * kH, 2, w->ha[l], 2 * kH, w->wu + (size_t)l * 2 * kH * kH, params + L.b_u[l], true, out, nullptr,
         ldo, 0, false);
  }
  k_gnn_output<<<(g.nodes * kLpTerms + 255) / 256, 256, 0, s>>>(g.nodes, w->h3, params + L.w_out, params + L.b_out,
                                                               theta);
  return hipGetLastError() == hipSuccess;
}

bool gnn_loss_backward(hipStream_t s, GnnWorkspace* w, const GnnGraph& g, const float* params, const float* theta,
                       const float* qc, float inv_count, float* loss_out, float* grad) {
  if (!w || !w->training || g.nodes != w->nodes) return false;
  const Layout L = layout();
  const int nodes = g.nodes, M = w->m_pad;
  if (hipMemsetAsync(grad, 0, kGnnParams * sizeof(float), s) != hipSuccess) return false;
  k_gnn_loss<<<(nodes + 255) / 256, 256, 0, s>>>(nodes, theta, qc, inv_count, w->lpart, w->dtheta);
  // loss = sum of lpart (each already scaled by inv_count), fixed order
  col_sums(s, w, nodes, 1, nullptr, w->lpart, 1, loss_out);
  // Output layer.
  {
    const int chunks = (nodes + kChunk - 1) / kChunk;
    const int n = kH * kLpTerms + kLpTerms;
    k_gnn_output_wgrad<<<chunks, 192, 0, s>>>(nodes, w->h3, w->dtheta, w->part);
    k_sum_chunks<<<(n + 255) / 256, 256, 0, s>>>(chunks, n, w->part, grad + L.w_out);  // W_out then b_out
  }
  k_gnn_output_back<<<M, kH, 0, s>>>(nodes, M, w->dtheta, params + L.w_out, w->h3, w->dz);
  for (int l = kGnnLayers - 1; l >= 0; --l) {
    // h_l = relu([h_{l-1} | agg_l] W_u + b_u): dW_u = HA' dz, db_u = sum dz.
    gemm_tn(s, w, 4, 2, w->ha[l], 2 * kH, w->dz, kH, grad + L.w_u[l]);
    col_sums(s, w, nodes, kH, w->dz, nullptr, kH, grad + L.b_u[l]);
    // dHA = dz W_u' (FP32, 128 columns).
    gemm(s, w, kH, 4, w->dz, kH, w->wut + (size_t)l * 2 * kH * kH, nullptr, false, nullptr, w->dha, 2 * kH, 0, false);
    // Aggregation: dy and the gate parameter gradients.
    k_gnn_aggregate_back<<<M, kH, 0, s>>>(g, M, w->deg, w->y[l], w->dha, params + L.a_g[l], params + L.c_g[l], w->dy,
                                          w->gpart);
    col_sums(s, w, nodes, kLpEdgeFeatures + 1, nullptr, w->gpart, kLpEdgeFeatures + 1, grad + L.a_g[l]);  // a_g, c_g
    // y = h_{l-1} W_m + b_m: dW_m = h_{l-1}' dy, db_m = sum dy, dh_{l-1} += dy W_m'.
    gemm_tn(s, w, 2, 2, w->ha[l], 2 * kH, w->dy, kH, grad + L.w_m[l]);
    col_sums(s, w, nodes, kH, w->dy, nullptr, kH, grad + L.b_m[l]);
    gemm(s, w, kH, 2, w->dy, kH, w->wmt + (size_t)l * kH * kH, nullptr, false, nullptr, w->dha, 2 * kH, 0, true);
    if (l > 0) k_gnn_relu_back<<<M, kH, 0, s>>>(nodes, M, w->ha[l], w->dha, w->dz);
  }
  // Input layer (FP32).
  {
    const int chunks = (nodes + kChunk - 1) / kChunk;
    const int n = (kLpNodeFeatures + 1) * kH;
    k_gnn_input_back<<<chunks, kH, 0, s>>>(nodes, g.node_feat, w->ha[0], w->dha, w->part);
    k_sum_chunks<<<(n + 255) / 256, 256, 0, s>>>(chunks, n, w->part, grad + L.w_in);  // W_in then b_in
  }
  return hipGetLastError() == hipSuccess;
}

void launch_adam(hipStream_t s, int32_t n, float lr, int32_t step, const float* grad, float* params, float* m,
                 float* v) {
  const float c1 = 1.0f - powf(0.9f, (float)step), c2 = 1.0f - powf(0.999f, (float)step);
  k_adam<<<(n + 255) / 256, 256, 0, s>>>(n, lr, c1, c2, grad, params, m, v);
}

}  // namespace mlba

// =================================================================== C ABI

#include "../../include/maplab_ba_learned.h"

namespace {

std::string g_gnn_error;

int gnn_fail(int code, const char* msg) {
  g_gnn_error = msg;
  return code;
}

// A dataset copied to the device.
struct DeviceDataset {
  std::vector<void*> allocs;
  mlba::GnnGraph g{};
  float* qc = nullptr;
  float inv_count = 0.0f;
  ~DeviceDataset() {
    for (void* p : allocs) (void)hipFree(p);
  }
  // A device copy of n elements of src (zeros when src is NULL).
  template <typename T>
  T* put(const T* src, size_t n) {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return nullptr;
    allocs.push_back(p);
    if (!src) {
      if (hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return nullptr;
    } else if (n && hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
      return nullptr;
    }
    return static_cast<T*>(p);
  }
  int load(const mlba_gnn_dataset* d) {
    if (!d || d->nodes < 0 || d->slots < 0 || !d->row_ptr || (d->slots && (!d->col || !d->rev || !d->edge_features)) ||
        (d->nodes && (!d->node_features || !d->qc)))
      return gnn_fail(MLBA_EINVAL, "dataset: NULL array or negative size");
    if (d->row_ptr[0] != 0 || d->row_ptr[d->nodes] != d->slots)
      return gnn_fail(MLBA_EINVAL, "dataset: row_ptr does not span the slots");
    double cnt = 0.0;
    for (int i = 0; i < d->nodes; ++i) {
      if (d->row_ptr[i + 1] < d->row_ptr[i]) return gnn_fail(MLBA_EINVAL, "dataset: row_ptr not monotone");
      for (int e = d->row_ptr[i]; e < d->row_ptr[i + 1]; ++e) {
        const int j = d->col[e], r = d->rev[e];
        if (j < 0 || j >= d->nodes || r < 0 || r >= d->slots || d->col[r] != i || r < d->row_ptr[j] ||
            r >= d->row_ptr[j + 1])
          return gnn_fail(MLBA_EINVAL, "dataset: col / rev out of range or not the reverse slot");
      }
      cnt += d->qc[5 * (size_t)i + 4] != 0.0f;
    }
    inv_count = cnt > 0 ? (float)(1.0 / cnt) : 0.0f;
    g.nodes = d->nodes;
    g.slots = d->slots;
    g.row_ptr = put(d->row_ptr, (size_t)d->nodes + 1);
    g.col = put(d->col, d->slots);
    g.rev = put(d->rev, d->slots);
    g.node_feat = put(d->node_features, (size_t)d->nodes * mlba::kLpNodeFeatures);
    g.edge_feat = put(d->edge_features, (size_t)d->slots * mlba::kLpEdgeFeatures);
    qc = put(d->qc, (size_t)d->nodes * 5);
    if (!g.row_ptr || !g.col || !g.rev || !g.node_feat || !g.edge_feat || !qc)
      return gnn_fail(MLBA_ENOMEM, "dataset: device allocation failed");
    return MLBA_OK;
  }
};

struct Workspace {
  mlba::GnnWorkspace* w = nullptr;
  ~Workspace() { mlba::gnn_workspace_destroy(w); }
};

}  // namespace

extern "C" {

int32_t mlba_gnn_num_params(void) { return mlba::kGnnParams; }

const char* mlba_gnn_last_error(void) { return g_gnn_error.c_str(); }

int32_t mlba_gnn_evaluate(const mlba_gnn_dataset* data, const float* params, float* theta, float* loss,
                          float* grad) {
  if (!params) return gnn_fail(MLBA_EINVAL, "params is NULL");
  DeviceDataset ds;
  int rc = ds.load(data);
  if (rc) return rc;
  const int P = mlba::kGnnParams, n = data->nodes;
  float* dp = ds.put(params, P);
  float* dtheta = ds.put<float>(nullptr, 2 * (size_t)std::max(n, 1));
  float* dgrad = ds.put<float>(nullptr, P);
  float* dloss = ds.put<float>(nullptr, 1);
  if (!dp || !dtheta || !dgrad || !dloss) return gnn_fail(MLBA_ENOMEM, "device allocation failed");
  Workspace ws;
  ws.w = mlba::gnn_workspace_create(n, data->slots, loss || grad);
  if (!ws.w) return gnn_fail(MLBA_ENOMEM, "workspace allocation failed");
  hipStream_t s = nullptr;
  if (!mlba::gnn_forward(s, ws.w, ds.g, dp, dtheta)) return gnn_fail(MLBA_EDEVICE, "forward launch failed");
  if (loss || grad) {
    if (!mlba::gnn_loss_backward(s, ws.w, ds.g, dp, dtheta, ds.qc, ds.inv_count, dloss, dgrad))
      return gnn_fail(MLBA_EDEVICE, "backward launch failed");
  }
  if (theta && n && hipMemcpy(theta, dtheta, sizeof(float) * 2 * n, hipMemcpyDeviceToHost) != hipSuccess)
    return gnn_fail(MLBA_EDEVICE, "copy failed");
  if (loss && hipMemcpy(loss, dloss, sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    return gnn_fail(MLBA_EDEVICE, "copy failed");
  if (grad && hipMemcpy(grad, dgrad, sizeof(float) * P, hipMemcpyDeviceToHost) != hipSuccess)
    return gnn_fail(MLBA_EDEVICE, "copy failed");
  return MLBA_OK;
}

int32_t mlba_gnn_train(const mlba_gnn_dataset* data, int32_t epochs, float learning_rate, float* params_inout,
                       float* loss_log) {
  if (!params_inout || epochs < 0 || !(learning_rate > 0.0f)) return gnn_fail(MLBA_EINVAL, "params, epochs or rate");
  DeviceDataset ds;
  int rc = ds.load(data);
  if (rc) return rc;
  const int P = mlba::kGnnParams, n = data->nodes;
  float* dp = ds.put(params_inout, P);
  std::vector<float> zeros(P, 0.0f);
  float* m = ds.put(zeros.data(), P);
  float* v = ds.put(zeros.data(), P);
  float* dtheta = ds.put<float>(nullptr, 2 * (size_t)std::max(n, 1));
  float* dgrad = ds.put<float>(nullptr, P);
  float* dlosses = ds.put<float>(nullptr, (size_t)std::max(epochs, 1));
  if (!dp || !m || !v || !dtheta || !dgrad || !dlosses) return gnn_fail(MLBA_ENOMEM, "device allocation failed");
  Workspace ws;
  ws.w = mlba::gnn_workspace_create(n, data->slots, true);
  if (!ws.w) return gnn_fail(MLBA_ENOMEM, "workspace allocation failed");
  hipStream_t s = nullptr;
  for (int ep = 0; ep < epochs; ++ep) {
    if (!mlba::gnn_forward(s, ws.w, ds.g, dp, dtheta) ||
        !mlba::gnn_loss_backward(s, ws.w, ds.g, dp, dtheta, ds.qc, ds.inv_count, dlosses + ep, dgrad))
      return gnn_fail(MLBA_EDEVICE, "training launch failed");
    mlba::launch_adam(s, P, learning_rate, ep + 1, dgrad, dp, m, v);
  }
  if (hipDeviceSynchronize() != hipSuccess) return gnn_fail(MLBA_EDEVICE, "training failed");
  if (hipMemcpy(params_inout, dp, sizeof(float) * P, hipMemcpyDeviceToHost) != hipSuccess)
    return gnn_fail(MLBA_EDEVICE, "copy failed");
  if (loss_log && epochs &&
      hipMemcpy(loss_log, dlosses, sizeof(float) * epochs, hipMemcpyDeviceToHost) != hipSuccess)
    return gnn_fail(MLBA_EDEVICE, "copy failed");
  return MLBA_OK;
}

}  // extern "C"
