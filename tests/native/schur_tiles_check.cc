// Checks the reduced camera system's work decomposition (ba_tiles.h) on the
// CPU: a random problem (landmarks seen from many bands, two incidences of a
// landmark in one row, anchors, rows without anchors) is decomposed, the
// device kernel's staging and per-part products are replayed with bounds
// checks on every slot, pair and part, and S and its rhs are compared with a
// brute-force sum over incidence pairs (Q_x W_y + [x observation, y anchor]
// W_x^T A_l, rhs_x = r0_x - W_x^T Z g). Usage: schur_tiles_check Pvar L seed.
#include <algorithm>
#include <cassert>
#include <cmath>
#include <cstdio>
#include <random>
#include "ba_tiles.h"
using namespace mlba;
#define CHECK(c) do { if (!(c)) { printf("CHECK failed line %d: %s\n", __LINE__, #c); return 1; } } while (0)

int main(int argc, char** argv) {
  const int Pvar = argc > 1 ? atoi(argv[1]) : 75, L = argc > 2 ? atoi(argv[2]) : 3000;
  std::mt19937 rng(argc > 3 ? (unsigned)atoi(argv[3]) : 7u);
  std::uniform_real_distribution<double> U(-1, 1);
  // incidences: per landmark rows; observations have codes 0..N-1
  std::vector<std::vector<int>> row_inc(Pvar);  // codes per row
  std::vector<int> obs_lm;
  std::vector<int> anchor_row(L, -1);
  for (int l = 0; l < L; ++l) {
    int center = rng() % Pvar, span = (rng() % 10 == 0) ? Pvar : 1 + rng() % 40;
    int k = 1 + rng() % 14;
    std::vector<int> rows;
    int ar = (rng() % 4 == 0) ? -1 : std::max(0, std::min(Pvar - 1, center + (int)(rng() % span) - span / 2));
    anchor_row[l] = ar;
    for (int q = 0; q < k; ++q) {
      int r = std::max(0, std::min(Pvar - 1, center + (int)(rng() % span) - span / 2));
      if (r == ar) continue;
      int reps = (rng() % 7 == 0) ? 2 : 1;
      for (int z = 0; z < reps; ++z) { row_inc[r].push_back((int)obs_lm.size()); obs_lm.push_back(l); }
    }
  }
  for (int l = 0; l < L; ++l) if (anchor_row[l] >= 0) row_inc[anchor_row[l]].push_back(-1 - l);
  std::vector<int32_t> inc_ptr(Pvar + 1, 0), inc;
  for (int i = 0; i < Pvar; ++i) { for (int c : row_inc[i]) inc.push_back(c); inc_ptr[i + 1] = inc.size(); }
  const int N = obs_lm.size(), n_inc = inc.size();
  // data
  std::vector<double> Wx(18 * n_inc), r0x(6 * n_inc), Z(9 * L), g(3 * L), A(18 * L);
  for (auto& x : Wx) x = U(rng);
  for (auto& x : r0x) x = U(rng);
  for (auto& x : g) x = U(rng);
  for (auto& x : A) x = U(rng);
  for (int l = 0; l < L; ++l) {  // SPD Z = B B^T + I
    double B[9]; for (auto& x : B) x = U(rng);
    for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) {
      double s = a == b; for (int c = 0; c < 3; ++c) s += B[a * 3 + c] * B[b * 3 + c];
      Z[9 * l + a * 3 + b] = s; }
  }
  auto lm_of = [&](int code) { return code >= 0 ? obs_lm[code] : -1 - code; };
  std::vector<int> inc_row(n_inc);
  for (int i = 0; i < Pvar; ++i) for (int k = inc_ptr[i]; k < inc_ptr[i + 1]; ++k) inc_row[k] = i;
  // brute force S (upper, dense) and rhs
  std::vector<double> S(36 * Pvar * Pvar, 0.0), rhs(6 * Pvar, 0.0);
  auto Q = [&](int k, double* q) {  // 6x3 [c][a]
    int l = lm_of(inc[k]);
    for (int c = 0; c < 6; ++c) for (int a = 0; a < 3; ++a) {
      double v = inc[k] < 0 ? A[18 * l + a * 6 + c] : 0.0;
      for (int b = 0; b < 3; ++b) v -= Wx[18 * k + b * 6 + c] * Z[9 * l + b * 3 + a];
      q[c * 3 + a] = v; }
  };
  std::vector<std::vector<int>> lm_incs(L);
  for (int k = 0; k < n_inc; ++k) lm_incs[lm_of(inc[k])].push_back(k);
  for (int lx = 0; lx < L; ++lx) for (int x : lm_incs[lx]) for (int y : lm_incs[lx]) {
    if (inc_row[y] < inc_row[x]) continue;
    int l = lm_of(inc[x]), i = inc_row[x], j = inc_row[y];
    double q[18]; Q(x, q);
    double* s = &S[36 * (i * Pvar + j)];
    for (int r = 0; r < 6; ++r) for (int c = 0; c < 6; ++c) {
      double v = 0; for (int a = 0; a < 3; ++a) v += q[r * 3 + a] * Wx[18 * y + a * 6 + c];
      if (inc[x] >= 0 && inc[y] < 0) for (int a = 0; a < 3; ++a) v += Wx[18 * x + a * 6 + r] * A[18 * l + a * 6 + c];
      s[r * 6 + c] += v; }
  }
  for (int x = 0; x < n_inc; ++x) {
    int l = lm_of(inc[x]), i = inc_row[x];
    for (int c = 0; c < 6; ++c) {
      double v = r0x[6 * x + c];
      for (int a = 0; a < 3; ++a) { double zg = 0; for (int b = 0; b < 3; ++b) zg += Z[9 * l + a * 3 + b] * g[3 * l + b];
        v -= Wx[18 * x + a * 6 + c] * zg; }
      rhs[6 * i + c] += v; }
  }
  // tiles
  SchurTiles T; std::string err;
  CHECK(build_schur_tiles(Pvar, L, inc_ptr.data(), inc.data(), obs_lm.data(), &T, &err));
  const int nG = T.g_row.size();
  // merged factor blocks (g_inc codes -> incidence index by a code map)
  std::vector<int> code_k(N + L);  // obs code -> k, anchor -> N + l
  for (int k = 0; k < n_inc; ++k) code_k[inc[k] >= 0 ? inc[k] : N + (-1 - inc[k])] = k;
  std::vector<double> gw(18 * nG, 0.0), gr0(6 * nG, 0.0);
  for (int gg = 0; gg < nG; ++gg) for (int m = T.g_inc_ptr[gg]; m < T.g_inc_ptr[gg + 1]; ++m) {
    int c = T.g_inc[m]; int k = code_k[c >= 0 ? c : N + (-1 - c)];
    CHECK(inc_row[k] == T.g_row[gg]);
    for (int e = 0; e < 18; ++e) gw[18 * gg + e] += Wx[18 * k + e];
    for (int e = 0; e < 6; ++e) gr0[6 * gg + e] += r0x[6 * k + e];
  }
  const int nseg = T.segs.size() / 4, nbatch = T.batches.size() / 4;
  std::vector<double> partial((size_t)nseg * kTilePartial, 0.0);
  for (int sg = 0; sg < nseg; ++sg) {
    const int* S4 = &T.segs[4 * sg];
    const bool diag = S4[2] == S4[3];
    CHECK(S4[0] < S4[1] && S4[1] <= nbatch);
    std::vector<double> acc(kTileBlocks * kPartEntries, 0.0);
    const uint32_t* tw = &T.seg_thread[(size_t)sg * kTileBlocks];
    for (int b = S4[0]; b < S4[1]; ++b) {
      const int* B4 = &T.batches[4 * b];
      const int ns = B4[1];
      CHECK(ns >= 1 && ns <= kBatchSlots);
      static double sW[kBatchSlots][18], sQ[kBatchSlots][18], sR[kBatchSlots][6];
      static int sAnc[kBatchSlots];
      for (int s = 0; s < ns; ++s) {
        const int gg = T.slots[2 * (B4[0] + s)];
        const uint32_t lw = (uint32_t)T.slots[2 * (B4[0] + s) + 1];
        const int l = (int)(lw & ~kSlotAnchor);
        const bool anc = (lw & kSlotAnchor) != 0;
        CHECK(gg >= 0 && gg < nG && l >= 0 && l < L);
        CHECK(anc == (T.g_inc[T.g_inc_ptr[gg]] < 0));
        const int band = T.g_row[gg] / kTileBand;
        CHECK(band == S4[2] || band == S4[3]);
        const double* z = &Z[9 * l];
        double zg[3];
        for (int a = 0; a < 3; ++a) zg[a] = z[a * 3] * g[3 * l] + z[a * 3 + 1] * g[3 * l + 1] + z[a * 3 + 2] * g[3 * l + 2];
        const double* w = &gw[18 * gg];
        for (int a = 0; a < 3; ++a) for (int c = 0; c < 6; ++c) {
          double q = anc ? A[18 * l + a * 6 + c] : 0.0;
          q -= z[a * 3] * w[c] + z[a * 3 + 1] * w[6 + c] + z[a * 3 + 2] * w[12 + c];
          sQ[s][a * 6 + c] = q; sW[s][a * 6 + c] = w[a * 6 + c]; }
        sAnc[s] = anc;
        if (diag) for (int c = 0; c < 6; ++c)
          sR[s][c] = gr0[6 * gg + c] - (w[c] * zg[0] + w[6 + c] * zg[1] + w[12 + c] * zg[2]);
      }
      const uint16_t* off = &T.batch_off[(size_t)b * (kTileBlocks + 1)];
      CHECK(off[0] == 0);
      for (int k = 0; k < kTileBlocks; ++k) CHECK(off[k] <= off[k + 1]);
      for (int tid = 0; tid < kTileBlocks; ++tid) {
        const uint32_t w = tw[tid];
        const int np = seg_thread_parts(w);
        if (!np) continue;
        const int blk = seg_thread_block(w), part = seg_thread_part(w);
        const int bi = blk / kTileBand, bj = blk % kTileBand;
        if (diag) CHECK(bi <= bj);
        for (int p = off[blk] + part; p < off[blk + 1]; p += np) {
          const uint32_t pr = T.pairs[B4[2] + p];
          const int x = pr & 255, y = pr >> 8;
          CHECK(x < ns && y < ns);
          const bool sel = sAnc[y] && !sAnc[x];
          const double* Lt = sel ? sW[x] : sQ[x];
          const double* Rt = sel ? sQ[y] : sW[y];
          for (int a = 0; a < 3; ++a) for (int r = 0; r < 6; ++r) for (int c = 0; c < 6; ++c)
            acc[tid * kPartEntries + r * 6 + c] += Lt[a * 6 + r] * Rt[a * 6 + c];
          if (diag && bi == bj) { CHECK(x == y); for (int c = 0; c < 6; ++c) acc[tid * kPartEntries + 36 + c] += sR[x][c]; }
        }
      }
    }
    // every block with pairs has parts
    for (int b = S4[0]; b < S4[1]; ++b) {
      const uint16_t* off = &T.batch_off[(size_t)b * (kTileBlocks + 1)];
      for (int k = 0; k < kTileBlocks; ++k)
        if (off[k + 1] > off[k]) CHECK((T.seg_block[(size_t)sg * kTileBlocks + k] >> 8) > 0);
    }
    for (int e = 0; e < kTileBlocks * kPartEntries; ++e) partial[(size_t)sg * kTilePartial + e] = acc[e];
  }
  // finalize and compare
  double maxerr = 0, maxv = 0, rerr = 0;
  const int nb = T.nbands;
  auto block_sum = [&](int i, int j, int e) {
    const int64_t key = (int64_t)(i / kTileBand) * nb + j / kTileBand;
    const auto it = std::lower_bound(T.tile_key.begin(), T.tile_key.end(), key);
    if (it == T.tile_key.end() || *it != key) return 0.0;
    const size_t t = (size_t)(it - T.tile_key.begin());
    const int blk = (i % kTileBand) * kTileBand + (j % kTileBand);
    double v = 0;
    for (int sg = T.tile_seg[2 * t]; sg < T.tile_seg[2 * t + 1]; ++sg) {
      const uint16_t bw = T.seg_block[(size_t)sg * kTileBlocks + blk];
      for (int q = 0; q < (bw >> 8); ++q) v += partial[(size_t)sg * kTilePartial + ((bw & 255) + q) * kPartEntries + e];
    }
    return v;
  };
  for (int i = 0; i < Pvar; ++i) for (int j = i; j < Pvar; ++j)
    for (int e = 0; e < 36; ++e) {
      double v = block_sum(i, j, e), ref = S[36 * (i * Pvar + j) + e];
      maxerr = std::max(maxerr, std::fabs(v - ref)); maxv = std::max(maxv, std::fabs(ref));
    }
  for (int i = 0; i < Pvar; ++i)
    for (int c = 0; c < 6; ++c) rerr = std::max(rerr, std::fabs(block_sum(i, i, 36 + c) - rhs[6 * i + c]));
  printf("Pvar %d L %d n_inc %d nG %d visits %lld batches %d segs %d pairs %zu staged %lld\n", Pvar, L, n_inc, nG,
         (long long)T.visits, nbatch, nseg, T.pairs.size(), (long long)T.staged);
  printf("S max abs err %.3e (max |S| %.3e), rhs max abs err %.3e\n", maxerr, maxv, rerr);
  return maxerr < 1e-11 * maxv && rerr < 1e-9 ? 0 : 2;
}
