// ba_tiles.cc -- host construction of the reduced-camera-system work
// decomposition (factor blocks, tile visits, batches, pair lists, segments).
#include "ba_tiles.h"

#include <algorithm>
#include <limits>

namespace mlba {

namespace {

struct Run {
  int32_t band, g, n;  // band, first factor block, blocks
};

}  // namespace

bool build_schur_tiles(int32_t Pvar, int32_t L, const int32_t* inc_ptr, const int32_t* inc,
                       const int32_t* obs_landmark, SchurTiles* out, std::string* err) {
  SchurTiles& T = *out;
  T = SchurTiles();
  const int64_t n_inc = Pvar > 0 ? inc_ptr[Pvar] : 0;
  auto lm_of = [&](int32_t code) { return code >= 0 ? obs_landmark[code] : -1 - code; };
  // Incidences of each landmark in row order (the row lists are row-major).
  std::vector<int32_t> inc_row(n_inc), lm_ptr(L + 1, 0), lm_list(n_inc);
  for (int i = 0; i < Pvar; ++i)
    for (int k = inc_ptr[i]; k < inc_ptr[i + 1]; ++k) inc_row[k] = i;
  for (int64_t k = 0; k < n_inc; ++k) lm_ptr[lm_of(inc[k]) + 1]++;
  for (int l = 0; l < L; ++l) lm_ptr[l + 1] += lm_ptr[l];
  {
    std::vector<int32_t> fill(lm_ptr.begin(), lm_ptr.end() - 1);
    for (int64_t k = 0; k < n_inc; ++k) lm_list[fill[lm_of(inc[k])]++] = (int32_t)k;
  }
  // Factor blocks: the incidences of one landmark in one row.
  std::vector<int32_t>& lm_g = T.lm_g_ptr;
  lm_g.assign(L + 1, 0);
  T.g_row.reserve(n_inc);
  T.g_inc.reserve(n_inc);
  for (int l = 0; l < L; ++l) {
    int last = -1;
    for (int m = lm_ptr[l]; m < lm_ptr[l + 1]; ++m) {
      const int k = lm_list[m];
      if (inc_row[k] != last) {
        last = inc_row[k];
        T.g_row.push_back(last);
        T.g_inc_ptr.push_back((int32_t)T.g_inc.size());
      }
      T.g_inc.push_back(inc[k]);
    }
    lm_g[l + 1] = (int32_t)T.g_row.size();
  }
  T.g_inc_ptr.push_back((int32_t)T.g_inc.size());
  // An anchor incidence is alone in its row: the anchor vertex's own
  // observations of the landmark do not couple its pose.
  for (size_t g = 0; g + 1 < T.g_inc_ptr.size(); ++g) {
    const int b = T.g_inc_ptr[g], e = T.g_inc_ptr[g + 1];
    for (int k = b; k < e; ++k)
      if (T.g_inc[k] < 0 && e - b != 1) {
        *err = "internal: an anchor incidence shares its row with observations";
        return false;
      }
  }
  const int64_t nb = (Pvar + kTileBand - 1) / kTileBand;
  T.nbands = (int32_t)nb;
  std::vector<Run> runs;
  auto landmark_runs = [&](int l) {
    runs.clear();
    for (int g = lm_g[l]; g < lm_g[l + 1]; ++g) {
      const int band = T.g_row[g] / kTileBand;
      if (runs.empty() || runs.back().band != band) runs.push_back(Run{band, g, 0});
      runs.back().n++;
    }
  };
  // Visits (landmark, first block in I, first in J, blocks in I, in J),
  // grouped by tile in (I, J) order, landmarks ascending in a tile (a
  // landmark visits a tile at most once; only visited tiles exist).
  struct Visit {
    int64_t key;  // I * nb + J
    int32_t l, gi, gj, ni, nj;
  };
  std::vector<Visit> visits;
  for (int l = 0; l < L; ++l) {
    landmark_runs(l);
    for (size_t a = 0; a < runs.size(); ++a)
      for (size_t b = a; b < runs.size(); ++b) {
        const bool diag = a == b;
        visits.push_back(Visit{runs[a].band * nb + runs[b].band, l, runs[a].g, runs[b].g, runs[a].n,
                               diag ? 0 : runs[b].n});
      }
  }
  {
    // Stable LSD radix sort by key, 16 bits per pass (keys < nb^2).
    std::vector<Visit> tmp(visits.size());
    std::vector<int64_t> count(65537);
    for (int shift = 0; shift < 64 && (nb * nb - 1) >> shift; shift += 16) {
      std::fill(count.begin(), count.end(), 0);
      for (const Visit& v : visits) count[((v.key >> shift) & 65535) + 1]++;
      for (int k = 0; k < 65536; ++k) count[k + 1] += count[k];
      for (const Visit& v : visits) tmp[count[(v.key >> shift) & 65535]++] = v;
      visits.swap(tmp);
    }
  }
  const int64_t n_visits = (int64_t)visits.size();
  T.visits = n_visits;
  {
    int64_t staged = 0, npairs = 0;
    for (const Visit& v : visits) {
      staged += v.ni + v.nj;
      npairs += v.nj ? (int64_t)v.ni * v.nj : (int64_t)v.ni * (v.ni + 1) / 2;
    }
    T.slots.reserve(2 * staged);
    T.pairs.reserve(npairs + npairs / 64 + 64);
  }
  auto is_anchor = [&](int g) { return T.g_inc[T.g_inc_ptr[g]] < 0; };
  // Batches (LDS-sized groups of visits) with their slot lists and pair
  // lists by block, and segments (runs of batches of one tile) with their
  // thread parts.
  std::vector<uint32_t> bp_block;
  std::vector<uint16_t> bp_pair;
  int32_t cnt[kTileBlocks + 1];
  int64_t seg_cnt[kTileBlocks];
  int parts[kTileBlocks];
  auto close_segment = [&](int32_t b0, int32_t b1, int64_t I, int64_t J) {
    T.segs.push_back(b0);
    T.segs.push_back(b1);
    T.segs.push_back((int32_t)I);
    T.segs.push_back((int32_t)J);
    for (int k = 0; k < kTileBlocks; ++k) seg_cnt[k] = 0;
    for (int32_t b = b0; b < b1; ++b) {
      const uint16_t* o = &T.batch_off[(size_t)b * (kTileBlocks + 1)];
      for (int k = 0; k < kTileBlocks; ++k) seg_cnt[k] += o[k + 1] - o[k];
    }
    // One part per block with products, then the remaining threads one at a
    // time to the block with the most products per part (lowest index on
    // ties), at most 255 parts a block.
    int used = 0;
    for (int k = 0; k < kTileBlocks; ++k) {
      parts[k] = seg_cnt[k] > 0 ? 1 : 0;
      used += parts[k];
    }
    auto lower = [&](int a, int b) {  // a ranks below b
      const int64_t x = seg_cnt[a] * parts[b], y = seg_cnt[b] * parts[a];
      return x != y ? x < y : a > b;
    };
    int heap[kTileBlocks], nh = 0;
    for (int k = 0; k < kTileBlocks; ++k)
      if (parts[k]) heap[nh++] = k;
    std::make_heap(heap, heap + nh, lower);
    while (used < kTileBlocks && nh > 0) {
      std::pop_heap(heap, heap + nh, lower);
      const int best = heap[nh - 1];
      if (seg_cnt[best] <= parts[best]) break;
      parts[best]++;
      used++;
      if (parts[best] >= 255) --nh;
      else std::push_heap(heap, heap + nh, lower);
    }
    const size_t base = T.seg_thread.size();
    T.seg_thread.resize(base + kTileBlocks, 0u);
    T.seg_block.resize(base + kTileBlocks, 0);
    int t = 0;
    for (int k = 0; k < kTileBlocks; ++k) {
      if (!parts[k]) continue;
      T.seg_block[base + k] = (uint16_t)(t | (parts[k] << 8));
      for (int p = 0; p < parts[k]; ++p)
        T.seg_thread[base + t++] = (uint32_t)k | ((uint32_t)p << 8) | ((uint32_t)parts[k] << 16);
    }
  };
  for (int64_t tv = 0; tv < n_visits;) {
    const int64_t key = visits[tv].key, I = key / nb, J = key % nb;
    int64_t v = tv, ve = tv;
    while (ve < n_visits && visits[ve].key == key) ++ve;
    tv = ve;
    const bool diag = I == J;
    const int32_t seg_first = (int32_t)(T.segs.size() / 4);
    int32_t seg_batch = (int32_t)(T.batches.size() / 4);
    int64_t seg_slots = 0;
    while (v < ve) {
      const int64_t bv = v;
      int slots = 0;
      while (v < ve) {
        const int c = visits[v].ni + visits[v].nj;
        if (slots + c > kBatchSlots) break;
        slots += c;
        ++v;
      }
      const size_t slot0 = T.slots.size() / 2;
      if (slot0 + slots > (size_t)std::numeric_limits<int32_t>::max()) {
        *err = "reduced camera system has too many staged factor blocks";
        return false;
      }
      bp_block.clear();
      bp_pair.clear();
      int s = 0;
      for (int64_t u = bv; u < v; ++u) {
        const Visit& vr = visits[u];
        const int si = s, sj = diag ? s : s + vr.ni;
        for (int a = 0; a < vr.ni; ++a) {
          T.slots.push_back(vr.gi + a);
          T.slots.push_back(vr.l | (int32_t)(is_anchor(vr.gi + a) ? kSlotAnchor : 0u));
        }
        for (int b = 0; b < vr.nj; ++b) {
          T.slots.push_back(vr.gj + b);
          T.slots.push_back(vr.l | (int32_t)(is_anchor(vr.gj + b) ? kSlotAnchor : 0u));
        }
        const int nj = diag ? vr.ni : vr.nj;
        const int gj = diag ? vr.gi : vr.gj;
        for (int a = 0; a < vr.ni; ++a)
          for (int b = diag ? a : 0; b < nj; ++b) {
            bp_block.push_back((uint32_t)((T.g_row[vr.gi + a] - I * kTileBand) * kTileBand +
                                          (T.g_row[gj + b] - J * kTileBand)));
            bp_pair.push_back((uint16_t)((si + a) | ((sj + b) << 8)));
          }
        s += vr.ni + vr.nj;
      }
      // A visit has at most 16 x 16 pairs over 32 slots (8 per slot), or
      // 16 x 17 / 2 over 16 on a diagonal tile (8.5 per slot).
      if (bp_pair.size() > (size_t)kBatchPairs) {
        *err = "internal: a Schur batch has more pairs than its LDS holds";
        return false;
      }
      // Stable counting sort of the batch's pairs by block.
      for (int k = 0; k <= kTileBlocks; ++k) cnt[k] = 0;
      for (uint32_t b : bp_block) cnt[b + 1]++;
      for (int k = 0; k < kTileBlocks; ++k) cnt[k + 1] += cnt[k];
      if (T.pairs.size() % 2) T.pairs.push_back(0);  // batches start on a 32-bit word
      const size_t pb = T.pairs.size();
      if (pb + bp_pair.size() > (size_t)std::numeric_limits<int32_t>::max()) {
        *err = "reduced camera system has too many factor pairs";
        return false;
      }
      for (int k = 0; k <= kTileBlocks; ++k) T.batch_off.push_back((uint16_t)cnt[k]);
      T.pairs.resize(pb + bp_pair.size());
      for (size_t k = 0; k < bp_pair.size(); ++k) T.pairs[pb + cnt[bp_block[k]]++] = bp_pair[k];
      T.batches.push_back((int32_t)slot0);
      T.batches.push_back(slots);
      T.batches.push_back((int32_t)pb);
      T.batches.push_back((int32_t)bp_pair.size());
      T.staged += slots;
      seg_slots += slots;
      const int32_t nbatch = (int32_t)(T.batches.size() / 4);
      if (v == ve || seg_slots >= kSegmentSlots || nbatch - seg_batch >= kSegmentBatches) {
        close_segment(seg_batch, nbatch, I, J);
        seg_batch = nbatch;
        seg_slots = 0;
      }
    }
    T.tile_key.push_back(key);
    T.tile_seg.push_back(seg_first);
    T.tile_seg.push_back((int32_t)(T.segs.size() / 4));
  }
  return true;
}

}  // namespace mlba
