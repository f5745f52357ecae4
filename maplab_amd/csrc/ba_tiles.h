// ba_tiles.h -- work decomposition of the reduced camera system build.
//
// S = sum over landmarks l of the products of l's factor blocks: a factor
// block is the sum of the incidences of one landmark in one row of S
// (normally one observation, or the landmark's anchor). S is cut into tiles
// of kTileBand x kTileBand blocks (block rows I, block columns J >= I); a
// landmark whose factor blocks fall into bands I and J "visits" tile (I, J)
// and contributes the products of its blocks in I with its blocks in J. A
// tile's visits (in landmark order) are cut into batches whose factor blocks
// fit the LDS of one work-group (the batch's slot list names them), and the
// batches into segments (the work items). In a segment, each of the
// kTileBlocks threads owns one part of one block of the tile: blocks with
// more products get more parts (in proportion to their products in the
// segment), and part p of a block with n parts takes the block's products
// p, p + n, ... of every batch. Each part's sum is written to its own slot
// of the segment's partial; a block of S adds its parts in part order and
// its segments in segment order. Nothing here depends on the device, the
// launch or timing, so S repeats bit for bit.
#ifndef MAPLAB_AMD_BA_TILES_H_
#define MAPLAB_AMD_BA_TILES_H_

#include <stdint.h>

#include <string>
#include <vector>

namespace mlba {

constexpr int kTileBand = 16;                        // block rows (and columns) of a tile
constexpr int kTileBlocks = kTileBand * kTileBand;   // blocks of a tile = threads of a work-group
constexpr int kBatchSlots = 192;                     // factor blocks staged in LDS per batch (<= 256)
constexpr int kSegmentSlots = 5120;                  // staged factor blocks per segment (about)
constexpr int kSegmentBatches = 64;                  // batches per segment, at most
constexpr int kBatchPairs = 2048;                    // pairs per batch, at most
constexpr int kPairWords = kBatchPairs / 2 / kTileBlocks;  // 32-bit pair words a thread stages per batch
constexpr int kPartEntries = 42;                     // a part's block (36) and rhs (6) sums
constexpr int kTilePartial = kTileBlocks * kPartEntries;  // doubles of one segment's partial
constexpr uint32_t kSlotAnchor = 0x80000000u;        // slot word: the block is its landmark's anchor

// Segment thread word: block (bits 0-7), part (8-15), parts of the block
// (16-23); 0 parts: the thread idles. Segment block word: first thread of
// the block (bits 0-7) and its parts (8-15).
constexpr int seg_thread_block(uint32_t w) { return (int)(w & 255u); }
constexpr int seg_thread_part(uint32_t w) { return (int)((w >> 8) & 255u); }
constexpr int seg_thread_parts(uint32_t w) { return (int)((w >> 16) & 255u); }

struct SchurTiles {
  int32_t nbands = 0;
  std::vector<int32_t> g_row;       // row of each factor block (landmark-major, rows ascending)
  std::vector<int32_t> g_inc_ptr;   // incidences summed into each factor block
  std::vector<int32_t> g_inc;       // their codes (observation >= 0, anchor -1 - landmark)
  std::vector<int32_t> lm_g_ptr;    // factor blocks of each landmark (L + 1)
  std::vector<int32_t> slots;       // 2 per staged slot: factor block, landmark | kSlotAnchor
  std::vector<int32_t> batches;     // 4 per batch: first slot, slots, first pair (even), pairs
  std::vector<uint16_t> batch_off;  // kTileBlocks + 1 per batch: pair offsets of each block
  std::vector<uint16_t> pairs;      // staged slots x | y << 8, by block, then visit
  std::vector<int32_t> segs;        // 4 per segment: first batch, end batch, I, J
  std::vector<uint32_t> seg_thread; // kTileBlocks per segment: the thread's part
  std::vector<uint16_t> seg_block;  // kTileBlocks per segment: the block's parts
  std::vector<int64_t> tile_key;    // visited tiles, ascending I * nbands + J
  std::vector<int32_t> tile_seg;    // 2 per visited tile: its segment range
  int64_t visits = 0;               // landmark tile visits
  int64_t staged = 0;               // factor blocks staged over all batches
};

// inc_ptr / inc: incidences of each of the Pvar rows of S (codes as above);
// obs_landmark: landmark of each observation. Returns false (and a reason)
// when a count overflows the decomposition's index widths.
bool build_schur_tiles(int32_t Pvar, int32_t L, const int32_t* inc_ptr, const int32_t* inc,
                       const int32_t* obs_landmark, SchurTiles* out, std::string* err);

}  // namespace mlba

#endif  // MAPLAB_AMD_BA_TILES_H_
