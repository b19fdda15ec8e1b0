// Internal interface between the host library and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "mrg8_math.hpp"

namespace mrg8 {

constexpr int kLaneRows = 256;   // block rows kept on the device: up to 4 outputs per lane
// Device jumps take a 64-bit offset as kJumpPositions digits of kJumpBits
// bits; each non-zero digit is one product with a table matrix.
constexpr int kJumpBits = 8;
constexpr int kJumpPositions = 64 / kJumpBits;
constexpr int kJumpValues = (1 << kJumpBits) - 1;  // non-zero digit values
// Record index of row 0 of the identity, stored after the digit matrices.
constexpr int kIdentityRec = kJumpPositions * kJumpValues * 8;
// Bulk advances take offsets of up to kAdvanceWords 64-bit words; byte
// positions kJumpPositions .. kWidePositions-1 live in a second table built
// on first use.
constexpr int kAdvanceWords = 4;
constexpr int kWidePositions = 8 * kAdvanceWords;

// Largest batch tile (outputs): keeps head + (k + 1) * tile of the tile
// bounds within 64 bits.
constexpr uint64_t kMaxTile = 1ull << 40;

// Per-device constant tables.
struct DeviceTables {
  CoefRecord* lane = nullptr;  // kLaneRows records: row 0 of A^(k+1), support 0xff
  // kJumpPositions * kJumpValues * 8 records: row r of A^(d 2^(kJumpBits k))
  // at ((k * kJumpValues) + (d - 1)) * 8 + r; then the 8 rows of the
  // identity at kIdentityRec + r
  CoefRecord* jump = nullptr;
  // (kWidePositions - kJumpPositions) * kJumpValues * 8 records: row r of
  // A^(d 2^(kJumpBits k)) for k >= kJumpPositions at
  // (((k - kJumpPositions) * kJumpValues) + (d - 1)) * 8 + r; null until the
  // first bulk advance with offsets wider than 64 bits
  CoefRecord* wide = nullptr;
  // Matrix-core fill (k_fill_mfma): kMfmaClasses x 64 lanes x 8 bytes of
  // operand digits and kMfmaClasses x 4 x 4 accumulator start values; per
  // column length kMfmaColumnLen[i], the (kMfmaColumns - 1) x 8 records that
  // give a wave's first column states, and A^(tile outputs) in operand form
  // (kMfmaClasses x 64 x 2 words) with its combine constants (4 x 2): the
  // column states of a wave's next tile.
  const uint32_t* mfma_a = nullptr;
  const int32_t* mfma_c = nullptr;
  const CoefRecord* mfma_fan[2] = {nullptr, nullptr};
  const uint32_t* mfma_next_a[2] = {nullptr, nullptr};
  const uint32_t* mfma_next_k[2] = {nullptr, nullptr};
};

// Matrix-core fill.  A tile is kMfmaColumns columns of `column` consecutive
// outputs (kMfmaColumnLen: 64 or 128); one step of a wave makes 16
// consecutive outputs of each column with int8 MFMAs (see k_fill_mfma).  The
// shape value outputs_per_lane = column / 4 (16 or 32, the outputs each lane
// makes per tile) selects this fill.
constexpr int kMfmaColumns = 16;
constexpr int kMfmaColumnLen[2] = {64, 128};
constexpr int kMfmaClasses = 4;   // byte digits of a coefficient
// Word fills of at least this many outputs take the matrix-core fill by
// default (measured on MI355X below).
constexpr uint64_t kMfmaDefaultMinWords = 1ull << 22;
// Bias that keeps the combine's partial sums non-negative: |partial| <= 2^27 + 2^19.
constexpr uint32_t kMfmaBeta = (1u << 27) + (1u << 19);

// Wave-uniform starting point of a launch.
struct StreamStart {
  uint32_t state[8];  // newest word first
  uint64_t offset;    // jump applied before output 0
};

// One fill launch over a single stream: out[i] is stream position
// start.offset + i.  Wave 0 produces outputs [0, head + tile) and wave w >= 1
// outputs [head + w tile, head + (w + 1) tile), so every wave after the first
// starts on a 1 KiB boundary (head: the outputs before the first one).
struct FillArgs {
  void* out;                // element-aligned
  uint64_t n;
  uint64_t tile;            // multiple of 64 * outputs-per-lane
  uint64_t head;            // outputs before the first 1 KiB boundary of out
  StreamStart start;        // the state, and the position (offset) of out[0]
  const CoefRecord* lane;
  const CoefRecord* jump;
  uint32_t xcds;            // XCDs that blocks are dealt to in turn (block b runs on XCD b mod xcds)
  uint32_t xcd_group;       // consecutive blocks of output per XCD run (1: blocks in dispatch order)
};

// A matrix-core fill: `tiles` whole tiles of 16 x `column` outputs at
// out[head + t 16 column] (1 KiB aligned); wave w makes tiles
// w k .. w k + k - 1 (k = tiles_per_wave, the last wave fewer), moving its
// column states from one tile to the next with the device's step tables
// (A^(16 column) in operand form).  Up to two short runs (the outputs
// before the first tile and after the last) are made by extra waves with the
// vector kernel's loop.
struct MfmaFillArgs {
  void* out;
  uint64_t head;            // outputs before the first tile
  uint64_t tiles;
  uint64_t waves;           // ceil(tiles / tiles_per_wave)
  uint32_t tiles_per_wave;
  uint32_t column;          // outputs per column: 64 or 128
  uint32_t extra;           // 0..2 short runs
  uint32_t waves_per_block; // 4, 8 or 16
  uint64_t extra_first[2];  // first output of each short run
  uint64_t extra_count[2];
  StreamStart start;
  const CoefRecord* lane;
  const CoefRecord* jump;
  const uint32_t* mfma_a;
  const int32_t* mfma_c;
  const CoefRecord* mfma_fan;   // this column length's tables
  const uint32_t* mfma_next_a;  // kMfmaClasses x 64 x 2 words
  const uint32_t* mfma_next_k;  // 4 x 2
};

// Waves of a fill of n >= 1 outputs.
__host__ __device__ inline uint64_t fill_waves(uint64_t n, uint64_t head, uint64_t tile) {
  return n > head ? (n - head + tile - 1) / tile : 1;
}

// A segment of the batched kernel: `count` outputs whose first one is
// `offset` positions after the state of the segment's stream, written to
// `out` in `format`.  The state is `state` itself when `stream` is
// kInlineState, else word block `stream` of the resident state table.
constexpr uint32_t kInlineState = 0xffffffffu;

struct BatchSegment {
  uint32_t state[8];
  uint64_t offset;
  uint64_t count;
  void* out;
  int32_t format;
  uint32_t stream;
};
static_assert(sizeof(BatchSegment) == 64, "segment layout");

// Tiles of a batch segment.  The head is the run of outputs before the first
// tile-aligned address of `out` (the alignment is the largest power of two
// dividing tile * element size, e.g. 16 KiB for 2048 doubles); tile 0 covers
// the head plus `tile` more outputs and tile k >= 1 covers [head + k tile,
// head + (k + 1) tile), so every later tile starts on that boundary and
// needs no alignment peel.  Measured on MI355X with 4096 uneven float64
// segments packed back to back (one 2048-double tile per wave): 2.79-2.82 ms
// per launch against 2.85-2.87 ms with 128-byte alignment; a head as a tile
// of its own was slower (2.86-2.88 ms).  Shared by the host (tile map) and
// the kernel (tile bounds).
__host__ __device__ inline uint64_t tile_align_bytes(int32_t format, uint64_t tile) {
  const uint64_t b = tile * (format == 2 ? 4u : 8u);
  return b ? b & (~b + 1) : 128u;  // lowest set bit: every tile start stays aligned
}

__host__ __device__ inline uint64_t segment_head(const void* out, int32_t format, uint64_t count, uint64_t tile) {
  const uint64_t esz = format == 2 ? 4u : 8u;
  const uint64_t a = tile_align_bytes(format, tile);
  const uint64_t h = ((a - (reinterpret_cast<uintptr_t>(out) & (a - 1))) & (a - 1)) / esz;
  return h < count ? h : count;
}

__host__ __device__ inline uint64_t segment_tiles(const void* out, int32_t format, uint64_t count, uint64_t tile) {
  if (count == 0) return 0;
  const uint64_t h = segment_head(out, format, count, tile);
  return count == h ? 1 : (count - h + tile - 1) / tile;
}

// Outputs [*first, *end) of tile k of a segment of `count` outputs with head h.
__host__ __device__ inline void tile_bounds(uint64_t h, uint64_t count, uint64_t tile, uint64_t k, uint64_t* first,
                                            uint64_t* end) {
  *first = k ? h + k * tile : 0;
  *end = count < h + (k + 1) * tile ? count : h + (k + 1) * tile;
}

// A batch block, copied to the device in one piece per run:
//   BatchHeader | BatchSegment[nsegs] | uint64 tile_begin[nsegs + 1] |
//   uint32 tile_seg[ntiles]
// (tile_begin: prefix sum of tiles per segment; tile_seg: segment of each
// tile).  The header gives the counts and where the two tables start, so the
// launch reads everything through one fixed pointer and the graph holding it
// never needs re-capturing.
struct BatchHeader {
  uint32_t nsegs;
  uint32_t reserved;
  uint64_t ntiles;
  uint64_t tile_begin_off;  // bytes from the block start
  uint64_t tile_seg_off;
  uint64_t pad0[12];
  // Work counters of k_batch_queue: the next group of tiles to hand out and
  // the blocks that have finished.  Zero when the block is staged; the last
  // block of every launch sets them back to zero.  They have a 128-byte
  // cache line of their own: every block's atomic adds on them would
  // otherwise contend with the reads of the header and of the first
  // segments (a single 2^28-double segment: 1.46 ms with the counters
  // beside the header, 0.35 ms on a line of their own).
  uint32_t queue_next;
  uint32_t queue_done;
  uint64_t pad1[15];
};
static_assert(sizeof(BatchHeader) == 256 && offsetof(BatchHeader, queue_next) == 128, "header layout");

struct BatchArgs {
  const uint8_t* block;
  uint64_t tile;               // outputs per tile
  const uint32_t* states;      // resident per-stream states (8 words each), may be null
  const CoefRecord* lane;
  const CoefRecord* jump;
  uint32_t xcds;               // as in FillArgs
  uint32_t xcd_group;
};

// A bulk advance: out[8i..8i+7] = in[8i..8i+7] jumped ahead by the
// little-endian `words`-word offset offsets[words i .. words i + words - 1].
struct AdvanceArgs {
  const uint32_t* in;
  const uint64_t* offsets;
  uint32_t* out;
  uint64_t n;
  uint32_t words;  // 1..kAdvanceWords
  const CoefRecord* jump;
  const CoefRecord* wide;  // needed when words > 1
};

// Launchers (mrg8_device.hip).  All are asynchronous on `stream`.
hipError_t launch_fill(const FillArgs& a, int format, int per_lane, int block_threads, hipStream_t stream);
// Blocks of a.waves_per_block waves.
hipError_t launch_fill_mfma(const MfmaFillArgs& a, int format, hipStream_t stream);
hipError_t launch_batch(const BatchArgs& a, int grid, int block_threads, hipStream_t stream);
// Persistent grid of `grid` blocks taking groups of (waves per block)
// consecutive tiles, in address order, from the counters in the batch
// block's header.  Launches of one batch block must not overlap.
hipError_t launch_batch_queue(const BatchArgs& a, int grid, int block_threads, hipStream_t stream);
// One tile per wave: ceil(ntiles / waves per block) blocks.
hipError_t launch_batch_tiles(const BatchArgs& a, uint64_t ntiles, int block_threads, hipStream_t stream);
// Eight lanes per state; a.n * 8 must stay below 2^32.
hipError_t launch_advance(const AdvanceArgs& a, hipStream_t stream);
// Copies `bytes` (a multiple of 4) from coherent pinned host memory `src` to
// device memory `dst`, both 16-byte aligned, with a kernel on `stream`.
hipError_t launch_stage(const void* src, void* dst, size_t bytes, hipStream_t stream);

}  // namespace mrg8
