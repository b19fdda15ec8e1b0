// A batch slot: everything one in-flight batched launch needs.
//
// One device block holds the batch block of mrg8_device.hpp (header,
// segments, tile prefix sums, tile -> segment map), sized for `cap` segments
// and `max_tiles` tiles; a coherent pinned host block of the same size
// stages it.  A run copies the used prefix to the device with a small kernel
// (k_stage) when it changed since the last run, on the launch stream, and
// launches the kernel directly or
// replays a captured graph holding it; the kernel reads the counts and table
// offsets from the header, so a graph is captured once per launch
// configuration.
//
// Completion of a run is an event recorded behind it on its stream.  A
// tracked run records it at once (the serving loop polls it); otherwise it
// is recorded only when something needs it -- a copy issued on another
// stream, a wait, a re-capture, release -- since every event record is one
// more packet the stream must process between kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "mrg8_internal.hpp"

namespace mrg8 {

struct LaunchConfig {
  int block_threads = 256;
  int blocks_per_cu = 1;  // 0: one tile per wave (k_batch_tiles); > 0: persistent grid (k_batch);
                          // < 0: persistent grid taking tiles in address order (k_batch_queue)
  uint64_t tile = 4096;
};

// Bytes of a batch block with `count` segments and `ntiles` tile-map entries.
size_t batch_block_bytes(uint32_t count, uint64_t ntiles);
// Lays out the batch block at `mem` whose `count` segments are already
// written after its header: prefix sums of tiles of `tile` outputs, the
// tile -> segment map (at most `max_tiles` entries) and the header.  Sets
// *ntiles to the tile count and returns the bytes in use.
size_t layout_batch_block(uint8_t* mem, uint32_t count, uint64_t tile, uint64_t max_tiles, uint64_t* ntiles);

class BatchSlot {
 public:
  BatchSlot() = default;
  BatchSlot(const BatchSlot&) = delete;
  BatchSlot& operator=(const BatchSlot&) = delete;
  ~BatchSlot() { release(); }

  // Allocates room for `cap` segments and `max_tiles` tiles; `states` is the
  // resident state table the segments may refer to (may be null).
  int init(DeviceContext* ctx, uint32_t cap, uint64_t max_tiles, const uint32_t* states);
  void release();

  // Host staging (valid to write only when idle()).
  BatchSegment* segs() { return h_segs_; }
  uint32_t capacity() const { return cap_; }
  uint64_t max_tiles() const { return max_tiles_; }
  // Waits until the previous run no longer reads the staging block.
  int wait_staging();
  // Finalises `count` staged segments for tiles of `tile` outputs (prefix
  // sums, tile -> segment map).  Returns the number of tiles; the caller
  // keeps it <= max_tiles.
  uint64_t finalise(uint32_t count, uint64_t tile);
  uint64_t staged_tiles() const { return ntiles_; }
  uint32_t staged_count() const { return count_; }

  // Copies the staged block (when it changed) and launches on `stream`:
  // a plain kernel launch, or with `graph` a replay of the captured graph;
  // with `track` the run's completion event is recorded at once.
  int run(const LaunchConfig& cfg, hipStream_t stream, bool graph, bool track);
  // The last run finished (non-blocking query).
  bool done();
  // Waits for the last run.  If the stream it ran on no longer accepts
  // work, waits for the whole device.
  int wait_done();

 private:
  int capture(const LaunchConfig& cfg);
  // Records the completion event behind the last run if it is not there yet.
  hipError_t mark_last_run();
  BatchArgs args(const LaunchConfig& cfg) const;
  hipError_t launch(const LaunchConfig& cfg, hipStream_t stream) const;

  DeviceContext* ctx_ = nullptr;
  uint32_t cap_ = 0;
  uint64_t max_tiles_ = 0;
  const uint32_t* states_ = nullptr;
  uint8_t* dmem_ = nullptr;
  uint8_t* hmem_ = nullptr;
  size_t used_bytes_ = 0;  // staged prefix to copy
  BatchSegment* h_segs_ = nullptr;
  hipEvent_t staged_ = nullptr, run_done_ = nullptr;
  bool copy_pending_ = false, run_pending_ = false, dirty_ = false;
  hipStream_t last_stream_ = nullptr;  // stream of the last run
  bool last_marked_ = true;            // run_done_ is recorded behind the last run
  bool last_queue_ = false;            // the last run was an address-order grid (work counters)
  uint32_t count_ = 0;
  uint64_t ntiles_ = 0;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
  LaunchConfig captured_{};
  uint64_t captured_ntiles_ = 0;
  bool have_graph_ = false;
};

}  // namespace mrg8
