// A batch slot: everything one in-flight persistent launch needs.
//
// Device block (one allocation):
//   BatchHeader | BatchSegment[cap] | uint64 tile_begin[cap+1] |
//   uint32 tile_seg[max_tiles]
// Host side: a pinned staging copy of the same block.  A run copies the used
// parts host-to-device when they changed since the last run and replays a
// captured graph holding the persistent kernel; the kernel reads the counts
// from the device header, so the graph is captured once per launch
// configuration.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "mrg8_internal.hpp"

namespace mrg8 {

struct LaunchConfig {
  int block_threads = 256;
  int blocks_per_cu = 8;
  uint64_t tile = 65536;
};

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

  // Copies the staged block and replays the launch on `stream`.
  int run(const LaunchConfig& cfg, hipStream_t stream);
  // The last run finished (non-blocking query).
  bool done();
  int wait_done();
  hipEvent_t done_event() const { return run_done_; }

 private:
  int capture(const LaunchConfig& cfg);

  DeviceContext* ctx_ = nullptr;
  uint32_t cap_ = 0;
  uint64_t max_tiles_ = 0;
  const uint32_t* states_ = nullptr;
  char* dmem_ = nullptr;
  char* hmem_ = nullptr;
  BatchHeader* d_hdr_ = nullptr;
  BatchSegment* d_segs_ = nullptr;
  uint64_t* d_tiles_ = nullptr;
  uint32_t* d_tile_seg_ = nullptr;
  BatchHeader* h_hdr_ = nullptr;
  BatchSegment* h_segs_ = nullptr;
  uint64_t* h_tiles_ = nullptr;
  uint32_t* h_tile_seg_ = nullptr;
  hipEvent_t staged_ = nullptr, run_done_ = nullptr;
  bool copy_pending_ = false, run_pending_ = false, dirty_ = false;
  uint32_t count_ = 0;
  uint64_t ntiles_ = 0;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
  LaunchConfig captured_{};
  bool have_graph_ = false;
};

// RAII device switch.
class DeviceGuard {
 public:
  explicit DeviceGuard(int dev) : dev_(dev) {
    (void)hipGetDevice(&prev_);
    if (prev_ != dev_) (void)hipSetDevice(dev_);
  }
  ~DeviceGuard() {
    if (prev_ != dev_) (void)hipSetDevice(prev_);
  }

 private:
  int prev_ = 0, dev_ = 0;
};

}  // namespace mrg8
