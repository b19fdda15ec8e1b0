// Shared host-side pieces of libmrg8_amd: per-device context, error helpers,
// single-stream fill.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>

#include "../../include/mrg8.h"
#include "mrg8_device.hpp"

namespace mrg8 {

struct TunedShape {
  bool valid = false;
  int block_threads = 0;
  int per_lane = 0;
  uint64_t tile = 0;
};

struct DeviceContext {
  int device = 0;
  int num_cu = 0;
  int num_xcc = 1;  // XCDs the dispatcher deals blocks to in turn
  DeviceTables tables;
  TunedShape tuned[3][64];  // [format][floor(log2 n)], set by mrg8_tune_fill
  std::mutex tuned_mu;      // guards `tuned` (fills on other handles read it)
  std::mutex wide_mu;       // guards the first-use build of tables.wide
  uint8_t* noalloc_scratch = nullptr;  // this device's copy of the fixed scratch area
};

// Blocks per XCD run for kernels that map blocks through xcd_block: runs of
// 32768 threads (64 blocks of 512) on a device with more than one XCD.
inline uint32_t xcd_run_blocks(const DeviceContext& ctx, int block_threads) {
  return ctx.num_xcc > 1 && block_threads > 0 ? (uint32_t)std::max(1, 32768 / block_threads) : 1;
}

// Builds the table of byte positions kJumpPositions..kWidePositions-1 on the
// context's device if it is not there yet.
int ensure_wide_table(DeviceContext& ctx);

// Records `fmt` as this thread's last error message and returns `code`.
int fail(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);
void clear_error();

const JumpTable& host_jump_table();
int device_context(int device, DeviceContext** out);
// The device's context if mrg8_init_device (or any call that needs it) has
// created it, else null; creates nothing.
DeviceContext* existing_device_context(int device);
bool state_valid(const uint32_t st[8]);
int format_size(int format);

// Fills dptr[0..n) with stream positions offset..offset+n-1 of the stream
// whose state is `state`; stream-ordered.  block_threads/tile/per_lane 0 =
// default.
int fill_device(DeviceContext& ctx, const uint32_t state[8], uint64_t offset, void* dptr, uint64_t n, int format,
                hipStream_t stream, int block_threads, uint64_t tile, int per_lane);

// Batch descriptors (include/mrg8.h): checks a launch configuration; checks
// `count` descriptors and sums their tiles (of `tile` outputs) and outputs;
// converts them to the device segment layout (inline states).
int validate_batch_config(const mrg8_batch_config& c);
int check_descriptors(const mrg8_stream_desc* descs, uint32_t count, uint64_t tile, uint64_t* tiles,
                      uint64_t* total);
void write_segments(BatchSegment* segs, const mrg8_stream_desc* descs, uint32_t count);

// Device address of the library's fixed scratch area on the current device
// (taken once per device when its context is created).
int noalloc_scratch_address(uint8_t** ptr);

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

// Host wall clock in seconds (steady).
double now_seconds();

}  // namespace mrg8
