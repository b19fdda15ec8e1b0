// Batched generation that allocates nothing (mrg8_batch_generate_noalloc).
//
// The same generation as mrg8_batch_set + mrg8_batch_run: the batch block
// (header, segments, tile prefix sums, tile -> segment map; mrg8_device.hpp)
// is laid out by the same code and read by the same kernels.  It lives in
// `scratch`, a fixed device array of this file that every GPU the library
// runs on holds until the process ends, instead of a handle's allocation; the
// host side builds it in a fixed host array of the same size.  The array's
// name is a host-side stand-in, not a device address: each device's address
// is taken once, with hipGetSymbolAddress, when the device's context is
// created (mrg8_init_device), and only that address is handed to copies and
// kernels.
//
// Calls share the scratch area of their device and the host array: the host
// array is guarded by a mutex for the length of a call (the block is copied
// out of it before the call returns), but a kernel still reading a device's
// scratch area races with a later call on that device issued on another
// stream.  Nothing in the block needs clearing: the kernels read only what
// the header of the current call describes.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/mrg8.h"
#include "mrg8_internal.hpp"
#include "mrg8_slot.hpp"

namespace mrg8 {
namespace {

static __device__ __align__(256) uint8_t scratch[MRG8_NOALLOC_SCRATCH_BYTES];

std::mutex g_host_mu;  // guards g_host
alignas(256) uint8_t g_host[MRG8_NOALLOC_SCRATCH_BYTES];

}  // namespace

int noalloc_scratch_address(uint8_t** ptr) {
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
  *ptr = static_cast<uint8_t*>(p);
  return MRG8_OK;
}

}  // namespace mrg8

using namespace mrg8;

extern "C" int mrg8_batch_generate_noalloc(int device, const mrg8_stream_desc* descs, uint32_t count,
                                           const mrg8_batch_config* cfg, void* stream) {
  clear_error();
  if (count && !descs) return fail(MRG8_ERR_ARG, "null descriptors");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(MRG8_ERR_NODEVICE, "no HIP device available");
  if (device < 0 || device >= ndev) return fail(MRG8_ERR_ARG, "device %d out of range (have %d)", device, ndev);
  const mrg8_batch_config c = cfg ? *cfg : mrg8_batch_config{1024, 0, 2048};
  int rc = validate_batch_config(c);
  if (rc) return rc;
  uint64_t tiles = 0, total = 0;
  if ((rc = check_descriptors(descs, count, c.tile_outputs, &tiles, &total)) != MRG8_OK) return rc;
  const size_t bytes = batch_block_bytes(count, tiles);
  if (count > (MRG8_NOALLOC_SCRATCH_BYTES - sizeof(BatchHeader)) / sizeof(BatchSegment) ||
      bytes > MRG8_NOALLOC_SCRATCH_BYTES)
    return fail(MRG8_ERR_CAPACITY,
                "%u segments in %llu tiles need %zu bytes of batch block; the fixed scratch area holds %u",
                count, (unsigned long long)tiles, bytes, (unsigned)MRG8_NOALLOC_SCRATCH_BYTES);
  if (c.blocks_per_cu == 0) {  // one wave per tile: the grid is a 32-bit quantity
    const uint64_t wpb = (uint64_t)c.block_threads / 64;
    if ((tiles + wpb - 1) / wpb * (uint64_t)c.block_threads > 0xffffffffull)
      return fail(MRG8_ERR_CAPACITY, "%llu tiles need more than 2^32 - 1 work-items at one tile per wave",
                  (unsigned long long)tiles);
  }
  if (tiles == 0) return MRG8_OK;
  // the context, tables and scratch address come from mrg8_init_device:
  // creating them here would allocate
  DeviceContext* ctx = existing_device_context(device);
  if (!ctx || !ctx->noalloc_scratch)
    return fail(MRG8_ERR_STATE, "call mrg8_init_device(%d) before mrg8_batch_generate_noalloc", device);
  DeviceGuard dg(device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  std::lock_guard<std::mutex> lk(g_host_mu);
  write_segments(reinterpret_cast<BatchSegment*>(g_host + sizeof(BatchHeader)), descs, count);
  uint64_t nt = 0;
  const size_t used = layout_batch_block(g_host, count, c.tile_outputs, tiles, &nt);
  // pageable source: the runtime has read g_host when the call returns
  hipError_t e = hipMemcpyAsync(ctx->noalloc_scratch, g_host, used, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return hip_fail(e, "scratch staging copy");
  BatchArgs a{};
  a.block = ctx->noalloc_scratch;
  a.tile = c.tile_outputs;
  a.states = nullptr;
  a.lane = ctx->tables.lane;
  a.jump = ctx->tables.jump;
  a.xcds = (uint32_t)ctx->num_xcc;
  a.xcd_group = xcd_run_blocks(*ctx, c.block_threads);
  // the staging copy above wrote the header's work counters as zero
  e = c.blocks_per_cu == 0  ? launch_batch_tiles(a, nt, c.block_threads, s)
      : c.blocks_per_cu < 0 ? launch_batch_queue(a, ctx->num_cu * -c.blocks_per_cu, c.block_threads, s)
                            : launch_batch(a, ctx->num_cu * c.blocks_per_cu, c.block_threads, s);
  if (e != hipSuccess) return hip_fail(e, "batch launch (no-alloc)");
  return MRG8_OK;
}
