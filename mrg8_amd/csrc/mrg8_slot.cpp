#include "mrg8_slot.hpp"

#include <algorithm>
#include <cstring>

#include "../../include/mrg8.h"

namespace mrg8 {

namespace {
size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Byte offsets of the tables behind `n` segments (see mrg8_device.hpp).
size_t tile_begin_offset(uint32_t n) { return sizeof(BatchHeader) + sizeof(BatchSegment) * (size_t)n; }
size_t tile_seg_offset(uint32_t n) { return tile_begin_offset(n) + sizeof(uint64_t) * ((size_t)n + 1); }
}  // namespace

int BatchSlot::init(DeviceContext* ctx, uint32_t cap, uint64_t max_tiles, const uint32_t* states) {
  release();
  ctx_ = ctx;
  cap_ = cap;
  max_tiles_ = max_tiles;
  states_ = states;
  const size_t total = align_up(tile_seg_offset(cap) + sizeof(uint32_t) * (size_t)max_tiles, 256);
  DeviceGuard dg(ctx->device);
  hipError_t e = hipMalloc((void**)&dmem_, total);
  // coherent: the staging kernel reads it over the bus every run, never a
  // copy the GPU cached from an earlier run of this slot
  if (e == hipSuccess) e = hipHostMalloc((void**)&hmem_, total, hipHostMallocCoherent);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&staged_, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&run_done_, hipEventDisableTiming);
  if (e != hipSuccess) {
    int rc = hip_fail(e, "batch slot allocation");
    release();
    return rc;
  }
  std::memset(hmem_, 0, total);
  h_segs_ = reinterpret_cast<BatchSegment*>(hmem_ + sizeof(BatchHeader));
  return MRG8_OK;
}

void BatchSlot::release() {
  if (!ctx_) return;
  DeviceGuard dg(ctx_->device);
  if (copy_pending_ && staged_) (void)hipEventSynchronize(staged_);
  if (run_pending_ && run_done_) (void)wait_done();
  if (exec_) (void)hipGraphExecDestroy(exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  if (staged_) (void)hipEventDestroy(staged_);
  if (run_done_) (void)hipEventDestroy(run_done_);
  if (hmem_) (void)hipHostFree(hmem_);
  if (dmem_) (void)hipFree(dmem_);
  exec_ = nullptr;
  graph_ = nullptr;
  staged_ = run_done_ = nullptr;
  hmem_ = dmem_ = nullptr;
  copy_pending_ = run_pending_ = have_graph_ = false;
  last_stream_ = nullptr;
  last_marked_ = true;
  last_queue_ = false;
  ctx_ = nullptr;
}

int BatchSlot::wait_staging() {
  if (!copy_pending_) return MRG8_OK;
  DeviceGuard dg(ctx_->device);
  hipError_t e = hipEventSynchronize(staged_);
  if (e != hipSuccess) return hip_fail(e, "hipEventSynchronize(staging)");
  copy_pending_ = false;
  return MRG8_OK;
}

size_t batch_block_bytes(uint32_t count, uint64_t ntiles) {
  return tile_seg_offset(count) + sizeof(uint32_t) * (size_t)ntiles;
}

size_t layout_batch_block(uint8_t* mem, uint32_t count, uint64_t tile, uint64_t max_tiles, uint64_t* ntiles) {
  const size_t tb_off = tile_begin_offset(count), ts_off = tile_seg_offset(count);
  BatchHeader* hdr = reinterpret_cast<BatchHeader*>(mem);
  const BatchSegment* segs = reinterpret_cast<const BatchSegment*>(mem + sizeof(BatchHeader));
  uint64_t* tiles = reinterpret_cast<uint64_t*>(mem + tb_off);
  uint32_t* tile_seg = reinterpret_cast<uint32_t*>(mem + ts_off);
  uint64_t t = 0;
  for (uint32_t i = 0; i < count; ++i) {
    tiles[i] = t;
    const uint64_t n = segment_tiles(segs[i].out, segs[i].format, segs[i].count, tile);
    const uint64_t end = std::min<uint64_t>(t + n, max_tiles);
    if (t < end) std::fill(tile_seg + t, tile_seg + end, i);
    t += n;
  }
  tiles[count] = t;
  hdr->nsegs = count;
  hdr->ntiles = t;
  hdr->tile_begin_off = tb_off;
  hdr->tile_seg_off = ts_off;
  hdr->queue_next = 0;  // k_batch_queue's counters start at zero
  hdr->queue_done = 0;
  *ntiles = t;
  return batch_block_bytes(count, std::min<uint64_t>(t, max_tiles));
}

uint64_t BatchSlot::finalise(uint32_t count, uint64_t tile) {
  uint64_t t = 0;
  used_bytes_ = layout_batch_block(hmem_, count, tile, max_tiles_, &t);
  count_ = count;
  ntiles_ = t;
  dirty_ = true;
  return t;
}

BatchArgs BatchSlot::args(const LaunchConfig& cfg) const {
  BatchArgs a{};
  a.block = dmem_;
  a.tile = cfg.tile;
  a.states = states_;
  a.lane = ctx_->tables.lane;
  a.jump = ctx_->tables.jump;
  a.xcds = (uint32_t)ctx_->num_xcc;
  a.xcd_group = xcd_run_blocks(*ctx_, cfg.block_threads);
  return a;
}

hipError_t BatchSlot::launch(const LaunchConfig& cfg, hipStream_t stream) const {
  if (cfg.blocks_per_cu == 0) return launch_batch_tiles(args(cfg), ntiles_, cfg.block_threads, stream);
  if (cfg.blocks_per_cu < 0)
    return launch_batch_queue(args(cfg), ctx_->num_cu * -cfg.blocks_per_cu, cfg.block_threads, stream);
  return launch_batch(args(cfg), ctx_->num_cu * cfg.blocks_per_cu, cfg.block_threads, stream);
}

int BatchSlot::capture(const LaunchConfig& cfg) {
  // a one-tile-per-wave grid depends on the tile count, a persistent one of
  // either kind does not (its kernel reads the count from the device block)
  if (have_graph_ && cfg.block_threads == captured_.block_threads && cfg.blocks_per_cu == captured_.blocks_per_cu &&
      cfg.tile == captured_.tile && (cfg.blocks_per_cu != 0 || captured_ntiles_ == ntiles_))
    return MRG8_OK;
  if (int rc = wait_done()) return rc;  // the old graph may still be running
  if (exec_) (void)hipGraphExecDestroy(exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  exec_ = nullptr;
  graph_ = nullptr;
  have_graph_ = false;
  hipStream_t cs = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
  if (e != hipSuccess) return hip_fail(e, "hipStreamCreate(capture)");
  e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    hipError_t e1 = launch(cfg, cs);
    hipError_t e2 = hipStreamEndCapture(cs, &graph_);
    e = e1 != hipSuccess ? e1 : e2;
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0);
  (void)hipStreamDestroy(cs);
  if (e != hipSuccess) return hip_fail(e, "graph capture");
  captured_ = cfg;
  captured_ntiles_ = ntiles_;
  have_graph_ = true;
  return MRG8_OK;
}

hipError_t BatchSlot::mark_last_run() {
  if (!run_pending_ || last_marked_) return hipSuccess;
  hipError_t e = hipEventRecord(run_done_, last_stream_);
  if (e == hipSuccess) last_marked_ = true;
  return e;
}

int BatchSlot::run(const LaunchConfig& cfg, hipStream_t stream, bool graph, bool track) {
  if (count_ == 0 || ntiles_ == 0) return MRG8_OK;
  if (ntiles_ > max_tiles_) return fail(MRG8_ERR_CAPACITY, "staged work needs more than max_tiles tiles");
  const uint64_t wpb = (uint64_t)cfg.block_threads / 64;
  if (cfg.blocks_per_cu == 0) {  // one wave per tile: the grid (work-items in total) is a 32-bit quantity
    if ((ntiles_ + wpb - 1) / wpb * (uint64_t)cfg.block_threads > 0xffffffffull)
      return fail(MRG8_ERR_CAPACITY, "%llu tiles need more than 2^32 - 1 work-items at one tile per wave; use larger "
                  "tiles or a persistent grid", (unsigned long long)ntiles_);
  } else if (cfg.blocks_per_cu < 0) {  // the work counter is 32 bits: groups plus one last take per block
    if ((ntiles_ + wpb - 1) / wpb + (uint64_t)ctx_->num_cu * (uint64_t)-cfg.blocks_per_cu > 0xffffffffull)
      return fail(MRG8_ERR_CAPACITY, "%llu tiles are more groups than the address-order grid counts; use larger tiles",
                  (unsigned long long)ntiles_);
  }
  DeviceGuard dg(ctx_->device);
  int rc = graph ? capture(cfg) : MRG8_OK;
  if (rc) return rc;
  hipError_t e = hipSuccess;
  // The device block may be overwritten once this slot's previous launch has
  // finished, and an address-order grid (blocks_per_cu < 0) may start only
  // after the previous one has reset the work counters in it: on the same
  // stream both are ordered after the previous launch; on another stream
  // they wait for it.  Runs next to an address-order run wait as well, so
  // that two such runs are never apart by one unordered run in between.
  if (run_pending_ && stream != last_stream_ && (dirty_ || cfg.blocks_per_cu < 0 || last_queue_)) {
    if (mark_last_run() == hipSuccess) e = hipStreamWaitEvent(stream, run_done_, 0);
    else if ((rc = wait_done()) != MRG8_OK) return rc;  // the last stream no longer accepts work
    if (e != hipSuccess) return hip_fail(e, "hipStreamWaitEvent(previous run)");
  }
  if (dirty_) {  // runs of an unchanged batch copy nothing
    e = launch_stage(hmem_, dmem_, used_bytes_, stream);
    if (e == hipSuccess) e = hipEventRecord(staged_, stream);
    if (e != hipSuccess) return hip_fail(e, "batch staging copy");
    copy_pending_ = true;
    dirty_ = false;
  }
  e = graph ? hipGraphLaunch(exec_, stream) : launch(cfg, stream);
  if (e != hipSuccess) return hip_fail(e, graph ? "hipGraphLaunch" : "batch launch");
  run_pending_ = true;
  last_stream_ = stream;
  last_marked_ = false;
  last_queue_ = cfg.blocks_per_cu < 0;
  if (track && (e = mark_last_run()) != hipSuccess) return hip_fail(e, "hipEventRecord(run)");
  return MRG8_OK;
}

bool BatchSlot::done() {
  if (!run_pending_) return true;
  DeviceGuard dg(ctx_->device);
  if (mark_last_run() != hipSuccess) return false;
  if (hipEventQuery(run_done_) == hipSuccess) {
    run_pending_ = false;
    copy_pending_ = false;
    return true;
  }
  return false;
}

int BatchSlot::wait_done() {
  if (!run_pending_) return MRG8_OK;
  DeviceGuard dg(ctx_->device);
  hipError_t e = mark_last_run();
  e = e == hipSuccess ? hipEventSynchronize(run_done_) : hipDeviceSynchronize();
  if (e != hipSuccess) return hip_fail(e, "waiting for the last batch run");
  run_pending_ = false;
  copy_pending_ = false;
  return MRG8_OK;
}

}  // namespace mrg8
