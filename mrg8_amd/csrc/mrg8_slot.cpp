#include "mrg8_slot.hpp"

#include <cstring>

#include "../../include/mrg8.h"

namespace mrg8 {

namespace {
size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

struct Layout {
  size_t segs, tiles, tile_seg, total;
  Layout(uint32_t cap, uint64_t max_tiles) {
    segs = align_up(sizeof(BatchHeader), 64);
    tiles = segs + sizeof(BatchSegment) * (size_t)cap;
    tile_seg = align_up(tiles + sizeof(uint64_t) * ((size_t)cap + 1), 64);
    total = align_up(tile_seg + sizeof(uint32_t) * (size_t)max_tiles, 256);
  }
};
}  // namespace

int BatchSlot::init(DeviceContext* ctx, uint32_t cap, uint64_t max_tiles, const uint32_t* states) {
  release();
  ctx_ = ctx;
  cap_ = cap;
  max_tiles_ = max_tiles;
  states_ = states;
  const Layout l(cap, max_tiles);
  DeviceGuard dg(ctx->device);
  hipError_t e = hipMalloc((void**)&dmem_, l.total);
  if (e == hipSuccess) e = hipHostMalloc((void**)&hmem_, l.total, hipHostMallocDefault);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&staged_, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&run_done_, hipEventDisableTiming);
  if (e != hipSuccess) {
    int rc = hip_fail(e, "batch slot allocation");
    release();
    return rc;
  }
  std::memset(hmem_, 0, l.total);
  d_hdr_ = reinterpret_cast<BatchHeader*>(dmem_);
  d_segs_ = reinterpret_cast<BatchSegment*>(dmem_ + l.segs);
  d_tiles_ = reinterpret_cast<uint64_t*>(dmem_ + l.tiles);
  d_tile_seg_ = reinterpret_cast<uint32_t*>(dmem_ + l.tile_seg);
  h_hdr_ = reinterpret_cast<BatchHeader*>(hmem_);
  h_segs_ = reinterpret_cast<BatchSegment*>(hmem_ + l.segs);
  h_tiles_ = reinterpret_cast<uint64_t*>(hmem_ + l.tiles);
  h_tile_seg_ = reinterpret_cast<uint32_t*>(hmem_ + l.tile_seg);
  return MRG8_OK;
}

void BatchSlot::release() {
  if (!ctx_) return;
  DeviceGuard dg(ctx_->device);
  if (copy_pending_ && staged_) (void)hipEventSynchronize(staged_);
  if (run_pending_ && run_done_) (void)hipEventSynchronize(run_done_);
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

uint64_t BatchSlot::finalise(uint32_t count, uint64_t tile) {
  uint64_t t = 0;
  for (uint32_t i = 0; i < count; ++i) {
    h_tiles_[i] = t;
    const uint64_t n = segment_tiles(h_segs_[i].out, h_segs_[i].format, h_segs_[i].count, tile);
    for (uint64_t k = 0; k < n && t + k < max_tiles_; ++k) h_tile_seg_[t + k] = i;
    t += n;
  }
  h_tiles_[count] = t;
  h_hdr_->nsegs = count;
  h_hdr_->ntiles = t;
  count_ = count;
  ntiles_ = t;
  dirty_ = true;
  return t;
}

int BatchSlot::capture(const LaunchConfig& cfg) {
  if (have_graph_ && cfg.block_threads == captured_.block_threads && cfg.blocks_per_cu == captured_.blocks_per_cu &&
      cfg.tile == captured_.tile)
    return MRG8_OK;
  if (run_pending_) (void)hipEventSynchronize(run_done_);
  if (exec_) (void)hipGraphExecDestroy(exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  exec_ = nullptr;
  graph_ = nullptr;
  have_graph_ = false;
  hipStream_t cs = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
  if (e != hipSuccess) return hip_fail(e, "hipStreamCreate(capture)");
  BatchArgs a{};
  a.hdr = d_hdr_;
  a.segs = d_segs_;
  a.tile_begin = d_tiles_;
  a.tile_seg = d_tile_seg_;
  a.tile = cfg.tile;
  a.states = states_;
  a.lane = ctx_->tables.lane;
  a.jump = ctx_->tables.jump;
  const int grid = ctx_->num_cu * cfg.blocks_per_cu;
  e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    hipError_t e1 = launch_batch(a, grid, cfg.block_threads, cs);
    hipError_t e2 = hipStreamEndCapture(cs, &graph_);
    e = e1 != hipSuccess ? e1 : e2;
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0);
  (void)hipStreamDestroy(cs);
  if (e != hipSuccess) return hip_fail(e, "graph capture");
  captured_ = cfg;
  have_graph_ = true;
  return MRG8_OK;
}

int BatchSlot::run(const LaunchConfig& cfg, hipStream_t stream) {
  if (count_ == 0 || ntiles_ == 0) return MRG8_OK;
  if (ntiles_ > max_tiles_) return fail(MRG8_ERR_CAPACITY, "staged work needs more than max_tiles tiles");
  DeviceGuard dg(ctx_->device);
  int rc = capture(cfg);
  if (rc) return rc;
  hipError_t e = hipSuccess;
  if (dirty_) {  // replays of an unchanged batch copy nothing
    e = hipMemcpyAsync(d_hdr_, h_hdr_, sizeof(BatchHeader), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(d_segs_, h_segs_, sizeof(BatchSegment) * count_, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(d_tiles_, h_tiles_, sizeof(uint64_t) * (count_ + 1), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(d_tile_seg_, h_tile_seg_, sizeof(uint32_t) * ntiles_, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipEventRecord(staged_, stream);
    if (e != hipSuccess) return hip_fail(e, "batch staging copy");
    copy_pending_ = true;
    dirty_ = false;
  }
  e = hipGraphLaunch(exec_, stream);
  if (e == hipSuccess) e = hipEventRecord(run_done_, stream);
  if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
  run_pending_ = true;
  return MRG8_OK;
}

bool BatchSlot::done() {
  if (!run_pending_) return true;
  DeviceGuard dg(ctx_->device);
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
  hipError_t e = hipEventSynchronize(run_done_);
  if (e != hipSuccess) return hip_fail(e, "hipEventSynchronize(run)");
  run_pending_ = false;
  copy_pending_ = false;
  return MRG8_OK;
}

}  // namespace mrg8
