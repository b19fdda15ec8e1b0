// Batched generation: many stream segments of uneven length in one
// persistent launch, replayed from a captured HIP graph.
//
// Layout in device memory (one allocation):
//   BatchHeader | BatchSegment[max_streams] | uint64 tile_begin[max_streams+1] | uint32 cursor
// The host stages header, segments and tile prefix sums in pinned memory,
// copies them with one async H2D copy, and replays a graph holding
// [zero the cursor] -> [persistent kernel].  Waves pull fixed-size tiles
// from the cursor until it passes the tile count, so uneven segment lengths
// balance automatically.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/mrg8.h"
#include "mrg8_internal.hpp"

using namespace mrg8;

struct mrg8_batch {
  DeviceContext* ctx = nullptr;
  uint32_t max_streams = 0;
  uint64_t max_tiles = 0;
  // device
  char* dmem = nullptr;
  BatchHeader* d_hdr = nullptr;
  BatchSegment* d_segs = nullptr;
  uint64_t* d_tile_begin = nullptr;
  uint32_t* d_cursor = nullptr;
  size_t copy_bytes_max = 0;
  // pinned host staging, same layout as the device block up to the cursor
  char* hmem = nullptr;
  BatchHeader* h_hdr = nullptr;
  BatchSegment* h_segs = nullptr;
  uint64_t* h_tile_begin = nullptr;
  hipEvent_t staged_done = nullptr;  // the last H2D copy out of the staging block finished
  hipEvent_t run_done = nullptr;     // the last launch finished
  bool copy_pending = false;
  bool run_pending = false;
  // run state
  uint32_t count = 0;
  uint64_t ntiles = 0;
  uint64_t total = 0;
  mrg8_batch_config cfg{256, 8, 65536};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool graph_valid = false;
};

namespace {

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

struct Layout {
  size_t hdr, segs, tiles, cursor, copy_bytes, total;
};

Layout layout_for(uint32_t max_streams) {
  Layout l;
  l.hdr = 0;
  l.segs = align_up(sizeof(BatchHeader), 64);
  l.tiles = l.segs + sizeof(BatchSegment) * (size_t)max_streams;
  l.cursor = align_up(l.tiles + sizeof(uint64_t) * ((size_t)max_streams + 1), 256);
  l.copy_bytes = l.cursor;
  l.total = l.cursor + 256;
  return l;
}

class DeviceGuard {
 public:
  explicit DeviceGuard(int dev) {
    (void)hipGetDevice(&prev_);
    if (prev_ != dev) (void)hipSetDevice(dev);
    dev_ = dev;
  }
  ~DeviceGuard() {
    if (prev_ != dev_) (void)hipSetDevice(prev_);
  }

 private:
  int prev_ = 0, dev_ = 0;
};

int validate_config(const mrg8_batch_config& c) {
  if (c.block_threads < 64 || c.block_threads > 1024 || c.block_threads % 64)
    return fail(MRG8_ERR_ARG, "block_threads %d must be a multiple of 64 in [64, 1024]", c.block_threads);
  if (c.blocks_per_cu < 1 || c.blocks_per_cu * (c.block_threads / 64) > 32)
    return fail(MRG8_ERR_ARG, "blocks_per_cu %d gives more than 32 waves per CU", c.blocks_per_cu);
  if (c.tile_outputs < 256 || c.tile_outputs % 256)
    return fail(MRG8_ERR_ARG, "tile_outputs must be a positive multiple of 256");
  return MRG8_OK;
}

int capture(mrg8_batch* b) {
  if (b->graph_valid) return MRG8_OK;
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  if (b->graph) {
    (void)hipGraphDestroy(b->graph);
    b->graph = nullptr;
  }
  hipStream_t cs = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
  if (e != hipSuccess) return hip_fail(e, "hipStreamCreate(capture)");
  BatchArgs a{};
  a.hdr = b->d_hdr;
  a.segs = b->d_segs;
  a.tile_begin = b->d_tile_begin;
  a.tile = b->cfg.tile_outputs;
  a.cursor = b->d_cursor;
  a.lane = b->ctx->tables.lane;
  a.jump = b->ctx->tables.jump;
  const int grid = b->ctx->num_cu * b->cfg.blocks_per_cu;
  e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    hipError_t e1 = launch_zero_u32(b->d_cursor, 1, cs);
    hipError_t e2 = launch_batch(a, grid, b->cfg.block_threads, cs);
    hipError_t e3 = hipStreamEndCapture(cs, &b->graph);
    e = e1 != hipSuccess ? e1 : (e2 != hipSuccess ? e2 : e3);
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
  (void)hipStreamDestroy(cs);
  if (e != hipSuccess) return hip_fail(e, "graph capture");
  b->graph_valid = true;
  return MRG8_OK;
}

}  // namespace

extern "C" int mrg8_batch_create(int device, uint32_t max_streams, uint64_t max_tiles, mrg8_batch_t* out) {
  clear_error();
  if (!out) return fail(MRG8_ERR_ARG, "null output handle");
  *out = nullptr;
  if (max_streams == 0) return fail(MRG8_ERR_ARG, "max_streams must be positive");
  if (max_tiles == 0 || max_tiles > (1ull << 31)) return fail(MRG8_ERR_ARG, "max_tiles must be in [1, 2^31]");
  DeviceContext* ctx = nullptr;
  int rc = device_context(device, &ctx);
  if (rc) return rc;
  DeviceGuard dg(device);
  auto* b = new (std::nothrow) mrg8_batch;
  if (!b) return fail(MRG8_ERR_NOMEM, "host allocation failed");
  b->ctx = ctx;
  b->max_streams = max_streams;
  b->max_tiles = max_tiles;
  const Layout l = layout_for(max_streams);
  hipError_t e = hipMalloc((void**)&b->dmem, l.total);
  if (e == hipSuccess) e = hipHostMalloc((void**)&b->hmem, l.copy_bytes, hipHostMallocDefault);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->staged_done, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->run_done, hipEventDisableTiming);
  if (e != hipSuccess) {
    int r = hip_fail(e, "batch allocation");
    mrg8_batch_destroy(b);
    return r;
  }
  b->d_hdr = reinterpret_cast<BatchHeader*>(b->dmem + l.hdr);
  b->d_segs = reinterpret_cast<BatchSegment*>(b->dmem + l.segs);
  b->d_tile_begin = reinterpret_cast<uint64_t*>(b->dmem + l.tiles);
  b->d_cursor = reinterpret_cast<uint32_t*>(b->dmem + l.cursor);
  b->h_hdr = reinterpret_cast<BatchHeader*>(b->hmem + l.hdr);
  b->h_segs = reinterpret_cast<BatchSegment*>(b->hmem + l.segs);
  b->h_tile_begin = reinterpret_cast<uint64_t*>(b->hmem + l.tiles);
  b->copy_bytes_max = l.copy_bytes;
  std::memset(b->hmem, 0, l.copy_bytes);
  // Default: 256-thread blocks, 8 per CU (32 waves), 64Ki outputs per tile.
  b->cfg = mrg8_batch_config{256, 8, 65536};
  *out = b;
  return MRG8_OK;
}

extern "C" int mrg8_batch_destroy(mrg8_batch_t b) {
  clear_error();
  if (!b) return MRG8_OK;
  DeviceGuard dg(b->ctx ? b->ctx->device : 0);
  if (b->copy_pending && b->staged_done) (void)hipEventSynchronize(b->staged_done);
  if (b->run_pending && b->run_done) (void)hipEventSynchronize(b->run_done);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->staged_done) (void)hipEventDestroy(b->staged_done);
  if (b->run_done) (void)hipEventDestroy(b->run_done);
  if (b->hmem) (void)hipHostFree(b->hmem);
  if (b->dmem) (void)hipFree(b->dmem);
  delete b;
  return MRG8_OK;
}

extern "C" int mrg8_batch_get_config(mrg8_batch_t b, mrg8_batch_config* cfg) {
  clear_error();
  if (!b || !cfg) return fail(MRG8_ERR_ARG, "null argument");
  *cfg = b->cfg;
  return MRG8_OK;
}

extern "C" int mrg8_batch_set_config(mrg8_batch_t b, const mrg8_batch_config* cfg) {
  clear_error();
  if (!b || !cfg) return fail(MRG8_ERR_ARG, "null argument");
  int rc = validate_config(*cfg);
  if (rc) return rc;
  const bool tile_changed = cfg->tile_outputs != b->cfg.tile_outputs;
  b->cfg = *cfg;
  b->graph_valid = false;
  if (tile_changed && b->count) {
    // Recompute the prefix sums for the staged segments.
    if (b->copy_pending) {
      (void)hipEventSynchronize(b->staged_done);
      b->copy_pending = false;
    }
    uint64_t t = 0;
    for (uint32_t i = 0; i < b->count; ++i) {
      b->h_tile_begin[i] = t;
      t += (b->h_segs[i].count + b->cfg.tile_outputs - 1) / b->cfg.tile_outputs;
    }
    b->h_tile_begin[b->count] = t;
    if (t > b->max_tiles) return fail(MRG8_ERR_CAPACITY, "%llu tiles exceed max_tiles", (unsigned long long)t);
    b->ntiles = t;
    b->h_hdr->ntiles = t;
  }
  return MRG8_OK;
}

extern "C" int mrg8_batch_set(mrg8_batch_t b, const mrg8_stream_desc* descs, uint32_t count) {
  clear_error();
  if (!b) return fail(MRG8_ERR_ARG, "null batch");
  if (count && !descs) return fail(MRG8_ERR_ARG, "null descriptors");
  if (count > b->max_streams) return fail(MRG8_ERR_CAPACITY, "%u segments exceed max_streams %u", count, b->max_streams);
  // validate before touching the staging block
  uint64_t tiles = 0, total = 0;
  for (uint32_t i = 0; i < count; ++i) {
    const mrg8_stream_desc& d = descs[i];
    const int esz = format_size(d.format);
    if (!esz) return fail(MRG8_ERR_ARG, "segment %u: unknown format %d", i, d.format);
    if (d.reserved) return fail(MRG8_ERR_ARG, "segment %u: reserved field must be 0", i);
    if (!state_valid(d.state)) return fail(MRG8_ERR_ARG, "segment %u: state words must be at most 2^31 - 1", i);
    if (d.count && !d.out) return fail(MRG8_ERR_ARG, "segment %u: null output", i);
    if (reinterpret_cast<uintptr_t>(d.out) % (uintptr_t)esz)
      return fail(MRG8_ERR_ARG, "segment %u: output must be %d-byte aligned", i, esz);
    if (d.count > UINT64_MAX - d.offset) return fail(MRG8_ERR_ARG, "segment %u: offset + count overflows", i);
    tiles += (d.count + b->cfg.tile_outputs - 1) / b->cfg.tile_outputs;
    total += d.count;
  }
  if (tiles > b->max_tiles)
    return fail(MRG8_ERR_CAPACITY, "%llu tiles exceed max_tiles %llu", (unsigned long long)tiles,
                (unsigned long long)b->max_tiles);
  DeviceGuard dg(b->ctx->device);
  if (b->copy_pending) {  // the previous run may still be reading the staging block
    hipError_t e = hipEventSynchronize(b->staged_done);
    if (e != hipSuccess) return hip_fail(e, "hipEventSynchronize");
    b->copy_pending = false;
  }
  uint64_t t = 0;
  for (uint32_t i = 0; i < count; ++i) {
    const mrg8_stream_desc& d = descs[i];
    BatchSegment& s = b->h_segs[i];
    std::memcpy(s.state, d.state, sizeof s.state);
    s.offset = d.offset;
    s.count = d.count;
    s.out = d.out;
    s.format = d.format;
    s.pad = 0;
    b->h_tile_begin[i] = t;
    t += (d.count + b->cfg.tile_outputs - 1) / b->cfg.tile_outputs;
  }
  b->h_tile_begin[count] = t;
  b->h_hdr->nsegs = count;
  b->h_hdr->pad = 0;
  b->h_hdr->ntiles = t;
  b->count = count;
  b->ntiles = t;
  b->total = total;
  return MRG8_OK;
}

extern "C" int mrg8_batch_run(mrg8_batch_t b, void* stream) {
  clear_error();
  if (!b) return fail(MRG8_ERR_ARG, "null batch");
  if (b->count == 0 || b->ntiles == 0) return MRG8_OK;
  DeviceGuard dg(b->ctx->device);
  int rc = capture(b);
  if (rc) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipError_t e = hipMemcpyAsync(b->d_hdr, b->h_hdr, sizeof(BatchHeader), hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(b->d_segs, b->h_segs, sizeof(BatchSegment) * b->count, hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(b->d_tile_begin, b->h_tile_begin, sizeof(uint64_t) * (b->count + 1), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipEventRecord(b->staged_done, s);
  if (e != hipSuccess) return hip_fail(e, "batch staging copy");
  b->copy_pending = true;
  e = hipGraphLaunch(b->exec, s);
  if (e == hipSuccess) e = hipEventRecord(b->run_done, s);
  if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
  b->run_pending = true;
  return MRG8_OK;
}

// ---- autotuning ------------------------------------------------------------------

namespace {

std::string tune_key(const mrg8_batch* b) {
  hipDeviceProp_t prop;
  (void)hipGetDeviceProperties(&prop, b->ctx->device);
  const int seg_class = b->count ? (int)std::floor(std::log2((double)b->count)) : 0;
  const int out_class = b->total ? (int)std::floor(std::log2((double)b->total)) : 0;
  std::ostringstream k;
  k << prop.gcnArchName << ":" << prop.multiProcessorCount << ":segs2^" << seg_class << ":outs2^" << out_class;
  std::string s = k.str();
  for (char& c : s)
    if (c == ' ') c = '_';
  return s;
}

bool cache_lookup(const char* path, const std::string& key, mrg8_batch_config* cfg) {
  if (!path) return false;
  std::ifstream f(path);
  std::string k;
  mrg8_batch_config c{};
  double ms;
  while (f >> k >> c.block_threads >> c.blocks_per_cu >> c.tile_outputs >> ms) {
    if (k == key && validate_config(c) == MRG8_OK) {
      *cfg = c;
      return true;
    }
  }
  return false;
}

void cache_store(const char* path, const std::string& key, const mrg8_batch_config& c, double ms) {
  if (!path) return;
  std::ofstream f(path, std::ios::app);
  f << key << " " << c.block_threads << " " << c.blocks_per_cu << " " << c.tile_outputs << " " << ms << "\n";
}

}  // namespace

extern "C" int mrg8_batch_autotune(mrg8_batch_t b, void* stream, const char* cache_path, mrg8_batch_config* best) {
  clear_error();
  if (!b) return fail(MRG8_ERR_ARG, "null batch");
  if (b->count == 0) return fail(MRG8_ERR_STATE, "stage descriptors before autotuning");
  const std::string key = tune_key(b);
  mrg8_batch_config found{};
  if (cache_lookup(cache_path, key, &found)) {
    int rc = mrg8_batch_set_config(b, &found);
    if (rc) return rc;
    if (best) *best = found;
    return MRG8_OK;
  }
  DeviceGuard dg(b->ctx->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  // The threads-per-block x waves-per-CU x tile sweep mirrors the reference's
  // manual TNUM x BS sweep (cuda/tests/test_mrg8_outer.cu:50-53).
  const int blocks[] = {256, 512, 1024};
  const int waves_per_cu[] = {16, 32};
  const uint64_t tiles[] = {16384, 65536, 262144};
  double best_ms = 1e30;
  mrg8_batch_config best_cfg = b->cfg;
  int rc = MRG8_OK;
  for (int bt : blocks)
    for (int w : waves_per_cu)
      for (uint64_t t : tiles) {
        mrg8_batch_config c{bt, w / (bt / 64), t};
        if (c.blocks_per_cu < 1 || validate_config(c) != MRG8_OK) continue;
        uint64_t ntiles = 0;
        for (uint32_t i = 0; i < b->count; ++i) ntiles += (b->h_segs[i].count + t - 1) / t;
        if (ntiles > b->max_tiles) continue;
        if ((rc = mrg8_batch_set_config(b, &c)) != MRG8_OK) break;
        if ((rc = mrg8_batch_run(b, s)) != MRG8_OK) break;  // warm-up and capture
        (void)hipEventRecord(e0, s);
        const int reps = 3;
        for (int r = 0; r < reps && rc == MRG8_OK; ++r) rc = mrg8_batch_run(b, s);
        (void)hipEventRecord(e1, s);
        (void)hipEventSynchronize(e1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        ms /= reps;
        if (ms < best_ms) {
          best_ms = ms;
          best_cfg = c;
        }
      }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  rc = mrg8_batch_set_config(b, &best_cfg);
  if (rc) return rc;
  cache_store(cache_path, key, best_cfg, best_ms);
  if (best) *best = best_cfg;
  return MRG8_OK;
}
