// Batched generation: many stream segments of uneven length in one launch
// (a plain launch, or a replay of a captured HIP graph; see mrg8_slot.hpp).  Segments are cut
// into fixed-size tiles numbered by a prefix sum; either every wave takes one
// tile, or a persistent grid of W waves takes tiles w, w + W, ..., so uneven
// segment lengths balance without atomics.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

#include "../../include/mrg8.h"
#include "mrg8_internal.hpp"
#include "mrg8_slot.hpp"

using namespace mrg8;

struct mrg8_batch {
  DeviceContext* ctx = nullptr;
  uint64_t max_tiles = 0;
  BatchSlot slot;
  uint64_t total = 0;  // outputs staged
  // one 2048-output tile per wave in 1024-thread blocks: the fastest shape
  // for thousands of streams of uneven length on MI355X (74 % of HBM peak
  // against 68-71 % for the best persistent grid); batches of many short
  // segments with long jumps are faster on a persistent grid
  // (mrg8_batch_autotune tries both)
  mrg8_batch_config cfg{1024, 0, 2048};
  // Direct launches by default: on MI355X a replay of the captured one-kernel
  // graph costs ~10 us more per launch than launching the kernel (0.047 ms
  // against 0.037 ms for 65536 jumps + 256 words each).
  bool use_graph = false;
};

namespace mrg8 {

int validate_batch_config(const mrg8_batch_config& c) {
  if (c.block_threads < 64 || c.block_threads > 1024 || c.block_threads % 64)
    return fail(MRG8_ERR_ARG, "block_threads %d must be a multiple of 64 in [64, 1024]", c.block_threads);
  if (c.blocks_per_cu < -32 || c.blocks_per_cu > 32 ||
      (c.blocks_per_cu < 0 ? -c.blocks_per_cu : c.blocks_per_cu) * (c.block_threads / 64) > 32)
    return fail(MRG8_ERR_ARG,
                "blocks_per_cu %d must be 0 (one tile per wave) or give at most 32 waves per CU (either sign)",
                c.blocks_per_cu);
  if (c.tile_outputs < 256 || c.tile_outputs % 256 || c.tile_outputs > kMaxTile)
    return fail(MRG8_ERR_ARG, "tile_outputs must be a positive multiple of 256, at most 2^40");
  return MRG8_OK;
}

int check_descriptors(const mrg8_stream_desc* descs, uint32_t count, uint64_t tile, uint64_t* tiles,
                      uint64_t* total) {
  *tiles = *total = 0;
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
    *tiles += segment_tiles(d.out, d.format, d.count, tile);
    *total += d.count;
  }
  return MRG8_OK;
}

void write_segments(BatchSegment* segs, const mrg8_stream_desc* descs, uint32_t count) {
  for (uint32_t i = 0; i < count; ++i) {
    const mrg8_stream_desc& d = descs[i];
    BatchSegment& s = segs[i];
    std::memcpy(s.state, d.state, sizeof s.state);
    s.offset = d.offset;
    s.count = d.count;
    s.out = d.out;
    s.format = d.format;
    s.stream = kInlineState;
  }
}

}  // namespace mrg8

namespace {

LaunchConfig launch_config(const mrg8_batch_config& c) {
  LaunchConfig l;
  l.block_threads = c.block_threads;
  l.blocks_per_cu = c.blocks_per_cu;
  l.tile = c.tile_outputs;
  return l;
}

uint64_t count_tiles(const BatchSegment* segs, uint32_t n, uint64_t tile) {
  uint64_t t = 0;
  for (uint32_t i = 0; i < n; ++i) t += segment_tiles(segs[i].out, segs[i].format, segs[i].count, tile);
  return t;
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
  auto* b = new (std::nothrow) mrg8_batch;
  if (!b) return fail(MRG8_ERR_NOMEM, "host allocation failed");
  b->ctx = ctx;
  b->max_tiles = max_tiles;
  if ((rc = b->slot.init(ctx, max_streams, max_tiles, nullptr)) != MRG8_OK) {
    delete b;
    return rc;
  }
  *out = b;
  return MRG8_OK;
}

extern "C" int mrg8_batch_destroy(mrg8_batch_t b) {
  clear_error();
  if (!b) return MRG8_OK;
  delete b;  // the slot waits for its last run
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
  int rc = validate_batch_config(*cfg);
  if (rc) return rc;
  const uint32_t n = b->slot.staged_count();
  if (n) {
    // the tile map depends on the tile size
    if (count_tiles(b->slot.segs(), n, cfg->tile_outputs) > b->max_tiles)
      return fail(MRG8_ERR_CAPACITY, "staged segments need more than max_tiles tiles at this tile size");
    if ((rc = b->slot.wait_staging()) != MRG8_OK) return rc;
    b->slot.finalise(n, cfg->tile_outputs);
  }
  b->cfg = *cfg;
  return MRG8_OK;
}

extern "C" int mrg8_batch_set(mrg8_batch_t b, const mrg8_stream_desc* descs, uint32_t count) {
  clear_error();
  if (!b) return fail(MRG8_ERR_ARG, "null batch");
  if (count && !descs) return fail(MRG8_ERR_ARG, "null descriptors");
  if (count > b->slot.capacity())
    return fail(MRG8_ERR_CAPACITY, "%u segments exceed max_streams %u", count, b->slot.capacity());
  uint64_t tiles = 0, total = 0;
  int rc = check_descriptors(descs, count, b->cfg.tile_outputs, &tiles, &total);
  if (rc) return rc;
  if (tiles > b->max_tiles)
    return fail(MRG8_ERR_CAPACITY, "%llu tiles exceed max_tiles %llu", (unsigned long long)tiles,
                (unsigned long long)b->max_tiles);
  if ((rc = b->slot.wait_staging()) != MRG8_OK) return rc;  // the previous run may still read the staging block
  write_segments(b->slot.segs(), descs, count);
  b->slot.finalise(count, b->cfg.tile_outputs);
  b->total = total;
  return MRG8_OK;
}

extern "C" int mrg8_batch_run(mrg8_batch_t b, void* stream) {
  clear_error();
  if (!b) return fail(MRG8_ERR_ARG, "null batch");
  return b->slot.run(launch_config(b->cfg), static_cast<hipStream_t>(stream), b->use_graph, /*track=*/false);
}

extern "C" int mrg8_batch_use_graph(mrg8_batch_t b, int enable) {
  clear_error();
  if (!b) return fail(MRG8_ERR_ARG, "null batch");
  b->use_graph = enable != 0;
  return MRG8_OK;
}

// ---- autotuning ------------------------------------------------------------------

namespace {

std::string tune_key(const mrg8_batch* b) {
  hipDeviceProp_t prop;
  (void)hipGetDeviceProperties(&prop, b->ctx->device);
  const uint32_t n = b->slot.staged_count();
  const int seg_class = n ? (int)std::floor(std::log2((double)n)) : 0;
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
  bool found = false;
  while (f >> k >> c.block_threads >> c.blocks_per_cu >> c.tile_outputs >> ms)
    if (k == key && validate_batch_config(c) == MRG8_OK) {
      *cfg = c;  // the last entry for a key wins
      found = true;
    }
  clear_error();
  return found;
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
  if (b->slot.staged_count() == 0) return fail(MRG8_ERR_STATE, "stage descriptors before autotuning");
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
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t ce = hipEventCreate(&e0);
  if (ce == hipSuccess) ce = hipEventCreate(&e1);
  if (ce != hipSuccess) {
    if (e0) (void)hipEventDestroy(e0);
    return hip_fail(ce, "hipEventCreate");
  }
  // threads-per-block x waves-per-CU x tile size, the counterpart of the
  // reference's manual TNUM x BS sweep (cuda/tests/test_mrg8_outer.cu:50-53)
  // waves per CU 0: one tile per wave; > 0: persistent grid striding over
  // the tiles; < 0: persistent grid taking the tiles in address order
  const int blocks[] = {256, 512, 1024};
  const int waves_per_cu[] = {0, 4, 8, 16, 32, -16, -24};
  const uint64_t tiles[] = {2048, 4096, 8192, 16384};
  double best_ms = 1e30;
  mrg8_batch_config best_cfg = b->cfg;
  int rc = MRG8_OK;
  for (int bt : blocks)
    for (int w : waves_per_cu)
      for (uint64_t t : tiles) {
        if (rc) break;
        const mrg8_batch_config c{bt, w / (bt / 64), t};
        if ((w != 0 && (c.blocks_per_cu == 0 || w % (bt / 64))) || validate_batch_config(c) != MRG8_OK) continue;
        const uint64_t nt = count_tiles(b->slot.segs(), b->slot.staged_count(), t);
        if (nt > b->max_tiles) continue;
        if (w == 0 && (nt + bt / 64 - 1) / (bt / 64) * (uint64_t)bt > 0xffffffffull) continue;  // 32-bit grid
        if ((rc = mrg8_batch_set_config(b, &c)) != MRG8_OK) break;
        if ((rc = mrg8_batch_run(b, s)) != MRG8_OK) break;  // warm-up (and capture)
        hipError_t e = hipEventRecord(e0, s);
        const int reps = 3;
        for (int r = 0; r < reps && rc == MRG8_OK; ++r) rc = mrg8_batch_run(b, s);
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (rc == MRG8_OK && e != hipSuccess) rc = hip_fail(e, "autotune timing");
        if (rc) break;
        ms /= reps;
        if (ms < best_ms) {
          best_ms = ms;
          best_cfg = c;
        }
      }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  if ((rc = mrg8_batch_set_config(b, &best_cfg)) != MRG8_OK) return rc;
  cache_store(cache_path, key, best_cfg, best_ms);
  if (best) *best = best_cfg;
  return MRG8_OK;
}
