// C ABI of the MI355X-native MRG8 generator: generator handles, device
// generation, host-memory forms, jump-table output.  Declared in include/mrg8.h.
#include "../../include/mrg8.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "mrg8_device.hpp"
#include "mrg8_internal.hpp"

namespace mrg8 {

// ---- errors ----------------------------------------------------------------

namespace {
thread_local std::string g_last_error;
}

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(e == hipErrorOutOfMemory ? MRG8_ERR_NOMEM : MRG8_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

void clear_error() { g_last_error.clear(); }

// ---- host tables -------------------------------------------------------------

const JumpTable& host_jump_table() {
  static const JumpTable t;
  return t;
}

// ---- per-device state ------------------------------------------------------------

namespace {
std::mutex g_dev_mu;
std::vector<std::unique_ptr<DeviceContext>> g_devices;
}  // namespace

int device_context(int device, DeviceContext** out) {
  std::lock_guard<std::mutex> lk(g_dev_mu);
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count <= 0) return fail(MRG8_ERR_NODEVICE, "no HIP device available");
  if (device < 0 || device >= count) return fail(MRG8_ERR_ARG, "device %d out of range (have %d)", device, count);
  if ((int)g_devices.size() < count) g_devices.resize(count);
  if (g_devices[device]) {
    *out = g_devices[device].get();
    return MRG8_OK;
  }
  hipDeviceProp_t prop;
  if ((e = hipGetDeviceProperties(&prop, device)) != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(MRG8_ERR_NODEVICE, "device %d is %s; this library is built for gfx950 only", device, prop.gcnArchName);

  int prev = 0;
  (void)hipGetDevice(&prev);
  if ((e = hipSetDevice(device)) != hipSuccess) return hip_fail(e, "hipSetDevice");

  // Coefficient records: block rows (row 0 of A^(k+1)) and jump rows.
  std::vector<CoefRecord> lane(kLaneRows), jump((size_t)kIdentityRec + 8);
  const std::vector<uint32_t> rows = block_rows(kLaneRows);
  for (int k = 0; k < kLaneRows; ++k) {
    if (!all_nonzero(&rows[(size_t)k * 8])) {
      (void)hipSetDevice(prev);
      return fail(MRG8_ERR_STATE, "block row %d has a zero entry", k);
    }
    lane[k] = make_record(&rows[(size_t)k * 8]);
  }
  // Jump table: A^(d 2^(kJumpBits k)) for digit d = 1..kJumpValues at
  // position k, built from A^(2^(kJumpBits k)) (a table entry) by repeated
  // products.  Rows of the small powers (A^1..A^7) have structural zeros;
  // the support mask keeps their exact-zero results exact.
  const JumpTable& jt = host_jump_table();
  for (int k = 0; k < kJumpPositions; ++k) {
    std::vector<uint32_t> m(jt.at(kJumpBits * k), jt.at(kJumpBits * k) + 64), p = m;
    for (int d = 1; d <= kJumpValues; ++d) {
      if (d > 1) p = mat_mul(p.data(), m.data());
      for (int r = 0; r < 8; ++r)
        jump[((size_t)k * kJumpValues + (d - 1)) * 8 + r] = make_record(&p[(size_t)r * 8]);
    }
  }
  for (int r = 0; r < 8; ++r) {  // the identity: a digit 0 leaves the state as it is
    uint32_t row[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    row[r] = 1;
    jump[(size_t)kIdentityRec + r] = make_record(row);
  }

  // Matrix-core fill tables (operand digits, combine constants, column
  // start records); a fan-out row with a zero entry would need the exact-zero
  // test, so such a table turns the matrix-core fill off.
  MfmaHostTables mt[2];
  MfmaStrideTables mn[2];
  bool mfma_ok = true;
  for (int i = 0; i < 2; ++i)
    mfma_ok = mfma_ok && mfma_host_tables(kMfmaColumns, kMfmaColumnLen[i], kMfmaClasses, kMfmaBeta, &mt[i]) &&
              mfma_stride_tables(host_jump_table(), (uint64_t)kMfmaColumns * kMfmaColumnLen[i], kMfmaClasses, kMfmaBeta,
                                 &mn[i]);

  auto ctx = std::make_unique<DeviceContext>();
  ctx->device = device;
  ctx->num_cu = prop.multiProcessorCount;
  int xcc = 1;
  if (hipDeviceGetAttribute(&xcc, hipDeviceAttributeNumberOfXccs, device) != hipSuccess) {
    (void)hipGetLastError();  // not reported by a later launch check
    xcc = 1;
  }
  if (xcc < 1) xcc = 1;
  ctx->num_xcc = xcc;
  void* p = nullptr;
  const size_t rec_bytes = sizeof(CoefRecord) * (lane.size() + jump.size());
  // operand digits and accumulator start values do not depend on the column length
  size_t mfma_bytes = sizeof(uint32_t) * (mt[0].a.size() + mt[0].c.size());
  for (int i = 0; i < 2; ++i)
    mfma_bytes += sizeof(CoefRecord) * mt[i].fan.size() + sizeof(uint32_t) * (mn[i].a.size() + mn[i].k.size());
  if (!mfma_ok) mfma_bytes = 0;
  const size_t bytes = rec_bytes + mfma_bytes;
  if ((e = hipMalloc(&p, bytes)) != hipSuccess) {
    (void)hipSetDevice(prev);
    return hip_fail(e, "hipMalloc(tables)");
  }
  ctx->tables.lane = static_cast<CoefRecord*>(p);
  ctx->tables.jump = ctx->tables.lane + lane.size();
  e = hipMemcpy(ctx->tables.lane, lane.data(), sizeof(CoefRecord) * lane.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(ctx->tables.jump, jump.data(), sizeof(CoefRecord) * jump.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess && mfma_ok) {
    // records first (16-byte aligned after the 48-byte records), then the
    // operand digits, the accumulator start values and the step tables
    uint8_t* m = static_cast<uint8_t*>(p) + rec_bytes;
    auto put = [&](const void* src, size_t n) {
      uint8_t* at = m;
      if (e == hipSuccess) e = hipMemcpy(at, src, n, hipMemcpyHostToDevice);
      m += n;
      return at;
    };
    for (int i = 0; i < 2; ++i)
      ctx->tables.mfma_fan[i] =
          reinterpret_cast<const CoefRecord*>(put(mt[i].fan.data(), sizeof(CoefRecord) * mt[i].fan.size()));
    ctx->tables.mfma_a = reinterpret_cast<const uint32_t*>(put(mt[0].a.data(), sizeof(uint32_t) * mt[0].a.size()));
    ctx->tables.mfma_c = reinterpret_cast<const int32_t*>(put(mt[0].c.data(), sizeof(int32_t) * mt[0].c.size()));
    for (int i = 0; i < 2; ++i) {
      ctx->tables.mfma_next_a[i] =
          reinterpret_cast<const uint32_t*>(put(mn[i].a.data(), sizeof(uint32_t) * mn[i].a.size()));
      ctx->tables.mfma_next_k[i] =
          reinterpret_cast<const uint32_t*>(put(mn[i].k.data(), sizeof(uint32_t) * mn[i].k.size()));
    }
  }
  if (e != hipSuccess) {
    (void)hipFree(p);
    (void)hipSetDevice(prev);
    return hip_fail(e, "hipMemcpy(tables)");
  }
  // the fixed scratch area of mrg8_batch_generate_noalloc: its address on
  // this device, taken once, with the device current
  int rc = noalloc_scratch_address(&ctx->noalloc_scratch);
  (void)hipSetDevice(prev);
  if (rc) {
    (void)hipFree(p);
    return rc;
  }
  *out = ctx.get();
  g_devices[device] = std::move(ctx);
  return MRG8_OK;
}

DeviceContext* existing_device_context(int device) {
  std::lock_guard<std::mutex> lk(g_dev_mu);
  return device >= 0 && device < (int)g_devices.size() ? g_devices[device].get() : nullptr;
}

int ensure_wide_table(DeviceContext& ctx) {
  std::lock_guard<std::mutex> lk(ctx.wide_mu);
  if (ctx.tables.wide) return MRG8_OK;
  // A^(d 2^(kJumpBits k)) for k = kJumpPositions..kWidePositions-1 and
  // d = 1..kJumpValues; A^(2^j) past the last jump_matrix.txt entry by squaring.
  const JumpTable& jt = host_jump_table();
  const int npos = kWidePositions - kJumpPositions;
  std::vector<CoefRecord> wide((size_t)npos * kJumpValues * 8);
  for (int k = kJumpPositions; k < kWidePositions; ++k) {
    const int j = kJumpBits * k;
    std::vector<uint32_t> m(jt.at(std::min(j, kNumJump - 1)), jt.at(std::min(j, kNumJump - 1)) + 64);
    for (int q = kNumJump - 1; q < j; ++q) m = mat_mul(m.data(), m.data());
    std::vector<uint32_t> p = m;
    for (int d = 1; d <= kJumpValues; ++d) {
      if (d > 1) p = mat_mul(p.data(), m.data());
      for (int r = 0; r < 8; ++r)
        wide[((size_t)(k - kJumpPositions) * kJumpValues + (d - 1)) * 8 + r] = make_record(&p[(size_t)r * 8]);
    }
  }
  DeviceGuard dg(ctx.device);
  void* dev = nullptr;
  const size_t bytes = sizeof(CoefRecord) * wide.size();
  hipError_t e = hipMalloc(&dev, bytes);
  if (e != hipSuccess) return hip_fail(e, "hipMalloc(wide jump table)");
  e = hipMemcpy(dev, wide.data(), bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(dev);
    return hip_fail(e, "hipMemcpy(wide jump table)");
  }
  ctx.tables.wide = static_cast<CoefRecord*>(dev);
  return MRG8_OK;
}

bool state_valid(const uint32_t st[8]) {
  for (int i = 0; i < 8; ++i)
    if (st[i] > kP) return false;
  return true;
}

int format_size(int format) {
  switch (format) {
    case MRG8_F64:
    case MRG8_F64_01: return 8;
    case MRG8_U32: return 4;
    default: return 0;
  }
}

// Default launch shape for a single-stream fill of n outputs.  Measured on
// MI355X (2^28 outputs): doubles are fastest with 512-thread blocks in XCD
// runs (fill_xcd_group), 4 outputs per lane and 2048 outputs per wave, words
// with 256-thread blocks and 8192 per wave; small fills shrink the per-wave
// run (down to four iterations) so that every SIMD still gets waves.
void default_shape(const DeviceContext& ctx, uint64_t n, int format, int per_lane, int* block, uint64_t* tile) {
  const uint64_t quantum = 64ull * per_lane;
  uint64_t t = format == MRG8_U32 ? 8192 : 2048;
  const uint64_t cu = (uint64_t)ctx.num_cu;
  const uint64_t min_waves = cu * 4 * 4;  // four waves per SIMD
  while (t > 4 * quantum && (n + t - 1) / t < min_waves) t /= 2;
  *tile = std::max<uint64_t>(quantum, t / quantum * quantum);
  // Double fills take 512-thread blocks, and 256-thread blocks when every CU
  // gets at most 4 waves (on MI355X, against 1024-thread blocks: 2^20 and
  // 2^21 doubles 15-44 % faster, 2^22..2^25 0-4 % faster; from 2^25 on in
  // XCD runs, 2^26 the same, 2^27 1 % and 2^28 1-2 % faster at full clock
  // and 3.5-5 % faster in the first launches after an idle spell, while the
  // power manager holds the clock down).
  int b = format == MRG8_U32 ? 256 : 512;
  const uint64_t waves = (n + *tile - 1) / *tile;
  if (format != MRG8_U32 && waves <= cu * 4) b = 256;
  *block = b;
}

// Blocks per XCD run of a vector-kernel fill (1: blocks in dispatch order).
// The dispatcher deals blocks to the XCDs in turn; a double fill of 64 or
// more waves per CU instead gives each XCD runs of 32768 threads' worth of
// consecutive output (64 blocks of 512 threads).  Measured on MI355X, 2^28
// doubles at 2048 per wave: 512-thread blocks 366 us in dispatch order and
// 337-341 us in runs of 64; 256-thread blocks 391 and 340 us (runs of 128);
// 1024-thread blocks 338-349 us either way.  Word fills (8192 per wave) were
// 1-17 % slower in runs, and are left in dispatch order.
uint32_t fill_xcd_group(const DeviceContext& ctx, int format, uint64_t waves, int block_threads) {
  if (format == MRG8_U32 || ctx.num_xcc <= 1 || waves < (uint64_t)ctx.num_cu * 64) return 1;
  return xcd_run_blocks(ctx, block_threads);
}

namespace {

// Matrix-core fill of n outputs: whole tiles of 16 x `column` outputs from
// the first 1 KiB boundary of dptr on, `outputs_per_wave` / tile consecutive
// tiles per wave (0: 32 KiB of output) in blocks of `block_threads` (0:
// the format's default), the outputs before the first tile and after the last
// by two extra waves.  Returns MRG8_ERR_STATE (and launches nothing) when the
// fill has no whole tile or the state is all zero: the caller takes the
// vector kernel.
int fill_mfma(DeviceContext& ctx, const uint32_t state[8], uint64_t offset, void* dptr, uint64_t n, int format,
              hipStream_t stream, int column, int block_threads, uint64_t outputs_per_wave) {
  const int esz = format_size(format);
  uint32_t any = 0;
  for (int j = 0; j < 8; ++j) any |= state[j];
  if (!ctx.tables.mfma_a || any == 0) return MRG8_ERR_STATE;
  const uint64_t align = 1024;
  uint64_t head = ((align - (reinterpret_cast<uintptr_t>(dptr) & (align - 1))) & (align - 1)) / (uint64_t)esz;
  if (head >= n) return MRG8_ERR_STATE;
  const uint64_t tile = (uint64_t)kMfmaColumns * column;
  const uint64_t tiles = (n - head) / tile;
  if (tiles == 0) return MRG8_ERR_STATE;
  const int ci = column == 64 ? 0 : 1;
  MfmaFillArgs a{};
  a.out = dptr;
  a.head = head;
  a.tiles = tiles;
  a.column = (uint32_t)column;
  a.waves_per_block = (uint32_t)(block_threads ? block_threads / 64 : format == MRG8_U32 ? 8 : 16);
  // default: 32 KiB of output per wave, fewer tiles when that would leave
  // fewer than 8 waves per CU
  const uint64_t per = outputs_per_wave
                           ? outputs_per_wave / tile
                           : std::min<uint64_t>(32768 / (uint64_t)esz / tile, tiles / ((uint64_t)ctx.num_cu * 8));
  a.tiles_per_wave = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(per, 1), 1u << 20);
  a.waves = (tiles + a.tiles_per_wave - 1) / a.tiles_per_wave;
  const uint64_t tail_first = head + tiles * tile;
  if (head) {
    a.extra_first[a.extra] = 0;
    a.extra_count[a.extra++] = head;
  }
  if (tail_first < n) {
    a.extra_first[a.extra] = tail_first;
    a.extra_count[a.extra++] = n - tail_first;
  }
  std::memcpy(a.start.state, state, sizeof a.start.state);
  a.start.offset = offset;
  a.lane = ctx.tables.lane;
  a.jump = ctx.tables.jump;
  a.mfma_a = ctx.tables.mfma_a;
  a.mfma_c = ctx.tables.mfma_c;
  a.mfma_fan = ctx.tables.mfma_fan[ci];
  a.mfma_next_a = ctx.tables.mfma_next_a[ci];
  a.mfma_next_k = ctx.tables.mfma_next_k[ci];
  DeviceGuard dg(ctx.device);
  const hipError_t e = launch_fill_mfma(a, format, stream);
  if (e != hipSuccess) return hip_fail(e, "matrix-core fill launch");
  return MRG8_OK;
}

}  // namespace

int fill_device(DeviceContext& ctx, const uint32_t state[8], uint64_t offset, void* dptr, uint64_t n, int format,
                hipStream_t stream, int block_threads, uint64_t tile, int per_lane) {
  if (n == 0) return MRG8_OK;
  const int esz = format_size(format);
  const int cls = 63 - __builtin_clzll(n);
  TunedShape tuned;
  {
    std::lock_guard<std::mutex> lk(ctx.tuned_mu);
    tuned = ctx.tuned[format][cls];
  }
  // The shape: the caller's, else the tuned one for this size class, else
  // the default (the matrix-core fill for large word fills, the vector
  // kernel otherwise).
  int pl = per_lane;
  if (!pl) {
    if (tuned.valid) {
      pl = tuned.per_lane;
      if (block_threads <= 0) block_threads = tuned.block_threads;
      if (tile == 0) tile = tuned.tile;
    } else if (format == MRG8_U32 && n >= kMfmaDefaultMinWords && block_threads <= 0 && tile == 0) {
      pl = 32;
    }
  }
  if (pl == 16 || pl == 32) {  // the matrix-core fill, columns of 4 pl outputs
    const int column = 4 * pl;
    if (block_threads > 0 && block_threads != 256 && block_threads != 512 && block_threads != 1024)
      return fail(MRG8_ERR_ARG, "the matrix-core fill takes blocks of 256, 512 or 1024 threads");
    if (tile % ((uint64_t)kMfmaColumns * column))
      return fail(MRG8_ERR_ARG, "the matrix-core fill takes whole %d-output tiles per wave", kMfmaColumns * column);
    const int rc = fill_mfma(ctx, state, offset, dptr, n, format, stream, column, std::max(block_threads, 0), tile);
    if (rc != MRG8_ERR_STATE) return rc;
    pl = 0;  // too short for a tile, or the all-zero state: the vector kernel's default shape
    block_threads = 0;
    tile = 0;
  }
  // 4 outputs per lane (16-byte stores) whatever the destination's alignment:
  // each wave writes the outputs before its first 1 KiB boundary one per lane.
  if (pl && pl != 1 && pl != 2 && pl != 4) return fail(MRG8_ERR_ARG, "outputs_per_lane must be 1, 2, 4, 16 or 32");
  if (!pl) pl = 4;
  int def_block = 0;
  uint64_t def_tile = 0;
  default_shape(ctx, n, format, pl, &def_block, &def_tile);
  if (block_threads <= 0) block_threads = def_block;
  if (block_threads < 64 || block_threads > 1024 || block_threads % 64)
    return fail(MRG8_ERR_ARG, "block_threads %d must be a multiple of 64 in [64, 1024]", block_threads);
  const uint64_t quantum = 64ull * pl;
  if (tile == 0) tile = def_tile;
  if (tile > n) tile = n;  // one wave takes everything (and the rounding below cannot wrap)
  if (tile % quantum) tile = (tile + quantum - 1) / quantum * quantum;
  // Wave 0 also takes the outputs before the first 1 KiB boundary, so every
  // other wave starts on one and writes whole cache lines (a 2^28-double fill
  // 8 bytes off a boundary took 471 us with waves cut from the destination
  // as given, 346-359 us so; a head up to the next 16 KiB boundary left wave
  // 0 with up to two tiles of work and the fill 3 % slower).
  const uint64_t align = 1024;
  uint64_t head = ((align - (reinterpret_cast<uintptr_t>(dptr) & (align - 1))) & (align - 1)) / (uint64_t)esz;
  if (head > n) head = n;
  // The grid (work-items in total) is a 32-bit quantity.
  const uint64_t wpb = (uint64_t)block_threads / 64;
  if ((fill_waves(n, head, tile) + wpb - 1) / wpb * (uint64_t)block_threads > 0xffffffffull)
    return fail(MRG8_ERR_ARG, "%llu outputs at %llu per wave need more than 2^32 - 1 work-items; use more outputs per wave",
                (unsigned long long)n, (unsigned long long)tile);
  FillArgs a{};
  a.out = dptr;
  a.n = n;
  a.tile = tile;
  a.head = head;
  std::memcpy(a.start.state, state, sizeof a.start.state);
  a.start.offset = offset;
  a.lane = ctx.tables.lane;
  a.jump = ctx.tables.jump;
  a.xcds = (uint32_t)ctx.num_xcc;
  a.xcd_group = fill_xcd_group(ctx, format, fill_waves(n, head, tile), block_threads);
  int prev = 0;
  (void)hipGetDevice(&prev);
  if (prev != ctx.device) (void)hipSetDevice(ctx.device);
  hipError_t e = launch_fill(a, format, pl, block_threads, stream);
  if (prev != ctx.device) (void)hipSetDevice(prev);
  if (e != hipSuccess) return hip_fail(e, "fill launch");
  return MRG8_OK;
}

}  // namespace mrg8

using namespace mrg8;

struct mrg8_gen {
  DeviceContext* ctx = nullptr;
  uint32_t state[8] = {0};
  uint32_t seed = 0;
  // streams generated on since the last mrg8_synchronize
  std::vector<hipStream_t> streams;
  // host-memory forms: two pinned host chunks, a stream and one event per
  // chunk (allocated on first use)
  void* pin_chunk[2] = {nullptr, nullptr};
  hipEvent_t chunk_ready[2] = {nullptr, nullptr};
  hipStream_t host_stream = nullptr;
};

namespace {

void release_host_staging(mrg8_gen* g) {
  if (g->host_stream) (void)hipStreamSynchronize(g->host_stream);
  for (int i = 0; i < 2; ++i) {
    if (g->pin_chunk[i]) (void)hipHostFree(g->pin_chunk[i]);
    if (g->chunk_ready[i]) (void)hipEventDestroy(g->chunk_ready[i]);
    g->pin_chunk[i] = nullptr;
    g->chunk_ready[i] = nullptr;
  }
  if (g->host_stream) (void)hipStreamDestroy(g->host_stream);
  g->host_stream = nullptr;
}

}  // namespace

// ---- library ----------------------------------------------------------------

extern "C" int mrg8_abi_version(void) { return MRG8_ABI_VERSION; }

extern "C" const char* mrg8_strerror(int status) {
  switch (status) {
    case MRG8_OK: return "ok";
    case MRG8_ERR_ARG: return "invalid argument";
    case MRG8_ERR_HIP: return "HIP runtime error";
    case MRG8_ERR_NOMEM: return "out of memory";
    case MRG8_ERR_NODEVICE: return "no usable gfx950 device";
    case MRG8_ERR_CAPACITY: return "capacity exceeded";
    case MRG8_ERR_STATE: return "invalid state for this call";
    case MRG8_ERR_NOTFOUND: return "not found";
    case MRG8_ERR_IO: return "I/O error";
    case MRG8_ERR_BLOCKED: return "blocked until completed requests are released";
    default: return "unknown status";
  }
}

extern "C" const char* mrg8_last_error(void) { return g_last_error.c_str(); }

extern "C" int mrg8_init_device(int device) {
  clear_error();
  DeviceContext* ctx = nullptr;
  return device_context(device, &ctx);
}

// ---- generator ----------------------------------------------------------------

extern "C" int mrg8_create(uint32_t seed, mrg8_gen_t* out) {
  clear_error();
  if (!out) return fail(MRG8_ERR_ARG, "null output handle");
  *out = nullptr;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(MRG8_ERR_NODEVICE, "hipGetDevice: %s", hipGetErrorString(e));
  DeviceContext* ctx = nullptr;
  int rc = device_context(dev, &ctx);
  if (rc) return rc;
  mrg8_gen* g = new (std::nothrow) mrg8_gen;
  if (!g) return fail(MRG8_ERR_NOMEM, "host allocation failed");
  g->ctx = ctx;
  g->seed = seed;
  seed_state(seed, g->state);
  *out = g;
  return MRG8_OK;
}

extern "C" int mrg8_destroy(mrg8_gen_t g) {
  clear_error();
  if (!g) return MRG8_OK;
  // Kernels in flight read only their arguments and the device tables, which
  // live until the process ends: nothing of `g` needs to outlive this call
  // except the host-form staging, which is drained first.
  if (g->host_stream) {
    DeviceGuard dg(g->ctx->device);
    release_host_staging(g);
  }
  delete g;
  return MRG8_OK;
}

extern "C" int mrg8_seed_init(mrg8_gen_t g, uint32_t seed) {
  clear_error();
  if (!g) return fail(MRG8_ERR_ARG, "null generator");
  g->seed = seed;
  seed_state(seed, g->state);
  return MRG8_OK;
}

extern "C" int mrg8_get_state(mrg8_gen_t g, uint32_t state[8]) {
  clear_error();
  if (!g || !state) return fail(MRG8_ERR_ARG, "null argument");
  std::memcpy(state, g->state, sizeof g->state);
  return MRG8_OK;
}

extern "C" int mrg8_set_state(mrg8_gen_t g, const uint32_t state[8]) {
  clear_error();
  if (!g || !state) return fail(MRG8_ERR_ARG, "null argument");
  if (!state_valid(state)) return fail(MRG8_ERR_ARG, "state words must be at most 2^31 - 1");
  std::memcpy(g->state, state, sizeof g->state);
  return MRG8_OK;
}

extern "C" int mrg8_jump_ahead(mrg8_gen_t g, uint64_t step) {
  clear_error();
  if (!g) return fail(MRG8_ERR_ARG, "null generator");
  jump64(host_jump_table(), step, g->state);
  return MRG8_OK;
}

extern "C" int mrg8_jump_ahead_bits(mrg8_gen_t g, const short bits[200]) {
  clear_error();
  if (!g || !bits) return fail(MRG8_ERR_ARG, "null argument");
  jump200(host_jump_table(), bits, g->state);
  return MRG8_OK;
}

extern "C" int mrg8_jump_state(mrg8_gen_t g, uint64_t step, uint32_t new_state[8]) {
  clear_error();
  if (!g || !new_state) return fail(MRG8_ERR_ARG, "null argument");
  uint32_t st[8];
  std::memcpy(st, g->state, sizeof st);
  jump64(host_jump_table(), step, st);
  std::memcpy(new_state, st, sizeof st);
  return MRG8_OK;
}

extern "C" int mrg8_jump_state_bits(mrg8_gen_t g, const short bits[200], uint32_t new_state[8]) {
  clear_error();
  if (!g || !bits || !new_state) return fail(MRG8_ERR_ARG, "null argument");
  uint32_t st[8];
  std::memcpy(st, g->state, sizeof st);
  jump200(host_jump_table(), bits, st);
  std::memcpy(new_state, st, sizeof st);
  return MRG8_OK;
}

extern "C" int mrg8_seed_state(uint32_t seed, uint32_t state[8]) {
  clear_error();
  if (!state) return fail(MRG8_ERR_ARG, "null state");
  seed_state(seed, state);
  return MRG8_OK;
}

extern "C" int mrg8_advance_state(const uint32_t state[8], uint64_t step, uint32_t new_state[8]) {
  clear_error();
  if (!state || !new_state) return fail(MRG8_ERR_ARG, "null argument");
  if (!state_valid(state)) return fail(MRG8_ERR_ARG, "state words must be at most 2^31 - 1");
  uint32_t st[8];
  std::memcpy(st, state, sizeof st);
  jump64(host_jump_table(), step, st);
  std::memcpy(new_state, st, sizeof st);
  return MRG8_OK;
}

extern "C" int mrg8_advance_state_bits(const uint32_t state[8], const short bits[200], uint32_t new_state[8]) {
  clear_error();
  if (!state || !bits || !new_state) return fail(MRG8_ERR_ARG, "null argument");
  if (!state_valid(state)) return fail(MRG8_ERR_ARG, "state words must be at most 2^31 - 1");
  uint32_t st[8];
  std::memcpy(st, state, sizeof st);
  jump200(host_jump_table(), bits, st);
  std::memcpy(new_state, st, sizeof st);
  return MRG8_OK;
}

static int generate_impl(mrg8_gen_t g, void* dptr, uint64_t n, int format, void* stream, int block_threads,
                         uint64_t tile, int per_lane) {
  if (!g) return fail(MRG8_ERR_ARG, "null generator");
  const int esz = format_size(format);
  if (!esz) return fail(MRG8_ERR_ARG, "unknown format %d", format);
  if (n == 0) return MRG8_OK;
  if (!dptr) return fail(MRG8_ERR_ARG, "null output pointer");
  if (reinterpret_cast<uintptr_t>(dptr) % (uintptr_t)esz)
    return fail(MRG8_ERR_ARG, "output pointer must be %d-byte aligned", esz);
  if (n > (UINT64_MAX / 8)) return fail(MRG8_ERR_ARG, "n too large");
  hipStream_t s = static_cast<hipStream_t>(stream);
  int rc = fill_device(*g->ctx, g->state, 0, dptr, n, format, s, block_threads, tile, per_lane);
  if (rc) return rc;
  if (std::find(g->streams.begin(), g->streams.end(), s) == g->streams.end()) g->streams.push_back(s);
  jump64(host_jump_table(), n, g->state);  // the state after n outputs
  return MRG8_OK;
}

extern "C" int mrg8_generate(mrg8_gen_t g, void* dptr, uint64_t n, int format, void* stream) {
  clear_error();
  return generate_impl(g, dptr, n, format, stream, 0, 0, 0);
}

extern "C" int mrg8_generate_shape(mrg8_gen_t g, void* dptr, uint64_t n, int format, void* stream,
                                   int block_threads, uint64_t outputs_per_wave, int outputs_per_lane) {
  clear_error();
  return generate_impl(g, dptr, n, format, stream, block_threads, outputs_per_wave, outputs_per_lane);
}

extern "C" int mrg8_synchronize(mrg8_gen_t g) {
  clear_error();
  if (!g) return fail(MRG8_ERR_ARG, "null generator");
  if (g->streams.empty()) return MRG8_OK;
  DeviceGuard dg(g->ctx->device);
  for (hipStream_t s : g->streams) {
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
  }
  g->streams.clear();
  return MRG8_OK;
}

// The reference's launchers take int n, BS (threads per block) and TNUM
// (total threads); they block until the kernel is done.  TNUM = 0 takes the
// library's own launch shape, BS = 0 the reference's 64 threads per block.
static int drop_in_fill(mrg8_gen_t g, double* d, int n, int bs, int tnum) {
  if (!g) return fail(MRG8_ERR_ARG, "null generator");
  if (n < 0) return fail(MRG8_ERR_ARG, "n must be non-negative");
  if (bs < 0 || tnum < 0) return fail(MRG8_ERR_ARG, "BS and TNUM must be non-negative");
  if (n == 0) return MRG8_OK;
  if (tnum == 0) {
    int rc = generate_impl(g, d, (uint64_t)n, MRG8_F64_01, nullptr, 0, 0, 0);
    if (rc) return rc;
    return mrg8_synchronize(g);
  }
  if (bs == 0) bs = 64;
  // TNUM threads in waves of 64: each wave covers n / (TNUM / 64) outputs.
  const uint64_t waves = std::max<uint64_t>(1, (uint64_t)tnum / 64);
  const uint64_t quantum = 64ull * 4;  // 4 outputs per lane
  uint64_t tile = ((uint64_t)n + waves - 1) / waves;
  tile = std::max<uint64_t>(quantum, (tile + quantum - 1) / quantum * quantum);
  int block = bs < 64 ? 64 : (bs > 1024 ? 1024 : bs / 64 * 64);
  int rc = generate_impl(g, d, (uint64_t)n, MRG8_F64_01, nullptr, block, tile, 0);
  if (rc) return rc;
  return mrg8_synchronize(g);
}

extern "C" int mrg8_outer_32(mrg8_gen_t g, double* d_ran, int n, int BS, int TNUM) {
  clear_error();
  return drop_in_fill(g, d_ran, n, BS, TNUM);
}

extern "C" int mrg8_inner(mrg8_gen_t g, double* d_ran, int n, int TNUM) {
  clear_error();
  return drop_in_fill(g, d_ran, n, 64, TNUM);
}

extern "C" int mrg8_state_cpy_DtH(mrg8_gen_t g) {
  clear_error();
  return mrg8_synchronize(g);
}

// ---- host-memory forms ------------------------------------------------------------

// Copies `bytes` from `src` to `dst` with up to 8 host threads (the pinned
// chunk to the caller's pageable memory).
static void parallel_copy(void* dst, const void* src, size_t bytes) {
  const size_t min_part = (size_t)4 << 20;
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const size_t parts = std::min<size_t>(std::min(8u, hw), std::max<size_t>(1, bytes / min_part));
  if (parts <= 1) {
    std::memcpy(dst, src, bytes);
    return;
  }
  const size_t per = (bytes / parts + 63) / 64 * 64;
  std::vector<std::thread> th;
  for (size_t i = 1; i < parts; ++i) {
    const size_t off = i * per;
    if (off >= bytes) break;
    th.emplace_back([=] {
      std::memcpy(static_cast<char*>(dst) + off, static_cast<const char*>(src) + off, std::min(per, bytes - off));
    });
  }
  std::memcpy(dst, src, std::min(per, bytes));
  for (auto& t : th) t.join();
}

// Host-memory forms: the kernel writes each chunk straight into a pinned host
// buffer over PCIe, two chunks in flight: while the host copies chunk i into
// the caller's memory, the device writes chunk i + 1.  Measured on MI355X:
// ~51 GB/s for the kernel's own writes to host memory whatever stream it runs
// on, where a device chunk followed by hipMemcpyAsync ran at 31-54 GB/s
// depending on the stream.  The chunks are coherent host memory, which the
// host reads back about 4x faster than the default kind after device writes.
static int host_fill(mrg8_gen_t g, void* host, uint64_t n, int format) {
  if (!g) return fail(MRG8_ERR_ARG, "null generator");
  if (n == 0) return MRG8_OK;
  if (!host) return fail(MRG8_ERR_ARG, "null output pointer");
  const size_t esz = (size_t)format_size(format);
  const size_t chunk_bytes = (size_t)32 << 20;
  DeviceGuard dg(g->ctx->device);
  hipError_t e = hipSuccess;
  if (!g->host_stream) {
    e = hipStreamCreateWithFlags(&g->host_stream, hipStreamNonBlocking);
    for (int i = 0; i < 2 && e == hipSuccess; ++i) {
      e = hipHostMalloc(&g->pin_chunk[i], chunk_bytes, hipHostMallocCoherent);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&g->chunk_ready[i], hipEventDisableTiming);
    }
    if (e != hipSuccess) {
      release_host_staging(g);
      return hip_fail(e, "host staging allocation");
    }
  }
  const uint64_t per = chunk_bytes / esz;
  const uint64_t nchunks = (n + per - 1) / per;
  auto count = [&](uint64_t i) { return std::min<uint64_t>(per, n - i * per); };
  auto issue = [&](uint64_t i) -> int {
    const int b = (int)(i & 1);
    const uint64_t m = count(i);
    int rc = fill_device(*g->ctx, g->state, 0, g->pin_chunk[b], m, format, g->host_stream, 0, 0, 0);
    if (rc) return rc;
    jump64(host_jump_table(), m, g->state);
    hipError_t x = hipEventRecord(g->chunk_ready[b], g->host_stream);
    return x == hipSuccess ? MRG8_OK : hip_fail(x, "hipEventRecord(chunk)");
  };
  int rc = issue(0);
  if (rc == MRG8_OK && nchunks > 1) rc = issue(1);
  for (uint64_t i = 0; i < nchunks && rc == MRG8_OK; ++i) {
    const int b = (int)(i & 1);
    e = hipEventSynchronize(g->chunk_ready[b]);
    if (e != hipSuccess) {
      rc = hip_fail(e, "hipEventSynchronize(chunk)");
      break;
    }
    parallel_copy(static_cast<char*>(host) + i * per * esz, g->pin_chunk[b], count(i) * esz);
    if (i + 2 < nchunks) rc = issue(i + 2);
  }
  if (rc) (void)hipStreamSynchronize(g->host_stream);  // leave nothing in flight on an error
  return rc;
}

extern "C" int mrg8_randint(mrg8_gen_t g, uint32_t* host, uint64_t n) {
  clear_error();
  return host_fill(g, host, n, MRG8_U32);
}

extern "C" int mrg8_rand(mrg8_gen_t g, double* host, uint64_t n) {
  clear_error();
  return host_fill(g, host, n, MRG8_F64);
}

extern "C" int mrg8_mrg8dnz2(mrg8_gen_t g, double* host, uint64_t n) {
  clear_error();
  // cuda/mrg8.cpp:290-339 produces the same words and conversion as rand():
  // its 1024-word blocks only change the order of evaluation.
  return host_fill(g, host, n, MRG8_F64);
}

extern "C" int mrg8_randint1(mrg8_gen_t g, uint32_t* value) {
  clear_error();
  if (!g || !value) return fail(MRG8_ERR_ARG, "null argument");
  *value = step1(g->state);
  return MRG8_OK;
}

extern "C" int mrg8_rand1(mrg8_gen_t g, double* value) {
  clear_error();
  if (!g || !value) return fail(MRG8_ERR_ARG, "null argument");
  const double rnorm = 1.0 / 2147483647.0;
  *value = (double)step1(g->state) * rnorm;
  return MRG8_OK;
}

// ---- bulk advance of device-resident states ----------------------------------------

extern "C" int mrg8_advance_states(int device, const uint32_t* d_in, const uint64_t* d_offsets, uint32_t words,
                                   uint32_t* d_out, uint64_t n, void* stream) {
  clear_error();
  if (n == 0) return MRG8_OK;
  if (!d_in || !d_offsets || !d_out) return fail(MRG8_ERR_ARG, "null argument");
  if (words < 1 || words > (uint32_t)kAdvanceWords)
    return fail(MRG8_ERR_ARG, "words must be 1..%d (offsets of 64..%d bits)", kAdvanceWords, 64 * kAdvanceWords);
  if ((reinterpret_cast<uintptr_t>(d_in) | reinterpret_cast<uintptr_t>(d_out)) & 3u ||
      reinterpret_cast<uintptr_t>(d_offsets) & 7u)
    return fail(MRG8_ERR_ARG, "states must be 4-byte and offsets 8-byte aligned");
  if (n > UINT64_MAX / (8ull * words)) return fail(MRG8_ERR_ARG, "n too large");
  DeviceContext* ctx = nullptr;
  int rc = device_context(device, &ctx);
  if (rc) return rc;
  if (words > 1 && (rc = ensure_wide_table(*ctx)) != MRG8_OK) return rc;
  DeviceGuard dg(device);
  const uint64_t chunk = 1ull << 28;  // 8 lanes per state: 2^31 work-items per launch at most
  for (uint64_t first = 0; first < n; first += chunk) {
    AdvanceArgs a{};
    a.in = d_in + 8 * first;
    a.offsets = d_offsets + words * first;
    a.out = d_out + 8 * first;
    a.n = std::min(chunk, n - first);
    a.words = words;
    a.jump = ctx->tables.jump;
    a.wide = ctx->tables.wide;
    hipError_t e = launch_advance(a, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "advance launch");
  }
  return MRG8_OK;
}

// ---- jump table output -----------------------------------------------------------

extern "C" int mrg8_jump_matrix(uint32_t* table, size_t words) {
  clear_error();
  if (!table) return fail(MRG8_ERR_ARG, "null table");
  if (words < (size_t)kNumJump * 64) return fail(MRG8_ERR_ARG, "table needs %d words", kNumJump * 64);
  const JumpTable& t = host_jump_table();
  for (int j = 0; j < kNumJump; ++j)
    for (int r = 0; r < 8; ++r)
      for (int c = 0; c < 8; ++c) table[(size_t)j * 64 + r + 8 * c] = t.at(j)[r * 8 + c];  // column-major
  return MRG8_OK;
}

extern "C" int mrg8_read_jump_matrix(const char* path, uint32_t* table, size_t words) {
  clear_error();
  if (!path || !table) return fail(MRG8_ERR_ARG, "null argument");
  if (words < (size_t)kNumJump * 64) return fail(MRG8_ERR_ARG, "table needs %d words", kNumJump * 64);
  FILE* f = std::fopen(path, "r");
  if (!f) return fail(MRG8_ERR_IO, "cannot open %s", path);
  // original/mrg8.cpp:88-102: kNumJump lines of 64 unsigned words (column-major 8x8)
  int rc = MRG8_OK;
  for (size_t i = 0; i < (size_t)kNumJump * 64; ++i) {
    unsigned long long v = 0;
    if (std::fscanf(f, "%llu", &v) != 1) {
      rc = fail(MRG8_ERR_IO, "%s: word %zu of %d missing or malformed", path, i, kNumJump * 64);
      break;
    }
    if (v > kP) {
      rc = fail(MRG8_ERR_IO, "%s: word %zu (%llu) is above 2^31 - 1", path, i, v);
      break;
    }
    table[i] = (uint32_t)v;
  }
  std::fclose(f);
  return rc;
}

extern "C" int mrg8_check_jump_matrix(const char* path) {
  clear_error();
  std::vector<uint32_t> file((size_t)kNumJump * 64), mine((size_t)kNumJump * 64);
  int rc = mrg8_read_jump_matrix(path, file.data(), file.size());
  if (rc) return rc;
  mrg8_jump_matrix(mine.data(), mine.size());
  for (size_t i = 0; i < file.size(); ++i)
    if (file[i] != mine[i])
      return fail(MRG8_ERR_STATE, "%s: word %zu is %u, A^(2^%zu) has %u", path, i, file[i], i / 64, mine[i]);
  return MRG8_OK;
}

extern "C" int mrg8_write_jump_matrix(const char* path) {
  clear_error();
  if (!path) return fail(MRG8_ERR_ARG, "null path");
  std::vector<uint32_t> tab((size_t)kNumJump * 64);
  mrg8_jump_matrix(tab.data(), tab.size());
  FILE* f = std::fopen(path, "w");
  if (!f) return fail(MRG8_ERR_IO, "cannot open %s", path);
  for (int j = 0; j < kNumJump; ++j) {
    for (int k = 0; k < 64; ++k) std::fprintf(f, "%10u  ", tab[(size_t)j * 64 + k]);
    std::fputc('\n', f);
  }
  if (std::fclose(f) != 0) return fail(MRG8_ERR_IO, "cannot write %s", path);
  return MRG8_OK;
}

// ---- fill autotuning -------------------------------------------------------------

namespace {

std::string fill_key(const DeviceContext& ctx, int format, int cls) {
  hipDeviceProp_t prop;
  (void)hipGetDeviceProperties(&prop, ctx.device);
  char buf[256];
  std::snprintf(buf, sizeof buf, "%s:%d:fill:fmt%d:n2^%d", prop.gcnArchName, prop.multiProcessorCount, format, cls);
  std::string k = buf;
  for (char& c : k)
    if (c == ' ') c = '_';
  return k;
}

}  // namespace

extern "C" int mrg8_tune_fill(int device, int format, uint64_t n, const char* cache_path, mrg8_fill_shape* best) {
  clear_error();
  const int esz = format_size(format);
  if (!esz) return fail(MRG8_ERR_ARG, "unknown format %d", format);
  if (n < 4096) return fail(MRG8_ERR_ARG, "tune with at least 4096 outputs");
  DeviceContext* ctx = nullptr;
  int rc = device_context(device, &ctx);
  if (rc) return rc;
  const int cls = 63 - __builtin_clzll(n);
  const std::string key = fill_key(*ctx, format, cls);
  TunedShape found;
  if (cache_path) {
    std::FILE* f = std::fopen(cache_path, "r");
    if (f) {
      char k[256];
      int bt = 0, pl = 0;
      unsigned long long tile = 0;
      double ms = 0;
      while (std::fscanf(f, "%255s %d %d %llu %lf", k, &bt, &pl, &tile, &ms) == 5)
        if (key == k && bt >= 64 && bt <= 1024 && bt % 64 == 0 && (pl == 1 || pl == 2 || pl == 4 || pl == 16 || pl == 32) &&
            tile) {
          found.valid = true;
          found.block_threads = bt;
          found.per_lane = pl;
          found.tile = tile;
        }
      std::fclose(f);
    }
  }
  if (!found.valid) {
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(device);
    void* buf = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&buf, n * (uint64_t)esz);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e != hipSuccess) {
      if (buf) (void)hipFree(buf);
      (void)hipSetDevice(prev);
      return hip_fail(e, "tuning allocation");
    }
    uint32_t st[8];
    seed_state(13579u, st);
    double best_ms = 1e30;
    const int blocks[] = {256, 512, 1024};
    const uint64_t tiles[] = {2048, 4096, 8192, 16384};
    // vector kernel: 1, 2 or 4 outputs per lane per iteration; matrix-core
    // fill: 16 or 32 outputs per lane per tile (columns of 64 or 128)
    const int pls[] = {1, 2, 4, 16, 32};
    for (int pl : pls) {
      for (int bt : blocks)
        for (uint64_t tile : tiles) {
          if (pl <= 4 ? tile < 64ull * pl * 8 : (bt == 256 || tile % (64ull * pl) || tile * esz > 65536)) continue;
          rc = fill_device(*ctx, st, 0, buf, n, format, nullptr, bt, tile, pl);  // warm-up
          if (rc) break;
          (void)hipEventRecord(e0, nullptr);
          const int reps = 4;
          for (int r = 0; r < reps && rc == MRG8_OK; ++r) rc = fill_device(*ctx, st, 0, buf, n, format, nullptr, bt, tile, pl);
          (void)hipEventRecord(e1, nullptr);
          (void)hipEventSynchronize(e1);
          float ms = 0;
          (void)hipEventElapsedTime(&ms, e0, e1);
          ms /= reps;
          if (rc == MRG8_OK && ms < best_ms) {
            best_ms = ms;
            found.valid = true;
            found.block_threads = bt;
            found.per_lane = pl;
            found.tile = tile;
          }
        }
      if (rc) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(buf);
    (void)hipSetDevice(prev);
    if (rc) return rc;
    if (cache_path && found.valid) {
      std::FILE* f = std::fopen(cache_path, "a");
      if (f) {
        std::fprintf(f, "%s %d %d %llu %.6f\n", key.c_str(), found.block_threads, found.per_lane,
                     (unsigned long long)found.tile, best_ms);
        std::fclose(f);
      }
    }
  }
  {
    std::lock_guard<std::mutex> lk(ctx->tuned_mu);
    ctx->tuned[format][cls] = found;
  }
  if (best) {
    best->block_threads = found.block_threads;
    best->outputs_per_lane = found.per_lane;
    best->outputs_per_wave = found.tile;
  }
  return MRG8_OK;
}
