// Serving: per-stream states resident in device memory, requests admitted by
// output-buffer memory, continuous batching over several in-flight launches.
//
// Life of a request:
//   submit   -> QUEUED: positions [start, start+count) of the stream are
//               reserved at once (cursor += count), so the request's outputs
//               are fixed no matter what else is in flight.
//   admitted -> ACTIVE: once a request slot and `count * size` bytes of the
//               output arena are free (FIFO order).
//   issued   -> every output is part of a launched step; the request is
//               complete when the step holding its last chunk finishes.
//   released -> its arena bytes return to the pool.
// A step is one persistent batched launch (BatchSlot) over chunks of the
// active requests, up to `step_outputs` outputs and at most a quarter of a
// step per request.  Service order within a step: requests admitted more
// than `age_limit_us` ago first (admission order), then the rest by fewest
// outputs left, so small requests do not queue behind large ones and large
// ones cannot starve.  Up to `steps_in_flight` steps are in flight, each on
// its own stream, so that a step's staging copy and launch overlap the
// steps before it and a step starts while the previous one's last waves
// drain; a request arriving while they run joins the next one.  Steps launch
// the kernel directly: every step has a new block, and a plain launch costs
// the host less than a graph replay.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/mrg8.h"
#include "mrg8_internal.hpp"
#include "mrg8_slot.hpp"

using namespace mrg8;

namespace {

constexpr uint32_t kMaxSteps = 8;
constexpr size_t kAlign = 256;

size_t granule_bytes(size_t bytes) { return (bytes + kAlign - 1) / kAlign * kAlign; }

double wall_now() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// First-fit allocator over one device block, with coalescing frees.
class Arena {
 public:
  void init(size_t cap) {
    cap_ = cap;
    free_.clear();
    if (cap) free_[0] = cap;
    used_ = 0;
  }
  bool alloc(size_t bytes, size_t* off) {
    bytes = granule_bytes(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      *off = it->first;
      const size_t rest = it->second - bytes, at = it->first + bytes;
      free_.erase(it);
      if (rest) free_[at] = rest;
      used_ += bytes;
      return true;
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = granule_bytes(bytes);
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t used() const { return used_; }
  size_t capacity() const { return cap_; }

 private:
  size_t cap_ = 0, used_ = 0;
  std::map<size_t, size_t> free_;
};

enum class RState { Queued, Active, Issued, Done };

struct Request {
  uint64_t id = 0;
  uint32_t stream = 0;
  int format = 0;
  uint64_t start = 0, count = 0;
  size_t bytes = 0, arena_off = 0;
  char* out = nullptr;
  uint64_t issued = 0;
  RState st = RState::Queued;
  double t_submit = 0, t_admit = 0, t_complete = 0;
};

struct Step {
  BatchSlot slot;
  hipStream_t stream = nullptr;  // this step's launches: the next step can start during this one's tail
  std::vector<uint64_t> finishing;  // requests whose last chunk is in this step
  bool busy = false;
};

}  // namespace

struct mrg8_server {
  DeviceContext* ctx = nullptr;
  mrg8_server_config cfg{};
  LaunchConfig launch;
  uint32_t* d_states = nullptr;
  std::vector<uint32_t> h_states;
  std::vector<uint64_t> cursor;
  std::vector<uint32_t> pending_per_stream;  // requests not yet released, per stream
  char* d_arena = nullptr;
  Arena arena;
  std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
  std::deque<Request*> queued, active;
  std::vector<Request*> order;                       // scratch: service order of one step
  std::vector<std::pair<Request*, uint64_t>> taken;  // scratch: chunks of the step being built
  std::deque<Request*> done_unreported;
  std::deque<Step*> inflight;  // launch order
  std::vector<std::unique_ptr<Step>> steps;
  uint64_t next_id = 1;
  uint64_t n_completed = 0;
  uint32_t n_admitted = 0;  // active + issued (not yet complete)
  mrg8_server_counters counters{};
};

namespace {

int retire(mrg8_server* s) {
  while (!s->inflight.empty()) {
    Step* st = s->inflight.front();
    if (!st->slot.done()) break;
    const double t = wall_now();
    for (uint64_t id : st->finishing) {
      auto it = s->reqs.find(id);
      if (it == s->reqs.end()) continue;
      Request* r = it->second.get();
      r->st = RState::Done;
      r->t_complete = t;
      s->done_unreported.push_back(r);
      --s->n_admitted;
      ++s->n_completed;
    }
    st->finishing.clear();
    st->busy = false;
    s->inflight.pop_front();
  }
  return MRG8_OK;
}

void admit(mrg8_server* s) {
  while (!s->queued.empty() && s->n_admitted < s->cfg.max_inflight) {
    Request* r = s->queued.front();
    size_t off = 0;
    if (!s->arena.alloc(r->bytes, &off)) break;  // FIFO: wait for memory
    r->arena_off = off;
    r->out = s->d_arena + off;
    r->st = RState::Active;
    r->t_admit = wall_now();
    s->counters.max_queue_s = std::max(s->counters.max_queue_s, r->t_admit - r->t_submit);
    s->queued.pop_front();
    s->active.push_back(r);
    ++s->n_admitted;
  }
}

int launch_steps(mrg8_server* s) {
  while (!s->active.empty()) {
    Step* step = nullptr;
    for (auto& st : s->steps)
      if (!st->busy) {
        step = st.get();
        break;
      }
    if (!step) return MRG8_OK;
    const double t_build = wall_now();
    int rc = step->slot.wait_staging();
    if (rc) return rc;
    BatchSegment* segs = step->slot.segs();
    const uint32_t cap = step->slot.capacity();
    uint64_t budget = s->cfg.step_outputs;
    const uint64_t per_req = std::max<uint64_t>(s->launch.tile, s->cfg.step_outputs / 4);
    uint32_t n = 0;
    std::vector<Request*>& order = s->order;
    order.assign(s->active.begin(), s->active.end());
    const double now = wall_now();
    const double age_limit = 1e-6 * s->cfg.age_limit_us;
    auto mid = std::stable_partition(order.begin(), order.end(),
                                     [&](const Request* r) { return now - r->t_admit > age_limit; });
    std::stable_sort(mid, order.end(),
                     [](const Request* a, const Request* b) { return a->count - a->issued < b->count - b->issued; });
    std::vector<std::pair<Request*, uint64_t>>& taken = s->taken;
    taken.clear();
    for (auto it = order.begin(); it != order.end() && budget > 0 && n < cap; ++it) {
      Request* r = *it;
      const uint64_t left = r->count - r->issued;
      uint64_t take = std::min(left, std::min(budget, per_req));
      if (take < left && take > s->launch.tile) take = take / s->launch.tile * s->launch.tile;
      taken.emplace_back(r, take);
      BatchSegment& sg = segs[n++];
      std::memset(sg.state, 0, sizeof sg.state);
      sg.stream = r->stream;
      sg.offset = r->start + r->issued;
      sg.count = take;
      sg.out = r->out + r->issued * (uint64_t)format_size(r->format);
      sg.format = r->format;
      r->issued += take;
      budget -= take;
      if (r->issued == r->count) {
        r->st = RState::Issued;
        step->finishing.push_back(r->id);
      }
    }
    s->active.erase(std::remove_if(s->active.begin(), s->active.end(),
                                   [](const Request* r) { return r->st == RState::Issued; }),
                    s->active.end());
    if (n == 0) return MRG8_OK;
    step->slot.finalise(n, s->launch.tile);
    if ((rc = step->slot.run(s->launch, step->stream, /*graph=*/false, /*track=*/true)) != MRG8_OK) {
      // nothing of this step was launched: its chunks go back to their requests
      for (auto& [r, take] : taken) {
        if (r->st == RState::Issued) {
          r->st = RState::Active;
          s->active.push_back(r);
        }
        r->issued -= take;
      }
      step->finishing.clear();
      return rc;
    }
    step->busy = true;
    s->inflight.push_back(step);
    s->counters.steps += 1;
    s->counters.step_outputs += s->cfg.step_outputs - budget;
    s->counters.step_segments += n;
    s->counters.host_launch_s += wall_now() - t_build;
  }
  return MRG8_OK;
}

int pump(mrg8_server* s) {
  int rc = retire(s);
  if (rc) return rc;
  admit(s);
  return launch_steps(s);
}

}  // namespace

extern "C" int mrg8_server_create(int device, const mrg8_server_config* cfg, uint32_t base_seed, mrg8_server_t* out) {
  clear_error();
  if (!cfg || !out) return fail(MRG8_ERR_ARG, "null argument");
  *out = nullptr;
  if (cfg->num_streams == 0) return fail(MRG8_ERR_ARG, "num_streams must be positive");
  if (cfg->max_output_bytes < kAlign) return fail(MRG8_ERR_ARG, "max_output_bytes must be at least %zu", kAlign);
  DeviceContext* ctx = nullptr;
  int rc = device_context(device, &ctx);
  if (rc) return rc;
  auto s = std::make_unique<mrg8_server>();
  s->ctx = ctx;
  s->cfg = *cfg;
  // the arena hands out whole granules: a cap that is not a multiple would
  // accept requests whose rounded size can never fit
  s->cfg.max_output_bytes = s->cfg.max_output_bytes / kAlign * kAlign;
  if (!s->cfg.max_inflight) s->cfg.max_inflight = 4096;
  // Defaults measured on MI355X with the log-uniform request mix at 490-680
  // G samples/s offered: 4096-output tiles, one 256-thread block per CU (the
  // batched kernel's best shape), 2^26-output steps (16 tiles per wave),
  // three steps in flight, 1 ms age limit.  2^25-output steps stop keeping up
  // above ~600 G samples/s; 2^27-output steps add ~0.2-0.4 ms to p99.
  if (!s->cfg.tile_outputs) s->cfg.tile_outputs = 4096;
  if (!s->cfg.step_outputs) s->cfg.step_outputs = 1ull << 26;
  if (!s->cfg.block_threads) s->cfg.block_threads = 256;
  if (!s->cfg.blocks_per_cu) s->cfg.blocks_per_cu = 1;
  if (!s->cfg.steps_in_flight) s->cfg.steps_in_flight = 3;
  if (!s->cfg.age_limit_us) s->cfg.age_limit_us = 1000;
  if (s->cfg.steps_in_flight > kMaxSteps) return fail(MRG8_ERR_ARG, "steps_in_flight must be at most %u", kMaxSteps);
  if (s->cfg.tile_outputs % 256 || s->cfg.tile_outputs > kMaxTile || s->cfg.block_threads % 64 ||
      s->cfg.block_threads > 1024 ||
      s->cfg.blocks_per_cu < -32 ||
      (s->cfg.blocks_per_cu < 0 ? -s->cfg.blocks_per_cu : s->cfg.blocks_per_cu) * (s->cfg.block_threads / 64) > 32)
    return fail(MRG8_ERR_ARG, "invalid launch configuration");
  s->launch.block_threads = s->cfg.block_threads;
  s->launch.blocks_per_cu = s->cfg.blocks_per_cu;
  s->launch.tile = s->cfg.tile_outputs;

  DeviceGuard dg(device);
  const size_t nst = s->cfg.num_streams;
  s->h_states.resize(nst * 8);
  s->cursor.assign(nst, 0);
  s->pending_per_stream.assign(nst, 0);
  for (size_t i = 0; i < nst; ++i) seed_state(base_seed + (uint32_t)i, &s->h_states[i * 8]);
  hipError_t e = hipMalloc((void**)&s->d_states, nst * 8 * sizeof(uint32_t));
  if (e == hipSuccess)
    e = hipMemcpy(s->d_states, s->h_states.data(), nst * 8 * sizeof(uint32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc((void**)&s->d_arena, s->cfg.max_output_bytes);
  if (e != hipSuccess) {
    rc = hip_fail(e, "server allocation");
    mrg8_server_destroy(s.release());
    return rc;
  }
  s->arena.init(s->cfg.max_output_bytes);
  // Segments per step: every admitted request may contribute one chunk.
  const uint32_t seg_cap = std::min<uint32_t>(s->cfg.max_inflight, 1u << 16);
  // a chunk of c outputs takes at most ceil(c / tile) tiles
  const uint64_t tile_cap = s->cfg.step_outputs / s->cfg.tile_outputs + seg_cap + 1;
  for (uint32_t i = 0; i < s->cfg.steps_in_flight; ++i) {
    auto st = std::make_unique<Step>();
    e = hipStreamCreateWithFlags(&st->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      rc = hip_fail(e, "server stream");
      mrg8_server_destroy(s.release());
      return rc;
    }
    const bool ok = (rc = st->slot.init(ctx, seg_cap, tile_cap, s->d_states)) == MRG8_OK;
    s->steps.push_back(std::move(st));
    if (!ok) {
      mrg8_server_destroy(s.release());
      return rc;
    }
  }
  *out = s.release();
  return MRG8_OK;
}

extern "C" int mrg8_server_destroy(mrg8_server_t s) {
  clear_error();
  if (!s) return MRG8_OK;
  DeviceGuard dg(s->ctx ? s->ctx->device : 0);
  for (auto& st : s->steps)
    if (st->stream) (void)hipStreamSynchronize(st->stream);
  for (auto& st : s->steps) {
    st->slot.release();  // waits for its last run
    if (st->stream) (void)hipStreamDestroy(st->stream);
  }
  s->steps.clear();
  if (s->d_arena) (void)hipFree(s->d_arena);
  if (s->d_states) (void)hipFree(s->d_states);
  delete s;
  return MRG8_OK;
}

extern "C" int mrg8_server_set_stream(mrg8_server_t s, uint32_t stream_id, const uint32_t state[8], uint64_t position) {
  clear_error();
  if (!s || !state) return fail(MRG8_ERR_ARG, "null argument");
  if (stream_id >= s->cfg.num_streams) return fail(MRG8_ERR_NOTFOUND, "stream %u out of range", stream_id);
  if (!state_valid(state)) return fail(MRG8_ERR_ARG, "state words must be at most 2^31 - 1");
  if (s->pending_per_stream[stream_id])
    return fail(MRG8_ERR_STATE, "stream %u has unreleased requests", stream_id);
  DeviceGuard dg(s->ctx->device);
  std::memcpy(&s->h_states[(size_t)stream_id * 8], state, 8 * sizeof(uint32_t));
  s->cursor[stream_id] = position;
  // in-flight steps of other streams read the table: copy once they finish
  hipError_t e = hipSuccess;
  for (auto& st : s->steps)
    if (e == hipSuccess) e = hipStreamSynchronize(st->stream);
  if (e == hipSuccess)
    e = hipMemcpy(s->d_states + (size_t)stream_id * 8, state, 8 * sizeof(uint32_t), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "state upload");
  return MRG8_OK;
}

extern "C" int mrg8_server_set_streams(mrg8_server_t s, uint32_t first, uint32_t count, const uint32_t* states,
                                       const uint64_t* positions) {
  clear_error();
  if (!s || (count && !states)) return fail(MRG8_ERR_ARG, "null argument");
  if (first > s->cfg.num_streams || count > s->cfg.num_streams - first)
    return fail(MRG8_ERR_NOTFOUND, "streams [%u, %u) out of range", first, first + count);
  for (uint32_t i = 0; i < count; ++i) {
    if (!state_valid(states + 8ull * i)) return fail(MRG8_ERR_ARG, "stream %u: state words must be at most 2^31 - 1", first + i);
    if (s->pending_per_stream[first + i])
      return fail(MRG8_ERR_STATE, "stream %u has unreleased requests", first + i);
  }
  if (count == 0) return MRG8_OK;
  DeviceGuard dg(s->ctx->device);
  std::memcpy(&s->h_states[(size_t)first * 8], states, (size_t)count * 8 * sizeof(uint32_t));
  for (uint32_t i = 0; i < count; ++i) s->cursor[first + i] = positions ? positions[i] : 0;
  // in-flight steps of other streams read the table: copy once they finish
  hipError_t e = hipSuccess;
  for (auto& st : s->steps)
    if (e == hipSuccess) e = hipStreamSynchronize(st->stream);
  if (e == hipSuccess)
    e = hipMemcpy(s->d_states + (size_t)first * 8, states, (size_t)count * 8 * sizeof(uint32_t),
                  hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "state upload");
  return MRG8_OK;
}

extern "C" int mrg8_server_get_stream(mrg8_server_t s, uint32_t stream_id, uint32_t state[8], uint64_t* position) {
  clear_error();
  if (!s || !state || !position) return fail(MRG8_ERR_ARG, "null argument");
  if (stream_id >= s->cfg.num_streams) return fail(MRG8_ERR_NOTFOUND, "stream %u out of range", stream_id);
  std::memcpy(state, &s->h_states[(size_t)stream_id * 8], 8 * sizeof(uint32_t));
  *position = s->cursor[stream_id];
  return MRG8_OK;
}

extern "C" int mrg8_server_submit(mrg8_server_t s, uint32_t stream_id, uint64_t count, int format,
                                  uint64_t* request_id) {
  clear_error();
  if (!s || !request_id) return fail(MRG8_ERR_ARG, "null argument");
  if (stream_id >= s->cfg.num_streams) return fail(MRG8_ERR_NOTFOUND, "stream %u out of range", stream_id);
  const int esz = format_size(format);
  if (!esz) return fail(MRG8_ERR_ARG, "unknown format %d", format);
  if (count == 0) return fail(MRG8_ERR_ARG, "count must be positive");
  if (count > s->cfg.max_output_bytes / (uint64_t)esz ||
      granule_bytes(count * (uint64_t)esz) > s->cfg.max_output_bytes)
    return fail(MRG8_ERR_CAPACITY, "request of %llu outputs exceeds the output memory cap",
                (unsigned long long)count);
  if (count > UINT64_MAX - s->cursor[stream_id]) return fail(MRG8_ERR_CAPACITY, "stream position overflows");
  auto r = std::make_unique<Request>();
  r->id = s->next_id++;
  r->stream = stream_id;
  r->format = format;
  r->start = s->cursor[stream_id];
  r->count = count;
  r->bytes = count * (uint64_t)esz;
  r->t_submit = wall_now();
  s->cursor[stream_id] += count;
  ++s->pending_per_stream[stream_id];
  *request_id = r->id;
  s->queued.push_back(r.get());
  s->reqs.emplace(r->id, std::move(r));
  return MRG8_OK;
}

extern "C" int mrg8_server_progress(mrg8_server_t s, mrg8_completion* done, uint32_t max, uint32_t* ndone,
                                    int wait_ms) {
  clear_error();
  if (!s || !ndone || (max && !done)) return fail(MRG8_ERR_ARG, "null argument");
  *ndone = 0;
  const double deadline = wall_now() + (wait_ms > 0 ? wait_ms * 1e-3 : 0.0);
  for (;;) {
    int rc = pump(s);
    if (rc) return rc;
    while (*ndone < max && !s->done_unreported.empty()) {
      Request* r = s->done_unreported.front();
      s->done_unreported.pop_front();
      mrg8_completion& c = done[(*ndone)++];
      c.request_id = r->id;
      c.stream_id = r->stream;
      c.status = MRG8_OK;
      c.start = r->start;
      c.count = r->count;
      c.out = r->out;
      c.submit_time_s = r->t_submit;
      c.complete_time_s = r->t_complete;
    }
    if (*ndone > 0) return MRG8_OK;
    if (s->inflight.empty() && s->active.empty() && !s->queued.empty() && s->done_unreported.empty()) {
      // admit() left the head queued with nothing on the device: only
      // completed requests that the caller still holds can occupy the arena
      // and the slots, and no call but mrg8_server_release frees them
      const Request* h = s->queued.front();
      return fail(MRG8_ERR_BLOCKED,
                  "request %llu (%zu bytes) cannot be admitted: %zu of %llu arena bytes are held by completed "
                  "requests that have not been released",
                  (unsigned long long)h->id, granule_bytes(h->bytes), s->arena.used(),
                  (unsigned long long)s->cfg.max_output_bytes);
    }
    if (wait_ms <= 0 || wall_now() >= deadline) return MRG8_OK;
    if (s->inflight.empty() && s->active.empty() && s->queued.empty()) return MRG8_OK;
    std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
}

extern "C" int mrg8_server_release(mrg8_server_t s, uint64_t request_id) {
  clear_error();
  if (!s) return fail(MRG8_ERR_ARG, "null server");
  auto it = s->reqs.find(request_id);
  if (it == s->reqs.end()) return fail(MRG8_ERR_NOTFOUND, "unknown request %llu", (unsigned long long)request_id);
  Request* r = it->second.get();
  if (r->st == RState::Active || r->st == RState::Issued)
    return fail(MRG8_ERR_STATE, "request %llu is on the device", (unsigned long long)request_id);
  if (r->st == RState::Queued) {  // cancel before admission; its positions stay consumed
    s->queued.erase(std::find(s->queued.begin(), s->queued.end(), r));
  } else {
    auto d = std::find(s->done_unreported.begin(), s->done_unreported.end(), r);
    if (d != s->done_unreported.end()) s->done_unreported.erase(d);
    s->arena.release(r->arena_off, r->bytes);
  }
  --s->pending_per_stream[r->stream];
  s->reqs.erase(it);
  return MRG8_OK;
}

extern "C" int mrg8_server_stats(mrg8_server_t s, uint64_t* queued, uint64_t* inflight, uint64_t* completed,
                                 uint64_t* bytes_inflight) {
  clear_error();
  if (!s) return fail(MRG8_ERR_ARG, "null server");
  if (queued) *queued = s->queued.size();
  if (inflight) *inflight = s->n_admitted;
  if (completed) *completed = s->n_completed;
  if (bytes_inflight) *bytes_inflight = s->arena.used();
  return MRG8_OK;
}

extern "C" int mrg8_server_counters_get(mrg8_server_t s, mrg8_server_counters* c) {
  clear_error();
  if (!s || !c) return fail(MRG8_ERR_ARG, "null argument");
  *c = s->counters;
  return MRG8_OK;
}

// ---- trace replay (load generator for the served benchmark) ----------------------

extern "C" int mrg8_server_replay(mrg8_server_t s, const mrg8_trace_entry* trace, uint64_t n, double* latency_s,
                                  mrg8_trace_stats* stats) {
  clear_error();
  if (!s || (n && !trace)) return fail(MRG8_ERR_ARG, "null argument");
  for (uint64_t i = 1; i < n; ++i)
    if (trace[i].arrival_s < trace[i - 1].arrival_s) return fail(MRG8_ERR_ARG, "arrivals must be sorted");
  std::vector<double> lat(n, 0.0);
  std::unordered_map<uint64_t, uint64_t> index;  // request id -> trace entry
  index.reserve(2 * s->cfg.max_inflight + 16);
  std::vector<mrg8_completion> done(1024);
  uint64_t next = 0, finished = 0, outputs = 0;
  const double t0 = wall_now();
  double prev = 0.0, max_gap = 0.0, max_gap_at = 0.0;
  while (finished < n) {
    const double now = wall_now() - t0;
    if (now - prev > max_gap) {
      max_gap = now - prev;
      max_gap_at = now;
    }
    prev = now;
    while (next < n && trace[next].arrival_s <= now) {
      uint64_t id = 0;
      int rc = mrg8_server_submit(s, trace[next].stream_id, trace[next].count, trace[next].format, &id);
      if (rc) return rc;
      index[id] = next++;
    }
    uint32_t nd = 0;
    int rc = mrg8_server_progress(s, done.data(), (uint32_t)done.size(), &nd, 0);
    if (rc) return rc;
    for (uint32_t k = 0; k < nd; ++k) {
      const uint64_t i = index[done[k].request_id];
      index.erase(done[k].request_id);
      lat[i] = done[k].complete_time_s - (t0 + trace[i].arrival_s);
      outputs += done[k].count;
      ++finished;
      if ((rc = mrg8_server_release(s, done[k].request_id)) != MRG8_OK) return rc;
    }
    if (nd == 0) std::this_thread::sleep_for(std::chrono::microseconds(5));
  }
  const double wall = wall_now() - t0;
  if (latency_s) std::memcpy(latency_s, lat.data(), n * sizeof(double));
  if (stats) {
    std::vector<double> sorted = lat;
    std::sort(sorted.begin(), sorted.end());
    auto pct = [&](double q) {
      if (sorted.empty()) return 0.0;
      size_t k = (size_t)std::ceil(q * (double)sorted.size()) - 1;
      return sorted[std::min(k, sorted.size() - 1)];
    };
    double sum = 0;
    for (double v : lat) sum += v;
    stats->completed = finished;
    stats->outputs = outputs;
    stats->wall_s = wall;
    stats->p50_s = pct(0.50);
    stats->p90_s = pct(0.90);
    stats->p99_s = pct(0.99);
    stats->max_s = sorted.empty() ? 0.0 : sorted.back();
    stats->mean_s = n ? sum / (double)n : 0.0;
    stats->max_host_gap_s = max_gap;
    stats->max_host_gap_at_s = max_gap_at;
  }
  return MRG8_OK;
}
