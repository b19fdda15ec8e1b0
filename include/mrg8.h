/*
 * mrg8.h -- C ABI of the MI355X-native MRG8 generator (libmrg8_amd.so).
 *
 * MRG8 is the 8th-order multiple recursive generator modulo p = 2^31 - 1,
 *   S[n] = (a0 S[n-1] + a1 S[n-2] + ... + a7 S[n-8]) mod p,
 * whose 8-word state is kept newest word first (state[0] = S[n-1]).  Every
 * output of this library is bit-identical to the reference's serial
 * generator (original/mrg8.cpp) for the same state and position: words are in
 * [1, p] (a residue of 0 is emitted as p), and doubles are formed by
 * multiplying with the rounded reciprocal 1.0/2147483647.0.
 *
 * The entry points replace the reference's C++ classes `mrg8` (host) and
 * `mrg8_cuda` (device generation path); each cites the method it stands in
 * for, as file:line in the reference tree.  The header-only C++ shim
 * include/mrg8_cuda.hpp rebuilds those class names on top of this ABI.
 *
 * Conventions:
 *   - Every function returns an int status: MRG8_OK (0) or a negative
 *     MRG8_ERR_* code; mrg8_last_error() gives the calling thread's message.
 *     Nothing calls exit() (the reference exits on a missing table file).
 *   - Device pointers are plain addresses on the generator's device; `stream`
 *     arguments are hipStream_t values passed as void* (NULL = null stream).
 *   - Generation calls are stream-ordered and do not allocate: the host-side
 *     state advances immediately by n (as if the call had run), so successive
 *     calls continue the sequence without a device round trip.
 *   - Distinct handles may be used from different threads; one handle needs
 *     external synchronisation.
 */
#ifndef MRG8_AMD_H
#define MRG8_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MRG8_ABI_VERSION 1

/* Status codes. */
#define MRG8_OK 0
#define MRG8_ERR_ARG (-1)       /* invalid argument (null pointer, bad format, n out of range) */
#define MRG8_ERR_HIP (-2)       /* a HIP runtime call failed */
#define MRG8_ERR_NOMEM (-3)     /* device or host allocation failed */
#define MRG8_ERR_NODEVICE (-4)  /* no usable gfx950 device / kernels not loadable */
#define MRG8_ERR_CAPACITY (-5)  /* a fixed capacity (streams, tiles, bytes) would be exceeded */
#define MRG8_ERR_STATE (-6)     /* call not valid in the object's current state */
#define MRG8_ERR_NOTFOUND (-7)  /* unknown stream or request id */
#define MRG8_ERR_IO (-8)        /* file could not be read or written */
#define MRG8_ERR_BLOCKED (-9)   /* no progress possible until the caller releases completed requests */

/* Output formats. */
#define MRG8_F64 0     /* double(x) * (1/p), (0, 1]: original/mrg8.cpp:231-239 (mrg8::rand) */
#define MRG8_F64_01 1  /* double(x - 1) * (1/p), [0, 1): cuda/mrg8_cuda.cu:363 (mrg8_outer_32) */
#define MRG8_U32 2     /* raw word in [1, p]: original/mrg8.cpp:197-222 (mrg8::randint) */

typedef struct mrg8_gen* mrg8_gen_t;
typedef struct mrg8_batch* mrg8_batch_t;
typedef struct mrg8_server* mrg8_server_t;

/* ---- library ----------------------------------------------------------- */

int mrg8_abi_version(void);
const char* mrg8_strerror(int status);
/* Message for the last failing call on this thread ("" if none). */
const char* mrg8_last_error(void);
/* Loads the jump tables and kernels on `device` (called implicitly by
 * mrg8_create on the current device). */
int mrg8_init_device(int device);

/* ---- generator: construction and state (classes mrg8 / mrg8_cuda) ------- */

/* mrg8_cuda::mrg8_cuda(seed) and mrg8_cuda::mrg8_cuda() (cuda/mrg8_cuda.cu:22-63,
 * seeding by mcg64ni, original/mrg8.cpp:180-194; seed 0 means 97531).  The
 * generator is bound to the current HIP device.  The jump table is built in
 * memory from its closed form; no jump_matrix.txt is read. */
int mrg8_create(uint32_t seed, mrg8_gen_t* out);
/* Frees the handle.  Generation already issued completes on its streams:
 * its kernels read only their arguments and tables that live until the
 * process ends, so destroy does not wait for them. */
int mrg8_destroy(mrg8_gen_t g);
/* mrg8_cuda::seed_init (cuda/mrg8_cuda.cu:66-70). */
int mrg8_seed_init(mrg8_gen_t g, uint32_t seed);
/* mrg8::get_state / set_state (original/mrg8.cpp:36-39,56-59).  Unlike
 * mrg8_cuda (whose set_state leaves the device copy untouched), the state
 * set here is the one the next device generation starts from. */
int mrg8_get_state(mrg8_gen_t g, uint32_t state[8]);
int mrg8_set_state(mrg8_gen_t g, const uint32_t state[8]);
/* mrg8::jump_ahead(uint64_t) -- mutating form (original/mrg8.cpp:125-129). */
int mrg8_jump_ahead(mrg8_gen_t g, uint64_t step);
/* mrg8::jump_ahead(const short[200]) -- 200-bit jump, bits[0] the least
 * significant bit (original/mrg8.cpp:105-122). */
int mrg8_jump_ahead_bits(mrg8_gen_t g, const short bits[200]);
/* mrg8::jump_ahead(uint64_t, uint32_t* new_state) -- non-mutating form of
 * the CUDA/knl classes (cuda/mrg8.cpp:145-150). */
int mrg8_jump_state(mrg8_gen_t g, uint64_t step, uint32_t new_state[8]);
/* mrg8::jump_ahead(const short[200], uint32_t* new_state) (cuda/mrg8.cpp:127-141). */
int mrg8_jump_state_bits(mrg8_gen_t g, const short bits[200], uint32_t new_state[8]);

/* Pure state functions (no handle, no device): the state after seeding
 * (original/mrg8.cpp:180-194) and a state advanced by `step` positions
 * (the new_state form of jump_ahead, cuda/mrg8.cpp:127-150; 200-bit form
 * with bits[0] least significant). */
int mrg8_seed_state(uint32_t seed, uint32_t state[8]);
int mrg8_advance_state(const uint32_t state[8], uint64_t step, uint32_t new_state[8]);
int mrg8_advance_state_bits(const uint32_t state[8], const short bits[200], uint32_t new_state[8]);

/* ---- generator: device generation (the hot path) ------------------------ */

/* Fill `dptr` (device memory, 8-byte aligned for doubles, 4 for words) with
 * the next n outputs in `format`, stream-ordered on `stream`; the state
 * advances by n.  Replaces the rng_outer_32 / rng_inner launches
 * (cuda/mrg8_cuda.cu:282-399,201-280 via :402-436) with no size limit
 * (n is 64-bit; the reference takes int and jumps only 32 bits on device). */
int mrg8_generate(mrg8_gen_t g, void* dptr, uint64_t n, int format, void* stream);
/* Same, with an explicit launch shape; 0 means automatic for each, and the
 * output is identical for every shape.  outputs_per_lane 1, 2 or 4 selects the
 * vector kernel (threads per block 64..1024, a multiple of 64; outputs per
 * wave at most n, rounded up to 64 * outputs per lane); 16 or 32 selects the
 * matrix-core fill (int8 MFMAs; tiles of 16 columns of 4 * outputs_per_lane
 * outputs, threads per block 256, 512 or 1024, outputs per wave a multiple of
 * the tile), which falls back to the vector kernel when the fill holds no
 * whole tile or the state is all zero.  Automatic: the matrix-core fill for
 * word fills of 2^22 outputs or more, the vector kernel otherwise, unless
 * mrg8_tune_fill chose a shape for that size class. */
int mrg8_generate_shape(mrg8_gen_t g, void* dptr, uint64_t n, int format, void* stream,
                        int block_threads, uint64_t outputs_per_wave, int outputs_per_lane);
/* Launch shape of a single-stream fill. */
typedef struct mrg8_fill_shape {
  int32_t block_threads;      /* threads per block */
  int32_t outputs_per_lane;   /* 1, 2, 4: vector kernel; 16, 32: matrix-core fill */
  uint64_t outputs_per_wave;  /* consecutive outputs per wave (the jump stride between waves) */
} mrg8_fill_shape;

/* Autotuner for single-stream fills: times threads-per-block x outputs-per-
 * lane (vector kernel and matrix-core fill) x outputs-per-wave on `device` for
 * fills of about n outputs in
 * `format` (a scratch buffer of n outputs is allocated for the trials) and
 * makes the fastest the default shape of mrg8_generate for that format and
 * size class (floor(log2 n)) on that device.  `cache_path` (may be NULL) is a
 * text file of results keyed by device, format and size class; a cached
 * entry is used without timing.  The counterpart of the reference's manual
 * TNUM x BS sweep (cuda/tests/test_mrg8_outer.cu:50-53). */
int mrg8_tune_fill(int device, int format, uint64_t n, const char* cache_path, mrg8_fill_shape* best);

/* mrg8_cuda::mrg8_outer_32(d_ran, n[, TNUM[, BS]]) (cuda/mrg8_cuda.cu:414-436):
 * blocking, [0, 1) doubles (MRG8_F64_01).  BS/TNUM keep the reference's
 * meaning (threads per block / total threads) and pick the launch shape;
 * TNUM = 0 takes the library's own shape (what the C++ shim's two-argument
 * form passes: the reference's default of 64 * 1024 threads gives every wave
 * n / 1024 outputs), BS = 0 the reference's 64.  The outputs are the same
 * for every shape. */
int mrg8_outer_32(mrg8_gen_t g, double* d_ran, int n, int BS, int TNUM);
/* mrg8_cuda::mrg8_inner(d_ran, n, TNUM) (cuda/mrg8_cuda.cu:402-412): blocking,
 * [0, 1) doubles; the same sequence as mrg8_outer_32 (TNUM = 0: the
 * library's own shape). */
int mrg8_inner(mrg8_gen_t g, double* d_ran, int n, int TNUM);
/* mrg8_cuda::state_cpy_DtH (cuda/mrg8_cuda.cu:72-79).  The host copy of the
 * state is always current here; this waits for the generator's outstanding
 * work and returns. */
int mrg8_state_cpy_DtH(mrg8_gen_t g);
/* Waits until all generation issued through `g` has finished: synchronises
 * every stream `g` generated on since the last call (those streams must
 * still exist). */
int mrg8_synchronize(mrg8_gen_t g);

/* ---- generator: host-memory forms (class mrg8) --------------------------- */

/* mrg8::randint(uint32_t*, int) / mrg8::rand(double*, int)
 * (original/mrg8.cpp:197-239): generated on the device, copied to the host
 * buffer; blocking. */
int mrg8_randint(mrg8_gen_t g, uint32_t* host, uint64_t n);
int mrg8_rand(mrg8_gen_t g, double* host, uint64_t n);
/* mrg8::mrg8dnz2(double*, int) (cuda/mrg8.cpp:290-344): the same sequence and
 * conversion as mrg8_rand. */
int mrg8_mrg8dnz2(mrg8_gen_t g, double* host, uint64_t n);
/* mrg8::randint() / mrg8::rand() / operator() -- one value, computed on the
 * host (original/mrg8.cpp:224-228,241-245). */
int mrg8_randint1(mrg8_gen_t g, uint32_t* value);
int mrg8_rand1(mrg8_gen_t g, double* value);

/* ---- jump-table format (jump_matrix.txt) -------------------------------- */

/* Writes A^(2^j) mod p, j = 0..246, in the reference's text layout
 * (one matrix per line, column-major, original/mrg8.cpp:85-102). */
int mrg8_write_jump_matrix(const char* path);
/* Copies the 247*64 table words (column-major per matrix) to `table`. */
int mrg8_jump_matrix(uint32_t* table, size_t words);
/* read_jump_matrix (original/mrg8.cpp:88-102): parses a jump_matrix.txt into
 * the 247*64 words of `table` (layout of mrg8_jump_matrix).  A missing,
 * short or malformed file is MRG8_ERR_IO (the reference exits).  The library
 * never needs the file: its table is computed. */
int mrg8_read_jump_matrix(const char* path, uint32_t* table, size_t words);
/* Reads `path` and checks it word for word against the computed table:
 * MRG8_OK, MRG8_ERR_IO (unreadable) or MRG8_ERR_STATE (a word differs). */
int mrg8_check_jump_matrix(const char* path);

/* ---- bulk jumps of device-resident states -------------------------------- */

/* Advances n states on `device`: the 8 words at d_out + 8 i become those at
 * d_in + 8 i jumped ahead by offset i, the little-endian integer of `words`
 * 64-bit words at d_offsets + words i (words 1..4: offsets below 2^64 ..
 * 2^256; the reference's widest jump, jump_ahead(short[200]) at
 * original/mrg8.cpp:105-122 and cuda/mrg8.cpp:127-141, takes 200 bits).  The
 * device-side counterpart of calling mrg8_advance_state_bits n times; the
 * reference has no device jump past 32 bits (cuda/mrg8_cuda.cu:219,314).
 * All pointers are device memory; d_in and d_out may be the same buffer
 * (other overlaps are not allowed); state words must be at most 2^31 - 1.
 * Stream-ordered on `stream`.  The first call with words > 1 builds a
 * 2.3 MiB table for byte positions 8..31 on the device (kept for the
 * process). */
int mrg8_advance_states(int device, const uint32_t* d_in, const uint64_t* d_offsets, uint32_t words,
                        uint32_t* d_out, uint64_t n, void* stream);

/* ---- batched generation: many streams, one persistent launch ------------ */

/* One stream segment: `count` outputs starting `offset` outputs after the
 * position whose state is `state`.  `out` is device memory. */
typedef struct mrg8_stream_desc {
  uint32_t state[8]; /* newest word first, as mrg8_get_state returns */
  uint64_t offset;   /* first output index relative to `state` */
  uint64_t count;    /* outputs to write */
  void* out;         /* device pointer */
  int32_t format;    /* MRG8_F64, MRG8_F64_01 or MRG8_U32 */
  uint32_t reserved; /* must be 0 */
} mrg8_stream_desc;

/* Launch configuration (picked by mrg8_batch_autotune or set by hand; a new
 * batch starts with 1024 threads, one tile per wave, 2048-output tiles). */
typedef struct mrg8_batch_config {
  int32_t block_threads;  /* 64..1024, multiple of 64 */
  int32_t blocks_per_cu;  /* 0: one tile per wave (grid from the tile count; best for long
                             segments); < 0: persistent grid of CUs * -blocks_per_cu blocks
                             that take groups of consecutive tiles in address order from a
                             counter in the batch's device block (as fast or faster for long
                             segments; runs of one batch are serialised, also across
                             streams; the counter is 32 bits: a run with more groups of
                             tiles than that is MRG8_ERR_CAPACITY); > 0: persistent grid
                             of CUs * blocks_per_cu blocks
                             (best for many short segments with long jumps) */
  uint64_t tile_outputs;  /* outputs per work item, multiple of 256, at most 2^40 */
} mrg8_batch_config;

/* A batch object owns device descriptor and tile tables sized for
 * `max_streams` segments and `max_tiles` work items (no reference
 * counterpart: the reference generates one stream per blocking call).  A
 * segment of c outputs takes at most ceil(c / tile_outputs) tiles: its first
 * tile also takes the outputs before its first tile-aligned address (a power
 * of two dividing tile_outputs times the element size). */
int mrg8_batch_create(int device, uint32_t max_streams, uint64_t max_tiles, mrg8_batch_t* out);
int mrg8_batch_destroy(mrg8_batch_t b);
int mrg8_batch_get_config(mrg8_batch_t b, mrg8_batch_config* cfg);
int mrg8_batch_set_config(mrg8_batch_t b, const mrg8_batch_config* cfg);
/* Stages `count` descriptors (host memory) for the next run. */
int mrg8_batch_set(mrg8_batch_t b, const mrg8_stream_desc* descs, uint32_t count);
/* Launches the staged descriptors on `stream` (stream-ordered; the staged
 * block is copied to the device only when it changed).  The batch remembers
 * the stream of its last run: destroy the batch, or run it on another
 * stream, before destroying that stream. */
int mrg8_batch_run(mrg8_batch_t b, void* stream);
/* Runs replay a captured HIP graph holding the launch (enable != 0) instead
 * of launching the kernel directly (the default: on MI355X a one-kernel
 * graph replay costs ~10 us more per launch). */
int mrg8_batch_use_graph(mrg8_batch_t b, int enable);
/* Times candidate configurations on the staged descriptors (the outputs are
 * rewritten each trial) and keeps the fastest.  `cache_path` (may be NULL)
 * is a text file of results keyed by device and workload class. */
int mrg8_batch_autotune(mrg8_batch_t b, void* stream, const char* cache_path, mrg8_batch_config* best);

/* Batched generation that allocates nothing.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * same generation as mrg8_batch_set + mrg8_batch_run with configuration
 * `cfg` (NULL: the batch default, 1024 threads and one 2048-output tile per
 * wave): identical outputs, stream-ordered on `stream`.  The batch block
 * (segments, tile prefix sums, tile map) goes to a fixed device scratch area
 * of MRG8_NOALLOC_SCRATCH_BYTES that the library holds on every device it
 * runs on until the process ends; the call makes no hipMalloc/hipFree and no
 * heap allocation.  mrg8_init_device(device) must have been called
 * (MRG8_ERR_STATE otherwise: creating the device's tables allocates).  Work
 * whose block exceeds the area (segments x 64 B + tiles x 4 B, about) is
 * MRG8_ERR_CAPACITY.  All calls on a device share its scratch area: calls
 * that overlap on one device race, so issue them on one stream or wait in
 * between.  `descs` may be reused as soon as the call returns. */
#define MRG8_NOALLOC_SCRATCH_BYTES (4u << 20)
int mrg8_batch_generate_noalloc(int device, const mrg8_stream_desc* descs, uint32_t count,
                                const mrg8_batch_config* cfg, void* stream);

/* ---- serving: resident streams, continuous batching -------------------- */

typedef struct mrg8_server_config {
  uint32_t num_streams;        /* resident per-stream states in device memory */
  uint32_t max_inflight;       /* requests admitted at once */
  uint64_t max_output_bytes;   /* admission cap on output-buffer bytes in flight (rounded down to a
                                  multiple of 256, the arena's allocation granule; at least 256) */
  uint64_t step_outputs;       /* target outputs per launch (0 = automatic) */
  uint64_t tile_outputs;       /* outputs per work item, multiple of 256, at most 2^40
                                  (0 = automatic) */
  int32_t block_threads;       /* 0 = automatic */
  int32_t blocks_per_cu;       /* 0 = automatic (a persistent grid striding over the tiles, 1 block
                                  per CU); as in mrg8_batch_config otherwise: < 0 takes the tiles in
                                  address order (measured slower for the serving loop's short
                                  request chunks) */
  uint32_t steps_in_flight;    /* launches in flight at once, 1..8 (0 = 3) */
  uint32_t age_limit_us;       /* requests admitted longer ago than this are served first,
                                  in admission order; the rest shortest-remaining first (0 = 1000) */
} mrg8_server_config;

typedef struct mrg8_completion {
  uint64_t request_id;
  uint32_t stream_id;
  int32_t status;         /* MRG8_OK or an error */
  uint64_t start;         /* stream position of the first output */
  uint64_t count;
  void* out;              /* device buffer holding the outputs (owned by the server) */
  double submit_time_s;   /* host clock (seconds) at submit */
  double complete_time_s; /* host clock at which completion was observed */
} mrg8_completion;

/* Creates a server on `device` with `num_streams` resident states; stream i
 * starts from the state of seed (base_seed + i) unless set otherwise. */
int mrg8_server_create(int device, const mrg8_server_config* cfg, uint32_t base_seed, mrg8_server_t* out);
int mrg8_server_destroy(mrg8_server_t s);
int mrg8_server_set_stream(mrg8_server_t s, uint32_t stream_id, const uint32_t state[8], uint64_t position);
/* Streams first .. first+count-1 at once: states[8 i .. 8 i + 7] and
 * positions[i] (NULL: position 0); one upload. */
int mrg8_server_set_streams(mrg8_server_t s, uint32_t first, uint32_t count, const uint32_t* states,
                            const uint64_t* positions);
int mrg8_server_get_stream(mrg8_server_t s, uint32_t stream_id, uint32_t state[8], uint64_t* position);
/* Queues a request for the next `count` outputs of stream `stream_id`; the
 * request is given the stream's next `count` positions at submit time, so
 * its outputs do not depend on what else is in flight.  Admission to the GPU
 * happens inside mrg8_server_progress as output memory and slots free up.
 * A request whose output bytes, rounded up to the 256-byte granule, exceed
 * the cap is MRG8_ERR_CAPACITY (it could never be admitted). */
int mrg8_server_submit(mrg8_server_t s, uint32_t stream_id, uint64_t count, int format, uint64_t* request_id);
/* Admits queued requests, launches work, and collects finished requests.
 * Up to `max` completions are written to `done`; *ndone gets the number.
 * `wait_ms` > 0 blocks up to that long for at least one completion.
 * Admission is first come, first served: when nothing is on the device,
 * nothing completed is left to report and the oldest queued request does not
 * fit because completed requests the caller has not released hold the
 * arena (or its slots), the call returns MRG8_ERR_BLOCKED at once instead of
 * waiting: only mrg8_server_release can make progress possible. */
int mrg8_server_progress(mrg8_server_t s, mrg8_completion* done, uint32_t max, uint32_t* ndone, int wait_ms);
/* Returns a completed request's output buffer to the admission pool. */
int mrg8_server_release(mrg8_server_t s, uint64_t request_id);
/* Counters: queued, admitted (in flight), completed, output bytes in flight. */
int mrg8_server_stats(mrg8_server_t s, uint64_t* queued, uint64_t* inflight, uint64_t* completed,
                      uint64_t* bytes_inflight);

/* Scheduler counters since creation: launched steps, outputs and segments
 * they carried, and host seconds spent building and launching them. */
typedef struct mrg8_server_counters {
  uint64_t steps;
  uint64_t step_outputs;
  uint64_t step_segments;
  double host_launch_s;
  double max_queue_s;  /* longest time a request waited for admission */
} mrg8_server_counters;
int mrg8_server_counters_get(mrg8_server_t s, mrg8_server_counters* c);

/* Trace replay: a load generator for the served benchmark.  Requests are
 * submitted when the wall clock passes their arrival time (seconds from the
 * start of the replay, non-decreasing), each completion is released at once,
 * and latency is measured from arrival to observed completion, so it
 * includes queueing.  `latency_s` (n entries, may be NULL) receives the
 * per-request latencies in trace order. */
typedef struct mrg8_trace_entry {
  double arrival_s;
  uint32_t stream_id;
  int32_t format;
  uint64_t count;
} mrg8_trace_entry;

typedef struct mrg8_trace_stats {
  uint64_t completed;
  uint64_t outputs;
  double wall_s;
  double p50_s, p90_s, p99_s, max_s, mean_s;
  double max_host_gap_s;   /* longest time between two passes of the replay loop */
  double max_host_gap_at_s; /* when that gap ended, seconds from the start of the replay */
} mrg8_trace_stats;

int mrg8_server_replay(mrg8_server_t s, const mrg8_trace_entry* trace, uint64_t n, double* latency_s,
                       mrg8_trace_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* MRG8_AMD_H */
