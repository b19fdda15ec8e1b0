/*
 * mrg8_probe.h -- measurement kernels used by the benchmark (libmrg8_probe.so).
 * Not part of the generator's interface.  Status: 0 ok, -1 invalid argument,
 * -2 launch failed.  Pointers are device memory; `stream` is a hipStream_t.
 */
#ifndef MRG8_PROBE_H
#define MRG8_PROBE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* d[i] = (double)i for i < n: the write-bandwidth calibration of the
 * reference (cuda/stream.cu:8-15, one element per thread in 512-thread
 * blocks) with mode 0; with mode 1 the store pattern of the generator's
 * double fill (runs of `tile` doubles per wave, 0 = 2048, a multiple of 128;
 * 1 KiB non-temporal wave stores; d 1 KiB-aligned; 512-thread blocks, each
 * XCD taking runs of 64 consecutive blocks' output) and no arithmetic; mode 2
 * the same runs in 1024-thread blocks in dispatch order. */
int mrg8_probe_write(void* d, uint64_t n, int mode, uint64_t tile, void* stream);

/* `blocks` blocks of 256 threads, each thread running eight independent
 * chains of `iters` 32x32->64 signed multiply-adds (v_mad_i64_i32) and
 * writing one word to d_out[thread] (d_out: blocks * 256 words): the
 * lane-multiply-adds per second the VALU sustains, blocks * 256 * 8 * iters
 * divided by the kernel time. */
int mrg8_probe_mad64(uint64_t* d_out, uint32_t blocks, int iters, void* stream);

/* One wave spins ~20 us and writes d_out[0] = shader-clock ticks
 * (s_memtime) and d_out[1] = 100 MHz reference ticks (s_memrealtime) over
 * the same interval: clock in MHz = 100 * d_out[0] / d_out[1]. */
int mrg8_probe_sclk(uint64_t* d_out, void* stream);

/* `blocks` blocks of one wave each; block b writes d_out[b] = the XCD it ran
 * on (bits [3:0] of the XCC_ID hardware register).  The fills and the
 * batched kernels map their blocks to the output in XCD runs on the
 * assumption that blocks are dealt to the XCDs in turn: block b on XCD
 * (b + r) mod (number of XCDs), r depending on earlier launches. */
int mrg8_probe_xcc(uint32_t* d_out, uint32_t blocks, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MRG8_PROBE_H */
