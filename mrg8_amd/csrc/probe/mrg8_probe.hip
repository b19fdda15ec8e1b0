// Measurement kernels for the benchmark (libmrg8_probe.so, include/mrg8_probe.h).
//
//  * Write probe: d_a[i] = i over n doubles, the counterpart of the
//    reference's bandwidth calibration (cuda/stream.cu:8-15,42-56).  Mode 0
//    keeps its shape (one element per thread, 512-thread blocks, plain
//    stores); mode 1 writes with the store pattern of the generator's double
//    fill and no arithmetic (each wave one run of `tile` doubles, 16 bytes
//    per lane and 1 KiB per wave store, non-temporal; 512-thread blocks, each
//    XCD taking runs of 64 consecutive blocks' output); mode 2 the same runs
//    in 1024-thread blocks in dispatch order.
//  * 64-bit multiply-add peak: every lane runs eight independent chains of
//    v_mad_i64_i32 (the generator's multiply-add), the VALU-side ceiling of
//    its integer work.
//  * Shader-clock probe: one wave counts s_memtime (shader clock) ticks
//    against the constant 100 MHz s_memrealtime over about 20 us, so a
//    caller can see the clock the power manager has set between launches.
//  * XCD probe: each block of a grid records the XCD (XCC_ID hardware
//    register) it ran on, which shows how the dispatcher deals blocks to the
//    XCDs -- the assumption behind the fills' XCD runs.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../../include/mrg8_probe.h"

namespace {

typedef double d2_t __attribute__((ext_vector_type(2)));

__global__ __launch_bounds__(512) void k_write_elem(double* __restrict__ d, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = (double)i;
}

// Block b of nb writes the output of block xcd_block(b): blocks are dealt to
// the x XCDs in turn (from wherever the turn stood), and with g > 1 the
// blocks of one XCD take runs of g consecutive blocks' output within every
// span of x g blocks (blocks past the last whole span keep their own).  The
// generator's fill maps its blocks the same way.
__device__ __forceinline__ uint64_t xcd_block(uint32_t b, uint32_t nb, uint32_t x, uint32_t g) {
  if (g <= 1 || x == 0) return b;
  const uint32_t span = x * g, whole = nb / span * span;
  if (b >= whole) return b;
  const uint32_t q = b / x;
  return (uint64_t)(q / g * span + (b - q * x) * g + q % g);
}

// Wave w writes elements [w * tile, min((w + 1) * tile, n)); tile is a
// multiple of 128 and d is 1 KiB-aligned (checked by the caller), so every
// full wave store covers 1 KiB of whole cache lines.
__global__ __launch_bounds__(1024) void k_write_tiles(double* __restrict__ d, uint64_t n, uint64_t tile, uint32_t x,
                                                      uint32_t g) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t w = xcd_block(blockIdx.x, gridDim.x, x, g) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const uint64_t first = w * tile;
  if (first >= n) return;
  const uint64_t end = first + tile < n ? first + tile : n;
  uint64_t i = first;
  for (; i + 128 <= end; i += 128) {
    const uint64_t k = i + 2 * lane;
    d2_t v = {(double)k, (double)(k + 1)};
    __builtin_nontemporal_store(v, reinterpret_cast<d2_t*>(d + k));
  }
  for (uint64_t k = i + lane; k < end; k += 64) __builtin_nontemporal_store((double)k, d + k);
}

__global__ __launch_bounds__(256) void k_mad64(uint64_t* __restrict__ out, int iters) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  uint64_t acc[8];
  int32_t u[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    acc[j] = t + j;
    u[j] = (int32_t)(t ^ (j * 0x9e3779b9u));
  }
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int j = 0; j < 8; ++j)  // acc = lo32(acc) * u + acc: each chain depends on itself only
      acc[j] = (uint64_t)((int64_t)(int32_t)(uint32_t)acc[j] * (int64_t)u[j]) + acc[j];
  }
  uint64_t sum = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) sum += acc[j];
  out[t] = sum;  // keeps the chains live
}

__global__ void k_sclk(uint64_t* __restrict__ out) {
  const uint64_t r0 = __builtin_amdgcn_s_memrealtime();
  const uint64_t t0 = __builtin_amdgcn_s_memtime();
  uint64_t r1 = r0;
  // about 20 us of the 100 MHz counter; the iteration bound ends the loop
  // even if that counter did not advance
  for (int i = 0; i < (1 << 22) && r1 - r0 < 2000; ++i) r1 = __builtin_amdgcn_s_memrealtime();
  const uint64_t t1 = __builtin_amdgcn_s_memtime();
  if (threadIdx.x == 0) {
    out[0] = t1 - t0;
    out[1] = r1 - r0;
  }
}

// XCC_ID (hardware register 20 on gfx9.4/9.5): bits [3:0] are the XCD.
__global__ void k_xcc(uint32_t* __restrict__ out) {
  const uint32_t id = __builtin_amdgcn_s_getreg(20 | (0 << 6) | (3 << 11));
  if (threadIdx.x == 0) out[blockIdx.x] = id;
}

}  // namespace

extern "C" int mrg8_probe_write(void* d, uint64_t n, int mode, uint64_t tile, void* stream) {
  if (!d && n) return -1;
  if (n == 0) return 0;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (mode == 0) {
    const uint64_t blocks = (n + 511) / 512;
    if (blocks > 0xffffffffull) return -1;
    hipLaunchKernelGGL(k_write_elem, dim3((uint32_t)blocks), dim3(512), 0, s, static_cast<double*>(d), n);
  } else if (mode == 1 || mode == 2) {
    if (tile == 0) tile = 2048;
    if (tile % 128 || (reinterpret_cast<uintptr_t>(d) & 1023)) return -1;
    const uint32_t threads = mode == 1 ? 512 : 1024;
    uint32_t xcds = 1, group = 1;
    if (mode == 1) {
      int dev = 0, x = 1;
      if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&x, hipDeviceAttributeNumberOfXccs, dev) != hipSuccess) {
        (void)hipGetLastError();
        x = 1;
      }
      xcds = x > 1 ? (uint32_t)x : 1;
      group = xcds > 1 ? 32768 / threads : 1;
    }
    const uint64_t waves = (n + tile - 1) / tile;
    const uint64_t blocks = (waves + threads / 64 - 1) / (threads / 64);
    if (blocks > 0xffffffffull) return -1;
    hipLaunchKernelGGL(k_write_tiles, dim3((uint32_t)blocks), dim3(threads), 0, s, static_cast<double*>(d), n, tile,
                       xcds, group);
  } else {
    return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

extern "C" int mrg8_probe_mad64(uint64_t* d_out, uint32_t blocks, int iters, void* stream) {
  if (!d_out || blocks == 0 || iters <= 0) return -1;
  hipLaunchKernelGGL(k_mad64, dim3(blocks), dim3(256), 0, static_cast<hipStream_t>(stream), d_out, iters);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

extern "C" int mrg8_probe_xcc(uint32_t* d_out, uint32_t blocks, void* stream) {
  if (!d_out || blocks == 0) return -1;
  hipLaunchKernelGGL(k_xcc, dim3(blocks), dim3(64), 0, static_cast<hipStream_t>(stream), d_out);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

extern "C" int mrg8_probe_sclk(uint64_t* d_out, void* stream) {
  if (!d_out) return -1;
  hipLaunchKernelGGL(k_sclk, dim3(1), dim3(64), 0, static_cast<hipStream_t>(stream), d_out);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
