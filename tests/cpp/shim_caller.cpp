// A C++ caller written against the reference's class surface (include/mrg8_cuda.hpp),
// with the call shape of the reference's benchmark driver
// (cuda/tests/test_mrg8_outer.cu:46-61): construct mrg8_cuda with seed 13579,
// allocate the device output, then per call seed_init + mrg8_outer_32(d_ran,
// N, BS, TNUM), then state_cpy_DtH.  Used by tests/test_gpu_shim_caller.py.
//
//   shim_caller N BS TNUM
//
// Prints, one per line: "outer <fnv1a64 of the N doubles>" for
// mrg8_outer_32(d, N, BS, TNUM), "outer2 <hash>" for the two-argument form,
// "inner <hash>" for mrg8_inner(d, N, TNUM), the print_state() text after the
// last call, "host <hash>" for mrg8::rand(double*, int) into host memory
// from the same seed, and print_matrix() of the one-step matrix A.
#include <hip/hip_runtime_api.h>

#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "mrg8_cuda.hpp"

static uint64_t fnv1a64(const void* p, size_t nbytes) {
  const unsigned char* b = static_cast<const unsigned char*>(p);
  uint64_t h = 0xcbf29ce484222325ull;
  for (size_t i = 0; i < nbytes; ++i) {
    h ^= b[i];
    h *= 0x100000001b3ull;
  }
  return h;
}

int main(int argc, char** argv) {
  const int N = argc > 1 ? std::atoi(argv[1]) : (1 << 24);
  const int BS = argc > 2 ? std::atoi(argv[2]) : 256;
  const int TNUM = argc > 3 ? std::atoi(argv[3]) : 65536;
  const uint32_t iseed = 13579;
  if (N <= 0) return 2;
  std::vector<double> ran(N);
  double* d_ran = nullptr;
  if (hipMalloc((void**)&d_ran, sizeof(double) * N) != hipSuccess) {
    std::printf("hipMalloc failed\n");
    return 3;
  }
  try {
    mrg8_cuda m(iseed);
    auto fetch = [&](const char* tag) {
      if (hipMemcpy(ran.data(), d_ran, sizeof(double) * N, hipMemcpyDeviceToHost) != hipSuccess) {
        std::printf("hipMemcpy failed\n");
        std::exit(4);
      }
      std::printf("%s %016" PRIx64 "\n", tag, fnv1a64(ran.data(), sizeof(double) * N));
    };
    m.seed_init(iseed);
    m.mrg8_outer_32(d_ran, N, BS, TNUM);
    fetch("outer");
    m.seed_init(iseed);
    m.mrg8_outer_32(d_ran, N);
    fetch("outer2");
    m.seed_init(iseed);
    m.mrg8_inner(d_ran, N, TNUM);
    fetch("inner");
    m.state_cpy_DtH();
    std::fflush(stdout);
    m.print_state();

    mrg8 h(iseed);
    h.rand(ran.data(), N);
    std::printf("host %016" PRIx64 "\n", fnv1a64(ran.data(), sizeof(double) * N));
    std::fflush(stdout);
    // jump_matrix.txt's first matrix is A (column-major); print_matrix takes [row][col]
    std::vector<uint32_t> table(247 * 64);
    if (mrg8_jump_matrix(table.data(), table.size()) != MRG8_OK) return 5;
    uint32_t a[8][8];
    for (int r = 0; r < 8; ++r)
      for (int c = 0; c < 8; ++c) a[r][c] = table[c * 8 + r];
    h.print_matrix(a);
  } catch (const mrg8_amd_error& e) {
    std::printf("error %d: %s\n", e.code(), e.what());
    hipFree(d_ran);
    return 6;
  }
  hipFree(d_ran);
  return 0;
}
