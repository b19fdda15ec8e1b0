// mrg8_cuda.hpp -- header-only C++ classes with the reference's names and
// method signatures (cuda/mrg8.h:23-79, cuda/mrg8_cuda.h:20-45), implemented
// on the C ABI of libmrg8_amd.so (include/mrg8.h).  Code written against the
// reference's `mrg8` / `mrg8_cuda` classes compiles against this header
// unchanged; link with -lmrg8_amd.
//
// Differences from the reference, all deliberate:
//   * errors throw mrg8_amd_error instead of exit() / being ignored;
//   * set_state() is the state the next device generation starts from
//     (the reference's mrg8_cuda only updates its host copy);
//   * generation sizes may exceed int through the uint64_t overloads, which
//     take part only for arguments of exactly uint64_t, so a call that the
//     reference's single int signature accepts (unsigned, long, size_t
//     arguments) resolves to the int overload here too.
#ifndef MRG8_AMD_CUDA_HPP
#define MRG8_AMD_CUDA_HPP

#include <cstdint>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "mrg8.h"

class mrg8_amd_error : public std::runtime_error {
 public:
  mrg8_amd_error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

namespace mrg8_amd_detail {
inline void check(int rc, const char* fn) {
  if (rc != MRG8_OK) throw mrg8_amd_error(rc, std::string(fn) + ": " + mrg8_last_error());
}
}  // namespace mrg8_amd_detail

// Host-class surface (original/mrg8.h:21-56, cuda/mrg8.h:23-79).
class mrg8 {
 public:
  const uint32_t MAX_RND = 2147483646u;

  mrg8() : mrg8(0u) {}
  explicit mrg8(const uint32_t seed_val) { mrg8_amd_detail::check(mrg8_create(seed_val, &g_), "mrg8_create"); }
  virtual ~mrg8() { mrg8_destroy(g_); }
  mrg8(const mrg8&) = delete;
  mrg8& operator=(const mrg8&) = delete;

  void seed_init(const uint32_t seed_val) { mrg8_amd_detail::check(mrg8_seed_init(g_, seed_val), "seed_init"); }
  void get_state(uint32_t st[8]) const { mrg8_amd_detail::check(mrg8_get_state(g_, st), "get_state"); }
  void set_state(const uint32_t st[8]) { mrg8_amd_detail::check(mrg8_set_state(g_, st), "set_state"); }
  // original/mrg8.cpp:41-54 (cuda/mrg8.cpp:50-68): the same text on std::cout
  void print_state() const {
    uint32_t st[8];
    get_state(st);
    std::cout << "(S[n-1], S[n-2], ..., S[n-8])= (";
    for (int i = 0; i < 7; ++i) std::cout << std::setw(10) << st[i] << ", ";
    std::cout << st[7] << ")" << std::endl;
  }
  void print_matrix(const uint32_t jm[8][8]) const {
    for (int i = 0; i < 8; ++i) {
      for (int j = 0; j < 8; ++j) std::cout << std::setw(10) << jm[i][j] << " ";
      std::cout << std::endl;
    }
  }

  void randint(uint32_t* iran, int n) { randint64(iran, (uint64_t)(n < 0 ? 0 : n)); }
  template <typename N, typename std::enable_if<std::is_same<N, uint64_t>::value, int>::type = 0>
  void randint(uint32_t* iran, N n) { randint64(iran, n); }
  uint32_t randint() {
    uint32_t v;
    mrg8_amd_detail::check(mrg8_randint1(g_, &v), "randint");
    return v;
  }
  void rand(double* fran, int n) { rand64(fran, (uint64_t)(n < 0 ? 0 : n)); }
  template <typename N, typename std::enable_if<std::is_same<N, uint64_t>::value, int>::type = 0>
  void rand(double* fran, N n) { rand64(fran, n); }
  double rand() {
    double v;
    mrg8_amd_detail::check(mrg8_rand1(g_, &v), "rand");
    return v;
  }
  double operator()() { return rand(); }
  void mrg8dnz2(double* ran, int n) {
    mrg8_amd_detail::check(mrg8_mrg8dnz2(g_, ran, (uint64_t)(n < 0 ? 0 : n)), "mrg8dnz2");
  }

  // original/mrg8.h:36-37: mutating jumps
  void jump_ahead(const short jump_val_bin[200]) {
    mrg8_amd_detail::check(mrg8_jump_ahead_bits(g_, jump_val_bin), "jump_ahead");
  }
  void jump_ahead(const uint64_t jump_val) { mrg8_amd_detail::check(mrg8_jump_ahead(g_, jump_val), "jump_ahead"); }
  // cuda/mrg8.h:52-53: jumps into new_state, leaving the generator as it is
  void jump_ahead(const short jump_val_bin[200], uint32_t* new_state) const {
    mrg8_amd_detail::check(mrg8_jump_state_bits(g_, jump_val_bin, new_state), "jump_ahead");
  }
  void jump_ahead(const uint64_t jump_val, uint32_t* new_state) const {
    mrg8_amd_detail::check(mrg8_jump_state(g_, jump_val, new_state), "jump_ahead");
  }

  mrg8_gen_t handle() const { return g_; }

 protected:
  mrg8_gen_t g_ = nullptr;

 private:
  void randint64(uint32_t* iran, uint64_t n) { mrg8_amd_detail::check(mrg8_randint(g_, iran, n), "randint"); }
  void rand64(double* fran, uint64_t n) { mrg8_amd_detail::check(mrg8_rand(g_, fran, n), "rand"); }
};

// Device generation surface (cuda/mrg8_cuda.h:20-45).  Output is the [0, 1)
// conversion double(x - 1) / (2^31 - 1) of the reference's kernels.
class mrg8_cuda : public mrg8 {
 public:
  mrg8_cuda() : mrg8() {}
  explicit mrg8_cuda(const uint32_t seed_val) : mrg8(seed_val) {}

  void mrg8_inner(double* d_ran, const int n, const int TNUM) {
    mrg8_amd_detail::check(::mrg8_inner(g_, d_ran, n, TNUM), "mrg8_inner");
  }
  // the reference's default of 64 * 1024 threads (cuda/mrg8_cuda.cu:433-436)
  // was a CUDA launch choice; here the library picks the shape (same outputs)
  void mrg8_outer_32(double* d_ran, const int n) {
    mrg8_amd_detail::check(::mrg8_outer_32(g_, d_ran, n, 0, 0), "mrg8_outer_32");
  }
  void mrg8_outer_32(double* d_ran, const int n, const int TNUM) { mrg8_outer_32(d_ran, n, 64, TNUM); }
  void mrg8_outer_32(double* d_ran, const int n, const int BS, const int TNUM) {
    mrg8_amd_detail::check(::mrg8_outer_32(g_, d_ran, n, BS, TNUM), "mrg8_outer_32");
  }
  void state_cpy_DtH() { mrg8_amd_detail::check(::mrg8_state_cpy_DtH(g_), "state_cpy_DtH"); }

  // Stream-ordered generation in any format (no reference counterpart).
  void generate(void* dptr, uint64_t n, int format = MRG8_F64_01, void* stream = nullptr) {
    mrg8_amd_detail::check(::mrg8_generate(g_, dptr, n, format, stream), "mrg8_generate");
  }
};

#endif  // MRG8_AMD_CUDA_HPP
