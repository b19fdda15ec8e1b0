// Driver for the reference generator compiled from its own sources (test
// infrastructure only; see the Makefile).  Built three times, against
// original/mrg8.cpp (REF_ORIGINAL), cuda/mrg8.cpp (REF_CUDA_HOST) and the
// header-only knl/mrg8.h (REF_KNL).  Run it in a directory holding
// jump_matrix.txt; it reads commands on stdin, one per line, and answers one
// line per command:
//
//   S <seed>                          -> 8 state words after construction with <seed>
//   G <seed|-> <w0..w7> <step> <n> <fmt> <k>
//        start from the state after seeding (or set_state(w) when seed is "-"),
//        jump_ahead(step), generate n values in format fmt, print
//        "<fnv1a64 hex> <first k values> | <final state>".
//        fmt 0: rand() doubles, 2: randint(), 3: mrg8dnz2(); doubles print as
//        16-hex-digit bit patterns.
//   B <seed> <bit,bit,...>            -> state after the 200-bit jump_ahead with those bits set
//   T <seed> <n> <iters>              -> per-iteration milliseconds of rand(buf, n)
//   Q                                 -> exit
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <string>
#include <vector>

#include "mrg8.h"

namespace {

uint64_t fnv(uint64_t h, const void* p, size_t n) {
  const unsigned char* b = static_cast<const unsigned char*>(p);
  for (size_t i = 0; i < n; ++i) {
    h ^= b[i];
    h *= 0x100000001b3ull;
  }
  return h;
}

// The mutating jump of original/ versus the new_state form of cuda/ and knl/.
void jump(mrg8& m, uint64_t step) {
#if defined(REF_ORIGINAL)
  m.jump_ahead(step);
#else
  uint32_t ns[8];
  m.jump_ahead(step, ns);
  m.set_state(ns);
#endif
}

void jump_bits(mrg8& m, const short bits[200]) {
#if defined(REF_ORIGINAL)
  m.jump_ahead(bits);
#else
  uint32_t ns[8];
  m.jump_ahead(bits, ns);
  m.set_state(ns);
#endif
}

void print_state(const mrg8& m) {
  uint32_t st[8];
  m.get_state(st);
  for (int i = 0; i < 8; ++i) printf(" %u", st[i]);
}

void generate(mrg8& m, uint64_t n, int fmt, int k) {
  const uint64_t CH = 1u << 20;
  uint64_t h = 0xcbf29ce484222325ull;
  std::vector<double> d;
  std::vector<uint32_t> u;
  std::string head;
  char buf[64];
  if (fmt == 3) {
    // mrg8dnz2 in one call: its block structure depends on n.
    d.resize(n ? n : 1);
#if defined(REF_ORIGINAL)
    fprintf(stderr, "mrg8dnz2 is not part of original/\n");
    exit(3);
#else
    m.mrg8dnz2(d.data(), (int)n);
#endif
    h = fnv(h, d.data(), n * sizeof(double));
    for (uint64_t i = 0; i < n && (int)i < k; ++i) {
      uint64_t bits;
      memcpy(&bits, &d[i], 8);
      snprintf(buf, sizeof buf, " %016llx", (unsigned long long)bits);
      head += buf;
    }
  } else {
    for (uint64_t done = 0; done < n;) {
      uint64_t c = n - done < CH ? n - done : CH;
      if (fmt == 2) {
        u.resize(c);
        m.randint(u.data(), (int)c);
        h = fnv(h, u.data(), c * sizeof(uint32_t));
        for (uint64_t i = 0; i < c && (int)(done + i) < k; ++i) {
          snprintf(buf, sizeof buf, " %u", u[i]);
          head += buf;
        }
      } else {
        d.resize(c);
        m.rand(d.data(), (int)c);
        h = fnv(h, d.data(), c * sizeof(double));
        for (uint64_t i = 0; i < c && (int)(done + i) < k; ++i) {
          uint64_t bits;
          memcpy(&bits, &d[i], 8);
          snprintf(buf, sizeof buf, " %016llx", (unsigned long long)bits);
          head += buf;
        }
      }
      done += c;
    }
  }
  printf("%016llx%s |", (unsigned long long)h, head.c_str());
  print_state(m);
  printf("\n");
}

}  // namespace

int main() {
  char line[8192];
  while (fgets(line, sizeof line, stdin)) {
    char* save = nullptr;
    char* cmd = strtok_r(line, " \t\n", &save);
    if (!cmd) continue;
    if (!strcmp(cmd, "Q")) break;
    if (!strcmp(cmd, "S")) {
      uint32_t seed = (uint32_t)strtoul(strtok_r(nullptr, " \t\n", &save), nullptr, 10);
      mrg8 m(seed);
      print_state(m);
      printf("\n");
    } else if (!strcmp(cmd, "G")) {
      const char* s = strtok_r(nullptr, " \t\n", &save);
      uint32_t w[8];
      for (int i = 0; i < 8; ++i) w[i] = (uint32_t)strtoul(strtok_r(nullptr, " \t\n", &save), nullptr, 10);
      uint64_t step = strtoull(strtok_r(nullptr, " \t\n", &save), nullptr, 10);
      uint64_t n = strtoull(strtok_r(nullptr, " \t\n", &save), nullptr, 10);
      int fmt = atoi(strtok_r(nullptr, " \t\n", &save));
      int k = atoi(strtok_r(nullptr, " \t\n", &save));
      if (strcmp(s, "-") == 0) {
        mrg8 m(1);
        m.set_state(w);
        if (step) jump(m, step);
        generate(m, n, fmt, k);
      } else {
        mrg8 m((uint32_t)strtoul(s, nullptr, 10));
        if (step) jump(m, step);
        generate(m, n, fmt, k);
      }
    } else if (!strcmp(cmd, "B")) {
      uint32_t seed = (uint32_t)strtoul(strtok_r(nullptr, " \t\n", &save), nullptr, 10);
      short bits[200];
      memset(bits, 0, sizeof bits);
      const char* list = strtok_r(nullptr, " \t\n", &save);
      if (list && strcmp(list, "-") != 0) {
        std::string l(list);
        size_t pos = 0;
        while (pos < l.size()) {
          size_t e = l.find(',', pos);
          if (e == std::string::npos) e = l.size();
          int b = atoi(l.substr(pos, e - pos).c_str());
          if (b >= 0 && b < 200) bits[b] = 1;
          pos = e + 1;
        }
      }
      mrg8 m(seed);
      jump_bits(m, bits);
      print_state(m);
      printf("\n");
    } else if (!strcmp(cmd, "T")) {
      uint32_t seed = (uint32_t)strtoul(strtok_r(nullptr, " \t\n", &save), nullptr, 10);
      int n = atoi(strtok_r(nullptr, " \t\n", &save));
      int iters = atoi(strtok_r(nullptr, " \t\n", &save));
      std::vector<double> d(n > 0 ? n : 1);
      mrg8 m(seed);
      for (int it = 0; it < iters; ++it) {
        m.seed_init(seed);
        auto t0 = std::chrono::steady_clock::now();
        m.rand(d.data(), n);
        auto t1 = std::chrono::steady_clock::now();
        printf(" %.6f", std::chrono::duration<double, std::milli>(t1 - t0).count());
      }
      printf("\n");
    } else {
      fprintf(stderr, "unknown command %s\n", cmd);
      return 2;
    }
    fflush(stdout);
  }
  return 0;
}
