/*
 * MRG8 oracle: plain-C restatement of the reference's serial generator
 * (original/mrg8.cpp, cuda/mrg8.cpp, knl/mrg8.h).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the CPU
 * baseline timed beside it.  Nothing in the product library links or calls
 * this file.  See mrg8_oracle.h for how each function is pinned.
 */
#include "mrg8_oracle.h"

#include <omp.h>
#include <stdio.h>
#include <string.h>

/* original/mrg8.cpp:16-18 (COEFF0..COEFF7). */
const uint32_t mrg8o_coeff[8] = {1089656042u, 1906537547u, 1764115693u, 1304127872u,
                                 189748160u,  1984088114u, 626062218u,  1927846343u};

static const uint64_t MASK31 = 2147483647ull;

/* original/mrg8.cpp:72-82.  Two 4-term 64-bit partial sums, then three
 * Mersenne folds with no final conditional subtract.  For inputs below 2^31
 * the result is 0 when every product is 0, else the residue in [1, p]. */
uint32_t mrg8o_bigdot(const uint32_t x[8], const uint32_t y[8]) {
  uint64_t lo = 0, hi = 0;
  for (int q = 0; q < 4; ++q) {
    lo += (uint64_t)x[q] * y[q];
    hi += (uint64_t)x[q + 4] * y[q + 4];
  }
  uint64_t s = (lo & MASK31) + (lo >> 31) + (hi & MASK31) + (hi >> 31);
  s = (s & MASK31) + (s >> 31);
  s = (s & MASK31) + (s >> 31);
  return (uint32_t)s;
}

/* original/mrg8.cpp:180-194 (mcg64ni): seed 0 becomes 97531, then eight steps
 * of a 64-bit multiplicative congruential generator, keeping bits 33..63. */
void mrg8o_seed(uint32_t seed, uint32_t st[8]) {
  uint64_t x = seed ? seed : 97531u;
  for (int k = 0; k < 8; ++k) {
    x *= 6364136223846793005ull;
    st[k] = (uint32_t)(x >> 33);
  }
}

/* original/mrg8.cpp:197-222 (randint): y = bigdot(a, state); shift the state
 * (state[0] is the newest word) and emit y. */
void mrg8o_randint(uint32_t st[8], uint32_t* out, uint64_t n) {
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t y = mrg8o_bigdot(mrg8o_coeff, st);
    for (int k = 7; k > 0; --k) st[k] = st[k - 1];
    st[0] = y;
    out[i] = y;
  }
}

/* original/mrg8.cpp:231-239 (rand): double(x) * (1.0 / 2147483647.0), one
 * rounding of the product with the rounded reciprocal (not x / p).  Range (0, 1]. */
void mrg8o_rand(uint32_t st[8], double* out, uint64_t n) {
  const double rnorm = 1.0 / 2147483647.0;
  uint32_t buf[256];
  for (uint64_t i = 0; i < n;) {
    uint64_t m = n - i < 256 ? n - i : 256;
    mrg8o_randint(st, buf, m);
    for (uint64_t k = 0; k < m; ++k) out[i + k] = (double)buf[k] * rnorm;
    i += m;
  }
}

/* The [0, 1) conversion of the CUDA kernels and the knl path:
 * double(x - 1) * (1.0 / 2147483647.0)  (cuda/mrg8_cuda.cu:363,390; knl/mrg8.h:359). */
void mrg8o_rand01(uint32_t st[8], double* out, uint64_t n) {
  const double rnorm = 1.0 / 2147483647.0;
  uint32_t buf[256];
  for (uint64_t i = 0; i < n;) {
    uint64_t m = n - i < 256 ? n - i : 256;
    mrg8o_randint(st, buf, m);
    for (uint64_t k = 0; k < m; ++k) out[i + k] = (double)(uint32_t)(buf[k] - 1u) * rnorm;
    i += m;
  }
}

/* cuda/mrg8.cpp:290-339 (mrg8dnz2): the same sequence produced in blocks of
 * 1024 through a scratch array that runs backwards (x[k] is newer than
 * x[k+1]); the n % 1024 tail goes through rand().  Conversion x * (1/p). */
void mrg8o_mrg8dnz2(uint32_t st[8], double* out, uint64_t n) {
  enum { KMAX = 1024 };
  const double rnorm = 1.0 / 2147483647.0;
  uint32_t x[KMAX + 8];
  for (int k = 0; k < 8; ++k) x[KMAX + k] = st[k];
  uint64_t tail = n % KMAX;
  for (uint64_t j = 0; j < n - tail; j += KMAX) {
    for (int k = KMAX - 1; k >= 0; --k) {
      x[k] = mrg8o_bigdot(mrg8o_coeff, &x[k + 1]);
      out[j + KMAX - 1 - k] = (double)x[k] * rnorm;
    }
    for (int k = 0; k < 8; ++k) x[KMAX + k] = x[k];
  }
  for (int k = 0; k < 8; ++k) st[k] = x[KMAX + k];
  if (tail) mrg8o_rand(st, out + (n - tail), tail);
}

/* The table that jump_matrix.txt holds (original/mrg8.cpp:85-87): entry j is
 * A^(2^j) mod p stored column-major, (A^(2^j))(r,c) = tab[64 j + r + 8 c].
 * A is the companion matrix of the recurrence: row 0 holds the coefficients,
 * A(r, r-1) = 1.  Here it is rebuilt by repeated squaring with bigdot. */
void mrg8o_jump_table(uint32_t* tab) {
  memset(tab, 0, sizeof(uint32_t) * 64 * MRG8O_NJUMP);
  for (int c = 0; c < 8; ++c) tab[0 + 8 * c] = mrg8o_coeff[c];
  for (int r = 1; r < 8; ++r) tab[r + 8 * (r - 1)] = 1;
  for (int j = 1; j < MRG8O_NJUMP; ++j) {
    const uint32_t* m = tab + 64 * (j - 1);
    uint32_t* o = tab + 64 * j;
    for (int r = 0; r < 8; ++r) {
      uint32_t row[8];
      for (int q = 0; q < 8; ++q) row[q] = m[r + 8 * q];
      for (int c = 0; c < 8; ++c) o[r + 8 * c] = mrg8o_bigdot(row, m + 8 * c);
    }
  }
}

/* Text layout of jump_matrix.txt: one matrix per line, 64 values each printed
 * right-aligned in 10 columns followed by two spaces. */
int mrg8o_write_jump_table(const uint32_t* tab, const char* path) {
  FILE* f = fopen(path, "w");
  if (!f) return -1;
  for (int j = 0; j < MRG8O_NJUMP; ++j) {
    for (int k = 0; k < 64; ++k) fprintf(f, "%10u  ", tab[64 * j + k]);
    fputc('\n', f);
  }
  return fclose(f) == 0 ? 0 : -1;
}

/* original/mrg8.cpp:105-122 and 145-176: build J as the product of the
 * table matrices for the set bits (J <- J * M[b], lowest bit first, each
 * entry a bigdot), then state <- J * state.  Mutates st. */
void mrg8o_jump_bits(const uint32_t* tab, const short bits[200], uint32_t st[8]) {
  uint32_t J[8][8];
  for (int r = 0; r < 8; ++r)
    for (int c = 0; c < 8; ++c) J[r][c] = (r == c);
  for (int b = 0; b < 200; ++b) {
    if (!bits[b]) continue;
    const uint32_t* m = tab + 64 * b;
    uint32_t T[8][8];
    for (int r = 0; r < 8; ++r)
      for (int c = 0; c < 8; ++c) T[r][c] = mrg8o_bigdot(J[r], m + 8 * c);
    memcpy(J, T, sizeof J);
  }
  uint32_t ns[8];
  for (int r = 0; r < 8; ++r) ns[r] = mrg8o_bigdot(J[r], st);
  memcpy(st, ns, sizeof ns);
}

/* original/mrg8.cpp:125-141: the low 64 bits of a 200-bit value. */
void mrg8o_jump(const uint32_t* tab, uint64_t step, uint32_t st[8]) {
  short bits[200];
  memset(bits, 0, sizeof bits);
  for (int b = 0; b < 64; ++b) bits[b] = (short)((step >> b) & 1u);
  mrg8o_jump_bits(tab, bits, st);
}

/* knl/mrg8.h:212-238: the same jump as a chain of matrix-vector products
 * (one per set bit), writing the result to out and leaving in unchanged. */
void mrg8o_jump_vec(const uint32_t* tab, uint64_t step, const uint32_t in[8], uint32_t out[8]) {
  uint32_t v[8];
  memcpy(v, in, sizeof v);
  for (int b = 0; b < 64; ++b) {
    if (!((step >> b) & 1u)) continue;
    const uint32_t* m = tab + 64 * b;
    uint32_t nv[8];
    for (int r = 0; r < 8; ++r) {
      uint32_t row[8];
      for (int q = 0; q < 8; ++q) row[q] = m[r + 8 * q];
      nv[r] = mrg8o_bigdot(row, v);
    }
    memcpy(v, nv, sizeof v);
  }
  memcpy(out, v, sizeof v);
}

/* FNV-1a, 64-bit, over bytes. */
uint64_t mrg8o_fnv1a(uint64_t h, const void* buf, size_t nbytes) {
  const unsigned char* p = (const unsigned char*)buf;
  for (size_t i = 0; i < nbytes; ++i) {
    h ^= p[i];
    h *= 0x100000001b3ull;
  }
  return h;
}

uint64_t mrg8o_hash_stream(uint32_t st[8], uint64_t n, int format) {
  uint64_t h = MRG8O_FNV_INIT;
  enum { B = 4096 };
  double d[B];
  uint32_t u[B];
  for (uint64_t i = 0; i < n;) {
    uint64_t m = n - i < B ? n - i : B;
    if (format == 0) {
      mrg8o_rand(st, d, m);
      h = mrg8o_fnv1a(h, d, m * sizeof(double));
    } else if (format == 1) {
      mrg8o_rand01(st, d, m);
      h = mrg8o_fnv1a(h, d, m * sizeof(double));
    } else {
      mrg8o_randint(st, u, m);
      h = mrg8o_fnv1a(h, u, m * sizeof(uint32_t));
    }
    i += m;
  }
  return h;
}

uint64_t mrg8o_wsum(const void* buf, uint64_t n, int esz) {
  uint64_t total = 0;
#pragma omp parallel for reduction(+ : total) schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    uint64_t w;
    if (esz == 4)
      w = ((const uint32_t*)buf)[i];
    else
      memcpy(&w, (const char*)buf + 8 * (size_t)i, 8);
    total += w * (2 * (uint64_t)i + 1);
  }
  return total;
}

void mrg8o_stream_wsums(const uint32_t* tab, const uint32_t* states, const uint64_t* offsets,
                        const uint64_t* counts, const int32_t* formats, uint64_t n, uint64_t* sums) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t k = 0; k < (int64_t)n; ++k) {
    uint32_t st[8];
    mrg8o_jump_vec(tab, offsets[k], states + 8 * k, st);
    enum { B = 4096 };
    double d[B];
    uint32_t u[B];
    uint64_t total = 0, pos = 0;
    while (pos < counts[k]) {
      uint64_t m = counts[k] - pos < B ? counts[k] - pos : B;
      if (formats[k] == 2) {
        mrg8o_randint(st, u, m);
        for (uint64_t i = 0; i < m; ++i) total += (uint64_t)u[i] * (2 * (pos + i) + 1);
      } else {
        if (formats[k] == 0)
          mrg8o_rand(st, d, m);
        else
          mrg8o_rand01(st, d, m);
        for (uint64_t i = 0; i < m; ++i) {
          uint64_t w;
          memcpy(&w, &d[i], 8);
          total += w * (2 * (pos + i) + 1);
        }
      }
      pos += m;
    }
    sums[k] = total;
  }
}

void mrg8o_stream_windows(const uint32_t* tab, const uint32_t* states, const uint64_t* offsets, uint64_t n,
                          uint32_t m, int format, void* out) {
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t k = 0; k < (int64_t)n; ++k) {
    uint32_t st[8];
    mrg8o_jump_vec(tab, offsets[k], states + 8 * k, st);
    if (format == 2)
      mrg8o_randint(st, (uint32_t*)out + (size_t)k * m, m);
    else if (format == 0)
      mrg8o_rand(st, (double*)out + (size_t)k * m, m);
    else
      mrg8o_rand01(st, (double*)out + (size_t)k * m, m);
  }
}
