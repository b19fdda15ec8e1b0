/*
 * CPU baseline: vector + OpenMP restatement of the reference's knl path
 * (knl/mrg8_vec.h mrg8_vec_outer / mrg8_vec_outer_tp, knl/mrg8.h jump).
 *
 * TEST INFRASTRUCTURE ONLY: timed beside the GPU as a reported baseline; it
 * is never linked into the product library.
 *
 * Algorithm (knl/mrg8_vec.h:305-451): the state is held oldest-first in a
 * vector t (t[d] = S[n-8+d]).  Eight outputs at a time come from A^8:
 * S[n+i] = sum_d A8(7-i, 7-d) t[d], with the two 4-term 64-bit partial sums
 * and Mersenne folds of the serial generator, so the new oldest-first state
 * is the output vector itself.  Outputs use the [0,1) conversion
 * double(x-1) * (1/p) of knl/mrg8_vec.h:362-366.
 *
 * Threading (knl/mrg8_vec.h:500-525): thread k starts at k*each where
 * each = ((n / T) / 8) * 8, the last thread takes the rest; each thread
 * jumps with the matrix-vector chain of knl/mrg8.h:212-238.
 *
 * Deliberate difference: the reference's final partial block folds only once
 * (knl/mrg8_vec.h:395-399), which can leave a word above p; here the last
 * (n mod 8, or 8) outputs of each thread go through the serial recurrence,
 * so every output matches the serial generator.  The timing is unaffected.
 */
#include "mrg8_oracle.h"

#include <immintrin.h>
#include <omp.h>
#include <string.h>

static const uint64_t M31 = 2147483647ull;

/* B[d][i] = A^8(7-i, 7-d): column d of the oldest-first update. */
static void build_b(const uint32_t* tab, uint64_t B[8][8]) {
  const uint32_t* a8 = tab + 64 * 3; /* A^(2^3), column-major */
  for (int d = 0; d < 8; ++d)
    for (int i = 0; i < 8; ++i) B[d][i] = a8[(7 - i) + 8 * (7 - d)];
}

static inline uint64_t fold(uint64_t s) { return (s & M31) + (s >> 31); }

/* Scalar form of the outer-product loop (the reference's non-AVX branch,
 * knl/mrg8_vec.h:416-450); generates 8*blocks outputs. */
static void outer_scalar(const uint64_t B[8][8], uint32_t t[8], double* out, int64_t blocks) {
  const double rnorm = 1.0 / 2147483647.0;
  for (int64_t b = 0; b < blocks; ++b) {
    uint64_t s1[8] = {0}, s2[8] = {0};
    for (int d = 0; d < 4; ++d)
      for (int i = 0; i < 8; ++i) {
        s1[i] += B[d][i] * t[d];
        s2[i] += B[d + 4][i] * t[d + 4];
      }
    for (int i = 0; i < 8; ++i) {
      uint64_t s = fold(s1[i]) + fold(s2[i]);
      s = fold(fold(s));
      t[i] = (uint32_t)s;
      out[8 * b + i] = (double)(t[i] - 1u) * rnorm;
    }
  }
}

__attribute__((target("avx512f,avx512dq")))
static void outer_avx512(const uint64_t B[8][8], uint32_t t[8], double* out, int64_t blocks) {
  __m512i col[8];
  for (int d = 0; d < 8; ++d) col[d] = _mm512_loadu_si512((const void*)B[d]);
  const __m512i mask = _mm512_set1_epi64((long long)M31);
  const __m512i one = _mm512_set1_epi64(1);
  const __m512d rn = _mm512_set1_pd(1.0 / 2147483647.0);
  uint64_t t64[8];
  for (int i = 0; i < 8; ++i) t64[i] = t[i];
  __m512i tv = _mm512_loadu_si512((const void*)t64);
  for (int64_t b = 0; b < blocks; ++b) {
    __m512i s1 = _mm512_setzero_si512(), s2 = _mm512_setzero_si512();
    for (int d = 0; d < 4; ++d) {
      __m512i x = _mm512_permutexvar_epi64(_mm512_set1_epi64(d), tv);
      s1 = _mm512_add_epi64(s1, _mm512_mul_epu32(col[d], x));
      __m512i y = _mm512_permutexvar_epi64(_mm512_set1_epi64(d + 4), tv);
      s2 = _mm512_add_epi64(s2, _mm512_mul_epu32(col[d + 4], y));
    }
    s1 = _mm512_add_epi64(_mm512_and_si512(s1, mask), _mm512_srli_epi64(s1, 31));
    s2 = _mm512_add_epi64(_mm512_and_si512(s2, mask), _mm512_srli_epi64(s2, 31));
    __m512i s = _mm512_add_epi64(s1, s2);
    s = _mm512_add_epi64(_mm512_and_si512(s, mask), _mm512_srli_epi64(s, 31));
    s = _mm512_add_epi64(_mm512_and_si512(s, mask), _mm512_srli_epi64(s, 31));
    tv = s;
    __m512d r = _mm512_mul_pd(_mm512_cvtepi64_pd(_mm512_sub_epi64(s, one)), rn);
    _mm512_storeu_pd(out + 8 * b, r);
  }
  _mm512_storeu_si512((void*)t64, tv);
  for (int i = 0; i < 8; ++i) t[i] = (uint32_t)t64[i];
}

static int have_avx512(void) {
  __builtin_cpu_init();
  return __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512dq");
}

int mrg8o_vec_has_avx512(void) { return have_avx512(); }

/* One thread's share: st is newest-first on entry and exit. */
static void outer_range(const uint64_t B[8][8], int avx, uint32_t st[8], double* out, int64_t n) {
  if (n <= 0) return;
  int64_t tail = n % 8 ? n % 8 : 8; /* the reference's loop runs while i < n - 8 */
  int64_t blocks = (n - tail) / 8;
  uint32_t t[8];
  for (int i = 0; i < 8; ++i) t[i] = st[7 - i];
  if (avx)
    outer_avx512(B, t, out, blocks);
  else
    outer_scalar(B, t, out, blocks);
  for (int i = 0; i < 8; ++i) st[i] = t[7 - i];
  mrg8o_rand01(st, out + 8 * blocks, (uint64_t)tail);
}

int mrg8o_vec_outer_tp(const uint32_t* tab, uint32_t st[8], double* out, int64_t n, int nthreads) {
  uint64_t B[8][8];
  build_b(tab, B);
  int avx = have_avx512();
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  uint32_t next[8];
  memcpy(next, st, sizeof next);
  int used = nthreads;
#pragma omp parallel num_threads(nthreads)
  {
    int T = omp_get_num_threads();
    int k = omp_get_thread_num();
#pragma omp single
    used = T;
    int64_t each = ((n / T) / 8) * 8;
    int64_t start = each * k;
    int64_t len = (k == T - 1) ? n - start : each;
    uint32_t mine[8];
    mrg8o_jump_vec(tab, (uint64_t)start, st, mine);
    outer_range(B, avx, mine, out + start, len);
    if (k == T - 1) memcpy(next, mine, sizeof next);
  }
  memcpy(st, next, sizeof next);
  return used;
}
