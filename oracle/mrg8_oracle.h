/*
 * MRG8 oracle: a plain-C restatement of the reference's serial generator.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker that the HIP path is
 * compared against, and the CPU baseline that bench.py times beside it.  It
 * is never linked into, loaded by, or called from the product library
 * (mrg8_amd/), which must fail loudly rather than fall back to it.
 *
 * Every function cites the reference file:line it restates (paths are
 * relative to the kenmiura/mrg8 source tree).
 *
 * Pinning (what checks each function against something other than itself):
 *   mrg8o_bigdot       closed form "S==0 ? 0 : 1+(S-1) mod p" over random and
 *                      extreme inputs (Python big ints), plus the compiled
 *                      reference (oracle/_ref) golden vectors.
 *   mrg8o_seed         closed form (seed * 6364136223846793005^(k+1) mod 2^64)>>33
 *                      and golden vectors from the compiled reference.
 *   mrg8o_randint      an independent big-int recurrence (Python), golden
 *                      vectors and full-sequence hashes from the compiled reference.
 *   mrg8o_rand[01]     IEEE single-rounding products checked against Python
 *                      float arithmetic; golden bit patterns.
 *   mrg8o_jump_table   byte-identical md5 of the reference's jump_matrix.txt
 *                      when written in its text format; M[j+1] == M[j]^2.
 *   mrg8o_jump*        jump(n) == state after generating n outputs; jump(a)
 *                      then jump(b) == jump(a+b); jump(0) == identity; golden
 *                      vectors from the compiled reference.
 *   mrg8o_mrg8dnz2     equals mrg8o_rand for every n (the reference's own claim
 *                      that ver.2 is the same sequence), golden vectors.
 */
#ifndef MRG8_ORACLE_H
#define MRG8_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MRG8O_P 2147483647u          /* 2^31 - 1 */
#define MRG8O_NJUMP 247              /* A^(2^j), j = 0..246 */

/* Coefficients a0..a7 of S[n] = sum_k a_k S[n-1-k] mod p. */
extern const uint32_t mrg8o_coeff[8];

uint32_t mrg8o_bigdot(const uint32_t x[8], const uint32_t y[8]);
void mrg8o_seed(uint32_t seed, uint32_t st[8]);
void mrg8o_randint(uint32_t st[8], uint32_t* out, uint64_t n);
void mrg8o_rand(uint32_t st[8], double* out, uint64_t n);
void mrg8o_rand01(uint32_t st[8], double* out, uint64_t n);
void mrg8o_mrg8dnz2(uint32_t st[8], double* out, uint64_t n);
void mrg8o_jump_table(uint32_t* tab);
int mrg8o_write_jump_table(const uint32_t* tab, const char* path);
void mrg8o_jump_bits(const uint32_t* tab, const short bits[200], uint32_t st[8]);
void mrg8o_jump(const uint32_t* tab, uint64_t step, uint32_t st[8]);
void mrg8o_jump_vec(const uint32_t* tab, uint64_t step, const uint32_t in[8], uint32_t out[8]);

/* FNV-1a 64 over a byte buffer, continuing from h (start with MRG8O_FNV_INIT). */
#define MRG8O_FNV_INIT 0xcbf29ce484222325ull
uint64_t mrg8o_fnv1a(uint64_t h, const void* buf, size_t nbytes);

/* Convenience for hashing long sequences without holding them: generate
 * n outputs in the given format (0: f64 x/p, 1: f64 (x-1)/p, 2: u32) from st,
 * advance st, and return the FNV-1a 64 hash of the little-endian bytes. */
uint64_t mrg8o_hash_stream(uint32_t st[8], uint64_t n, int format);

/* Position-weighted checksum sum_i w_i (2i+1) mod 2^64 over n outputs of
 * esz bytes each (esz 4: words zero-extended; esz 8: IEEE bit patterns). */
uint64_t mrg8o_wsum(const void* buf, uint64_t n, int esz);

/* For each of n streams: jump states[8k..8k+7] by offsets[k], generate
 * counts[k] outputs in formats[k], write their wsum to sums[k] (OpenMP over
 * streams). */
void mrg8o_stream_wsums(const uint32_t* tab, const uint32_t* states, const uint64_t* offsets,
                        const uint64_t* counts, const int32_t* formats, uint64_t n, uint64_t* sums);

/* For each of n streams: jump states[8k..8k+7] by offsets[k] and write the
 * next k outputs in `format` to out[k*m .. k*m + m) (esz 8 for doubles, 4 for
 * words; OpenMP over streams).  The windows at given positions of many
 * streams, for checking served requests at their ends. */
void mrg8o_stream_windows(const uint32_t* tab, const uint32_t* states, const uint64_t* offsets, uint64_t n,
                          uint32_t m, int format, void* out);

/* CPU baselines (timed beside the GPU; reported, not optimised). */
/* Vector + OpenMP restatement of knl/mrg8_vec.h mrg8_vec_outer_tp; writes
 * (x-1)/p doubles, advances st by n.  Returns the number of threads used. */
int mrg8o_vec_outer_tp(const uint32_t* tab, uint32_t st[8], double* out, int64_t n, int nthreads);
/* 1 if the vector baseline was compiled with AVX-512, else 0 (scalar path). */
int mrg8o_vec_has_avx512(void);

#ifdef __cplusplus
}
#endif

#endif
