// Host-side exact arithmetic for MRG8 (mod p = 2^31 - 1): seeding, the
// reference's dot-product semantics, jump matrices, and the coefficient
// tables the gfx950 kernels consume.
//
// Representation (matches original/mrg8.cpp:72-82, bigDotProd): a dot product
// of words below 2^31 with exact integer value S yields 0 when S == 0 and
// otherwise the residue of S in [1, p] -- a residue of 0 is written as p.
#pragma once

#include <cstdint>
#include <cstring>
#include <vector>

namespace mrg8 {

constexpr uint32_t kP = 2147483647u;  // 2^31 - 1
constexpr int kNumJump = 247;         // A^(2^j), j = 0..246 (jump_matrix.txt, original/mrg8.cpp:85-87)

// Recurrence coefficients a0..a7 (original/mrg8.cpp:16-18).
constexpr uint32_t kCoeff[8] = {1089656042u, 1906537547u, 1764115693u, 1304127872u,
                                189748160u,  1984088114u, 626062218u,  1927846343u};

// Canonical residue of a 64-bit value, in [0, p).
inline uint32_t mod_p(uint64_t x) {
  x = (x & kP) + (x >> 31);  // < 2^34
  x = (x & kP) + (x >> 31);  // < 2^31 + 8
  return (uint32_t)(x >= kP ? x - kP : x);
}

inline uint32_t mul_mod(uint32_t a, uint32_t b) { return mod_p((uint64_t)a * b); }
inline uint32_t add_mod(uint32_t a, uint32_t b) { return mod_p((uint64_t)a + b); }

// The reference's dot-product semantics for words below 2^31: 0 for an exact
// sum of 0, else the residue in [1, p].
inline uint32_t dot8(const uint32_t* x, const uint32_t* y) {
  unsigned __int128 s = 0;
  for (int i = 0; i < 8; ++i) s += (uint64_t)x[i] * y[i];
  if (s == 0) return 0;
  uint64_t r = (uint64_t)((s - 1) % kP);
  return (uint32_t)(r + 1);
}

// Seeding of original/mrg8.cpp:180-194 (mcg64ni): seed 0 means 97531, then
// word k is bits 33..63 of seed * 6364136223846793005^(k+1) mod 2^64.
inline void seed_state(uint32_t seed, uint32_t st[8]) {
  uint64_t x = seed ? seed : 97531u;
  for (int k = 0; k < 8; ++k) {
    x *= 6364136223846793005ull;
    st[k] = (uint32_t)(x >> 33);
  }
}

// One step of the recurrence (original/mrg8.cpp:208-220): returns the new word
// and shifts it into the newest-first state.
inline uint32_t step1(uint32_t st[8]) {
  uint32_t y = dot8(kCoeff, st);
  std::memmove(st + 1, st, 7 * sizeof(uint32_t));
  st[0] = y;
  return y;
}

// A^(2^j), j = 0..kNumJump-1, row-major: m[j][r][c].
struct JumpTable {
  std::vector<uint32_t> m;  // kNumJump * 64
  JumpTable() : m((size_t)kNumJump * 64, 0) {
    uint32_t* a = m.data();
    for (int c = 0; c < 8; ++c) a[c] = kCoeff[c];
    for (int r = 1; r < 8; ++r) a[r * 8 + r - 1] = 1;
    for (int j = 1; j < kNumJump; ++j) {
      const uint32_t* s = at(j - 1);
      uint32_t* d = m.data() + (size_t)j * 64;
      for (int r = 0; r < 8; ++r)
        for (int c = 0; c < 8; ++c) {
          uint32_t col[8];
          for (int q = 0; q < 8; ++q) col[q] = s[q * 8 + c];
          d[r * 8 + c] = dot8(s + r * 8, col);
        }
    }
  }
  const uint32_t* at(int j) const { return m.data() + (size_t)j * 64; }
};

// Row-major 8x8 product with the reference's dot-product semantics.
inline std::vector<uint32_t> mat_mul(const uint32_t* x, const uint32_t* y) {
  std::vector<uint32_t> out(64);
  for (int r = 0; r < 8; ++r)
    for (int c = 0; c < 8; ++c) {
      uint32_t col[8];
      for (int q = 0; q < 8; ++q) col[q] = y[q * 8 + c];
      out[(size_t)r * 8 + c] = dot8(x + r * 8, col);
    }
  return out;
}

// state <- A^(2^b) state with the reference's per-row semantics.
inline void apply_pow2(const JumpTable& t, int b, uint32_t st[8]) {
  const uint32_t* m = t.at(b);
  uint32_t ns[8];
  for (int r = 0; r < 8; ++r) ns[r] = dot8(m + r * 8, st);
  std::memcpy(st, ns, sizeof ns);
}

// Jump by a 64-bit step as a chain of matrix-vector products (the form of
// knl/mrg8.h:212-238).  For tables without accidental zero entries this
// equals the reference's matrix-product form bit for bit (original/mrg8.cpp:105-176):
// the exact zero pattern of a result only depends on which entries are
// structurally non-zero, and that pattern is the same for both orders.
inline void jump64(const JumpTable& t, uint64_t step, uint32_t st[8]) {
  for (int b = 0; b < 64; ++b)
    if ((step >> b) & 1u) apply_pow2(t, b, st);
}

// 200-bit jump, bits[0] least significant (original/mrg8.cpp:105-122).
inline void jump200(const JumpTable& t, const short bits[200], uint32_t st[8]) {
  for (int b = 0; b < 200; ++b)
    if (bits[b]) apply_pow2(t, b, st);
}

// Row 0 of A^(k+1) for k = 0..count-1: output k of a block generated from a
// newest-first state s is dot(row_k, s).  Canonical residues in [0, p).
inline std::vector<uint32_t> block_rows(int count) {
  std::vector<uint32_t> rows((size_t)count * 8);
  uint32_t v[8];
  for (int c = 0; c < 8; ++c) v[c] = kCoeff[c];
  for (int k = 0; k < count; ++k) {
    std::memcpy(&rows[(size_t)k * 8], v, sizeof v);
    // v <- v * A:  (vA)_c = v_0 a_c + v_{c+1}, (vA)_7 = v_0 a_7.
    uint32_t nv[8];
    for (int c = 0; c < 8; ++c) {
      uint32_t t = mul_mod(v[0], kCoeff[c]);
      nv[c] = c < 7 ? add_mod(t, v[c + 1]) : t;
    }
    std::memcpy(v, nv, sizeof v);
  }
  return rows;
}

// ---- device coefficient records -------------------------------------------
//
// The kernels evaluate dot(row, s) for a wave-uniform state s with eight
// signed 32x32->64 multiply-adds.  Both sides are centred so that the exact
// 64-bit sum cannot wrap:
//   c_j = row_j       if row_j <= (p-1)/2, else row_j - p   (|c_j| < 2^30)
//   e_j = s_j - 2^30                                         (|e_j| <= 2^30)
//   V   = B0 + sum_j c_j e_j,  B0 = (largest multiple of p below 2^63)
//                                  + (2^30 sum_j c_j mod p) - 1
// so V is in (0, 2^64) and V + 1 == dot(row, s) (mod p).  The canonical
// residue r of V gives the reference's word directly as r + 1, in [1, p]
// (a residue of 0 becomes p), except when the exact sum is 0: that happens
// exactly when every word of s under a non-zero entry of the row is 0
// (`support` marks the non-zero entries), and the word is then 0.
constexpr uint32_t kCenter = 1u << 30;

struct CoefRecord {   // 48 bytes, the layout the kernels read
  int32_t c[8];
  uint32_t b0_lo, b0_hi;
  uint32_t support;  // bit j set when entry j is non-zero as an integer
  uint32_t pad;
};
static_assert(sizeof(CoefRecord) == 48, "record layout");

inline CoefRecord make_record(const uint32_t row[8]) {
  CoefRecord rec{};
  uint64_t sum_c = 0;  // sum of c_j mod p, kept non-negative
  for (int j = 0; j < 8; ++j) {
    uint32_t r = row[j] % kP;
    int32_t c = r <= (kP - 1) / 2 ? (int32_t)r : (int32_t)((int64_t)r - kP);
    rec.c[j] = c;
    sum_c += (uint64_t)(c < 0 ? (int64_t)c + kP : c);
  }
  uint32_t k = mul_mod(mod_p(sum_c), kCenter);  // 2^30 * sum c mod p
  const uint64_t two63 = 1ull << 63;
  uint64_t base = two63 - (two63 % kP);
  uint64_t b0 = base + k - 1;
  rec.b0_lo = (uint32_t)b0;
  rec.b0_hi = (uint32_t)(b0 >> 32);
  rec.support = 0;
  for (int j = 0; j < 8; ++j)
    if (row[j] != 0) rec.support |= 1u << j;
  rec.pad = 0;
  return rec;
}

inline bool all_nonzero(const uint32_t row[8]) {
  for (int j = 0; j < 8; ++j)
    if (row[j] % kP == 0) return false;
  return true;
}

}  // namespace mrg8
