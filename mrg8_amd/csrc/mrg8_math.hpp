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

// ---- matrix-core fill tables ------------------------------------------------
//
// The int8 MFMA form of a step (16 outputs of each of 16 columns, output j
// being dot(row_j, s) for the column's newest-first state s): with w_q =
// s_(7-q) (oldest first) and R_j[q] = row_j[7-q], each state word is
// centred (e = w - 2^30) and cut into balanced base-256 digits d_b
// (e = sum_b 2^(8b) d_b, d_b in [-128, 127]); M_b[j][q] = 2^(8b) R_j[q] mod p,
// centred, is cut the same way into digits m_(b,a).  Then
//   dot(row_j, s) - 1 == sum_a 2^(8a) C_a[j] + kappa_j  (mod p),
//   C_a[j] = sum_(q,b) m_(b,a)[j][q] d_b[q],  kappa_j = 2^30 sum_q R_j[q] - 1,
// |C_a| <= 32 * 2^14 = 2^19.
// Layout (v_mfma_i32_16x16x32_i8): lane l = n + 16 g holds the 8 bytes of
// K-group g for row/column n, result register i of lane l is row 4 g + i of
// column n.  K-group g, byte i is word q = W(g) + i / 4, digit b = i % 4,
// with W = (0, 4, 2, 6); row 4 g + i is step position P(g, i): 2 g + i for
// i < 2, 8 + W(g) + i - 2 for i >= 2.  So the step's last eight outputs
// (positions 8..15, the column's next state) land in registers 2 and 3 of
// exactly the lanes whose K-group needs them.
//
// Step constants: the MFMAs start from per-row accumulator values c_a (the
// C operand), so that t = C_0 + 2^8 C_1 and u = C_2 + 2^8 C_3 come out in
// [0, 2 beta + 2^16) and [0, 2 beta + 2^15) and t + 2^16 u == the word's
// residue + kappa_j:  with D = kappa_j - 65537 beta (mod p),
//   t0 = beta + (D & 0xffff), u0 = beta + (D >> 16),
//   c_0 = t0 & 0xff, c_1 = t0 >> 8, c_2 = u0 & 0xff, c_3 = u0 >> 8.
// The kernel then reduces x = t + ((u << 16) & p) + (u >> 15) (2^16 u mod p
// is a 31-bit rotation of u), x < 2 p.
//
// Stride tables: A^steps in the same digit form, row 4 g + i (i < 2) giving
// new word W(g) + i (oldest first), rows with i >= 2 zero, so that the new
// words land in registers 0 and 1 of the lanes whose K-group needs them;
// per lane group g and register i < 2, K in [beta, beta + p) with
// K == kappa - 2^16 beta (mod p), for the combine t + K + 2^16 (u + beta) of
// accumulators that start from 0.
//
// Fan-out records: the start state of column c >= 1 of a tile is the tile's
// outputs L c - 8 .. L c - 1 (L outputs per column), row 0 of A^(k+1) for
// k = L c - 8 + q.
constexpr int kMfmaWordOf[4] = {0, 4, 2, 6};  // W(g)

inline int mfma_position(int g, int i) { return i < 2 ? 2 * g + i : 8 + kMfmaWordOf[g] + i - 2; }

struct MfmaHostTables {
  std::vector<uint32_t> a;      // [class][lane][2] words of packed digits
  std::vector<int32_t> c;       // [class][lane group g][register i]
  std::vector<CoefRecord> fan;  // [(c - 1) * 8 + q]
};

struct MfmaStrideTables {
  std::vector<uint32_t> a;  // [class][lane][2]
  std::vector<uint32_t> k;  // [g][i], i < 2
};

inline uint32_t centred_digits(int32_t m) { return (uint32_t)m + 0x80808080u; }  // bytes e_a: digit e_a - 128

// Packed digit bytes of class a for K-group g of the row vector R
// (oldest-first words): byte 4 v + b is digit a of centred(2^(8b) R[W(g) + v] mod p).
inline void mfma_operand(const uint32_t R[8], int g, int a, uint32_t out[2]) {
  for (int v = 0; v < 2; ++v) {
    uint32_t word = 0;
    for (int b = 0; b < 4; ++b) {
      const uint32_t m = mul_mod(R[kMfmaWordOf[g] + v] % kP, 1u << (8 * b));
      const int32_t mc = m <= (kP - 1) / 2 ? (int32_t)m : (int32_t)((int64_t)m - kP);
      const uint32_t byte = ((centred_digits(mc) >> (8 * a)) & 0xffu) ^ 0x80u;
      word |= byte << (8 * b);
    }
    out[v] = word;
  }
}

inline uint32_t mfma_kappa(const uint32_t R[8]) {
  uint32_t sum = 0;
  for (int q = 0; q < 8; ++q) sum = add_mod(sum, R[q] % kP);
  return add_mod(mul_mod(sum, 1u << 30), kP - 1);  // 2^30 sum - 1
}

inline bool mfma_host_tables(int columns, int column, int classes, uint32_t beta, MfmaHostTables* out) {
  const std::vector<uint32_t> rows = block_rows(columns * column);
  auto Rrow = [&](int j, uint32_t R[8]) {
    for (int q = 0; q < 8; ++q) R[q] = rows[(size_t)j * 8 + 7 - q];
  };
  out->a.assign((size_t)classes * 64 * 2, 0);
  for (int a = 0; a < classes; ++a)
    for (int l = 0; l < 64; ++l) {
      const int m = l & 15;  // row m = 4 g' + i' is position P(g', i')
      uint32_t R[8];
      Rrow(mfma_position(m >> 2, m & 3), R);
      mfma_operand(R, l >> 4, a, &out->a[((size_t)a * 64 + l) * 2]);
    }
  out->c.assign((size_t)classes * 16, 0);
  const uint32_t b65537 = mul_mod(mod_p(beta), 65537u);
  for (int g = 0; g < 4; ++g)
    for (int i = 0; i < 4; ++i) {
      uint32_t R[8];
      Rrow(mfma_position(g, i), R);
      const uint32_t D = add_mod(mfma_kappa(R), kP - b65537);
      const uint32_t t0 = beta + (D & 0xffffu), u0 = beta + (D >> 16);
      const int32_t cv[4] = {(int32_t)(t0 & 0xffu), (int32_t)(t0 >> 8), (int32_t)(u0 & 0xffu), (int32_t)(u0 >> 8)};
      for (int a = 0; a < classes; ++a) out->c[((size_t)a * 4 + g) * 4 + i] = cv[a];
    }
  out->fan.assign((size_t)(columns - 1) * 8, CoefRecord{});
  for (int c = 1; c < columns; ++c)
    for (int q = 0; q < 8; ++q) {
      const uint32_t* row = &rows[((size_t)c * column - 8 + q) * 8];
      if (!all_nonzero(row)) return false;
      out->fan[(size_t)(c - 1) * 8 + q] = make_record(row);
    }
  for (int j = 0; j < 16; ++j)
    if (!all_nonzero(&rows[(size_t)j * 8])) return false;
  return true;
}

// A^steps (row-major, rows giving the newest-first words) through the table
// of A^(2^b).
inline std::vector<uint32_t> matrix_power(const JumpTable& jt, uint64_t steps) {
  std::vector<uint32_t> m(64, 0);
  for (int r = 0; r < 8; ++r) m[(size_t)r * 9] = 1;
  for (int b = 0; b < 64; ++b)
    if ((steps >> b) & 1u) m = mat_mul(jt.at(b), m.data());
  return m;
}

// Stride tables for A^steps; false when a row has an entry == 0 (mod p),
// whose exact-zero case the residue arithmetic does not carry.
inline bool mfma_stride_tables(const JumpTable& jt, uint64_t steps, int classes, uint32_t beta,
                               MfmaStrideTables* out) {
  const std::vector<uint32_t> m = matrix_power(jt, steps);
  for (int r = 0; r < 8; ++r)
    if (!all_nonzero(&m[(size_t)r * 8])) return false;
  // new word q (oldest first) = newest-first word 7 - q: R_q[q'] = m[7 - q][7 - q']
  auto Rrow = [&](int q, uint32_t R[8]) {
    for (int k = 0; k < 8; ++k) R[k] = m[(size_t)(7 - q) * 8 + 7 - k];
  };
  out->a.assign((size_t)classes * 64 * 2, 0);
  for (int a = 0; a < classes; ++a)
    for (int l = 0; l < 64; ++l) {
      const int row = l & 15;
      if ((row & 3) >= 2) continue;  // rows 4 g + 2, 4 g + 3: zero
      uint32_t R[8];
      Rrow(kMfmaWordOf[row >> 2] + (row & 3), R);
      mfma_operand(R, l >> 4, a, &out->a[((size_t)a * 64 + l) * 2]);
    }
  out->k.assign(8, 0);
  for (int g = 0; g < 4; ++g)
    for (int i = 0; i < 2; ++i) {
      uint32_t R[8];
      Rrow(kMfmaWordOf[g] + i, R);
      const uint32_t k = add_mod(mfma_kappa(R), kP - mul_mod(mod_p(beta), 1u << 16));  // kappa - 2^16 beta
      out->k[(size_t)g * 2 + i] = k < beta ? k + kP : k;
    }
  return true;
}

}  // namespace mrg8
