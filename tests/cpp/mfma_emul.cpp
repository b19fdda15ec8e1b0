// CPU emulation of the matrix-core fill (k_fill_mfma) against the oracle.
//
// Builds the library's own MFMA tables (mrg8_amd/csrc/mrg8_math.hpp), then
// replays a wave's work lane by lane with the operand and result layouts of
// v_mfma_i32_16x16x32_i8 as the kernel uses them (lane l = n + 16 g holds
// K-group g of row/column n; result register i of lane l is row 4 g + i of
// column n): the first tile's column states through the fan-out records,
// each step from accumulators started at the step constants, the combine
// exactly as the kernel writes it in 32/64-bit unsigned arithmetic, the
// state hand-off between steps in registers 2 and 3, and the stride product
// that gives the column states of the wave's next tile.  Tiles 0, W and 2 W
// of a wave are compared with the oracle's serial generator
// (oracle/mrg8_oracle.c, original/mrg8.cpp's randint) for several strides W
// and states with zero and p words, and every partial sum is checked
// against the bound the kernel's arithmetic relies on.  Prints "mfma
// emulation ok" and exits 0 on success.
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "mrg8_math.hpp"

extern "C" {
#include "mrg8_oracle.h"
}

using namespace mrg8;

namespace {

constexpr int kColumns = 16, kClasses = 4, kRows = 16;
constexpr uint32_t kBeta = (1u << 27) + (1u << 19);

int failures = 0;

void check(bool ok, const char* what, long long a = 0, long long b = 0) {
  if (!ok && failures++ < 20) std::printf("FAIL %s (%lld, %lld)\n", what, a, b);
}

uint32_t digits(uint32_t w) { return (w + 0x40808080u) ^ 0x80808080u; }

int8_t byte_of(const uint32_t* v, int i) { return (int8_t)((v[i / 4] >> (8 * (i % 4))) & 0xffu); }

// dot_residue of mrg8_device.hip: V = B0 + sum c_j e_j (mod 2^64), then the fold.
uint32_t record_residue(const CoefRecord& k, const uint32_t s[8]) {
  uint64_t v = (uint64_t)k.b0_lo | ((uint64_t)k.b0_hi << 32);
  for (int j = 0; j < 8; ++j) v += (uint64_t)((int64_t)k.c[j] * (int64_t)((int32_t)(s[j] - kCenter)));
  const uint64_t w = 2 * (v >> 32) + (uint32_t)v;
  const uint32_t x = ((uint32_t)w & kP) + (uint32_t)(w >> 31);
  return x < kP ? x : x - kP;
}

// The step combine: accumulators that started from the step constants.
uint32_t combine_step(int32_t a0, int32_t a1, int32_t a2, int32_t a3) {
  const int64_t ts = (int64_t)a0 + 256 * (int64_t)a1, us = (int64_t)a2 + 256 * (int64_t)a3;
  check(ts >= 0 && ts < 2 * (int64_t)kBeta + 65536, "t range", ts);
  check(us >= 0 && us < 2 * (int64_t)kBeta + 32768, "u range", us);
  const uint32_t t = (uint32_t)a0 + ((uint32_t)a1 << 8);
  const uint32_t u = (uint32_t)a2 + ((uint32_t)a3 << 8);
  const uint64_t xs = (uint64_t)t + ((u << 16) & kP) + (u >> 15);
  check(xs < 2ull * kP, "x range", (long long)xs);
  const uint32_t x = t + ((u << 16) & kP) + (u >> 15);
  return x < x - kP ? x : x - kP;  // min(x, x - p) in 32 bits
}

// The stride combine: accumulators from 0, the constant K in [beta, beta + p).
uint32_t combine_stride(int32_t a0, int32_t a1, int32_t a2, int32_t a3, uint32_t k) {
  const int64_t ts = (int64_t)a0 + 256 * (int64_t)a1, us = (int64_t)a2 + 256 * (int64_t)a3;
  check(ts >= -(int64_t)kBeta && ts <= (int64_t)kBeta, "stride t range", ts);
  check(us >= -(int64_t)kBeta && us <= (int64_t)kBeta, "stride u range", us);
  check(k >= kBeta && (uint64_t)k < (uint64_t)kBeta + kP, "K range", k);
  const uint32_t t = (uint32_t)a0 + ((uint32_t)a1 << 8);
  const uint32_t u = (uint32_t)a2 + ((uint32_t)a3 << 8) + kBeta;
  const uint64_t y = (uint64_t)u * 65536u + (uint64_t)(uint32_t)(t + k);
  check(y < (1ull << 45), "y range", (long long)y);
  const uint32_t x = ((uint32_t)y & kP) + (uint32_t)(y >> 31);
  return x < x - kP ? x : x - kP;
}

// C[a][m][n] += sum over K-groups g and bytes i of A_a[m + 16 g][i] * B[n + 16 g][i]
void mfma(const std::vector<uint32_t>& A, const uint32_t B[64][2], int32_t C[kClasses][kRows][16]) {
  for (int a = 0; a < kClasses; ++a)
    for (int m = 0; m < kRows; ++m)
      for (int n = 0; n < 16; ++n) {
        int32_t acc = 0;
        for (int g = 0; g < 4; ++g) {
          const uint32_t* av = &A[((size_t)a * 64 + m + 16 * g) * 2];
          for (int i = 0; i < 8; ++i) acc += (int32_t)byte_of(av, i) * (int32_t)byte_of(B[n + 16 * g], i);
        }
        check(acc >= -(1 << 19) && acc <= (1 << 19), "acc range", acc);
        C[a][m][n] += acc;
      }
}

// A wave's first tile: column operands from the newest-first tile state s
// through the fan-out records (lane n + 16 g: words W(g), W(g) + 1 of column n).
void fan_out(const MfmaHostTables& mt, const uint32_t s[8], uint32_t B[64][2]) {
  for (int l = 0; l < 64; ++l) {
    const int c = l & 15, g = l >> 4;
    for (int v = 0; v < 2; ++v) {
      const int q = kMfmaWordOf[g] + v;
      const uint32_t w = c == 0 ? s[7 - q] : record_residue(mt.fan[(size_t)(c - 1) * 8 + q], s) + 1u;
      B[l][v] = digits(w);
    }
  }
}

// One tile from the column operands B (16 column words to out); B becomes
// the operands of the tile `stride` tiles later.
void emulate_tile(const MfmaHostTables& mt, const MfmaStrideTables& st, int column, uint32_t B[64][2], uint32_t* out) {
  uint32_t B0[64][2];
  std::memcpy(B0, B, sizeof B0);
  static int32_t C[kClasses][kRows][16];
  for (int step = 0; step < column / 16; ++step) {
    for (int a = 0; a < kClasses; ++a)
      for (int l = 0; l < 64; ++l)
        for (int i = 0; i < 4; ++i) C[a][4 * (l >> 4) + i][l & 15] = mt.c[((size_t)a * 4 + (l >> 4)) * 4 + i];
    mfma(mt.a, B, C);
    for (int l = 0; l < 64; ++l) {
      const int c = l & 15, g = l >> 4;
      uint32_t R[4];
      for (int i = 0; i < 4; ++i) {
        const int m = 4 * g + i;
        R[i] = combine_step(C[0][m][c], C[1][m][c], C[2][m][c], C[3][m][c]);
        out[c * column + 16 * step + mfma_position(g, i)] = R[i] + 1u;
      }
      for (int v = 0; v < 2; ++v) B[l][v] = digits(R[2 + v] + 1u);
    }
  }
  std::memset(C, 0, sizeof C);
  mfma(st.a, B0, C);
  for (int l = 0; l < 64; ++l) {
    const int c = l & 15, g = l >> 4;
    for (int i = 0; i < 2; ++i) {
      const int m = 4 * g + i;  // register i < 2 of lane group g: new word W(g) + i
      B[l][i] = digits(combine_stride(C[0][m][c], C[1][m][c], C[2][m][c], C[3][m][c], st.k[(size_t)g * 2 + i]) + 1u);
    }
  }
}

uint64_t rng = 0x9e3779b97f4a7c15ull;
uint32_t next_word() {
  rng ^= rng << 13;
  rng ^= rng >> 7;
  rng ^= rng << 17;
  return (uint32_t)(rng % ((uint64_t)kP + 1));  // [0, p]
}

}  // namespace

// Checks the tables and tiles of one column length; returns the tiles compared.
int check_column(int column, const std::vector<std::vector<uint32_t>>& states) {
  MfmaHostTables mt;
  if (!mfma_host_tables(kColumns, column, kClasses, kBeta, &mt)) {
    check(false, "table build (a fan-out or block row has a zero entry)", column);
    return 0;
  }
  // The operand digits reassemble the centred, shifted rows.
  for (int l = 0; l < 64; ++l)
    for (int i = 0; i < 8; ++i) {
      int64_t m = 0;
      for (int a = kClasses - 1; a >= 0; --a) m = m * 256 + byte_of(&mt.a[((size_t)a * 64 + l) * 2], i);
      check(m >= -(int64_t)(kP / 2) && m <= (int64_t)(kP / 2), "digit value range", m);
    }
  std::vector<uint32_t> got((size_t)kColumns * column), want((size_t)kColumns * column);
  std::vector<uint32_t> tab((size_t)247 * 64);
  mrg8o_jump_table(tab.data());
  const JumpTable jt;
  int tiles = 0;
  for (uint64_t stride : {1ull, 7ull, 2048ull, 65535ull}) {
    MfmaStrideTables st;
    check(mfma_stride_tables(jt, stride * kColumns * column, kClasses, kBeta, &st), "stride rows", (long long)stride);
    for (size_t si = 0; si < states.size(); ++si) {
      uint32_t so[8], B[64][2];
      std::memcpy(so, states[si].data(), sizeof so);
      fan_out(mt, states[si].data(), B);
      for (int t = 0; t < 3; ++t, ++tiles) {  // tiles 0, stride, 2 stride of a wave
        emulate_tile(mt, st, column, B, got.data());
        uint32_t sw[8];
        std::memcpy(sw, so, sizeof sw);
        mrg8o_randint(sw, want.data(), want.size());
        int bad = 0;
        for (size_t i = 0; i < got.size(); ++i)
          if (got[i] != want[i] && bad++ < 3)
            std::printf("FAIL column %d stride %llu state %zu tile %d output %zu: %u != %u\n", column,
                        (unsigned long long)stride, si, t, i, got[i], want[i]);
        failures += bad;
        mrg8o_jump(tab.data(), stride * kColumns * column, so);
      }
    }
  }
  return tiles;
}

int main() {
  std::vector<std::vector<uint32_t>> states;
  uint32_t st[8];
  mrg8o_seed(13579, st);
  states.emplace_back(st, st + 8);
  for (int t = 0; t < 6; ++t) {
    std::vector<uint32_t> v(8);
    for (auto& x : v) x = next_word();
    states.push_back(v);
  }
  states.push_back({kP, kP, kP, kP, kP, kP, kP, kP});  // every word == 0 (mod p)
  states.push_back({0, 0, 0, 0, 0, 0, 0, 1});          // one non-zero word
  states.push_back({kP, 0, 1, kP - 1, 0, kP, 2, 0});   // zeros, p and p - 1
  states.push_back({1, 1, 1, 1, 1, 1, 1, 1});
  int tiles = 0;
  for (int column : {64, 128}) tiles += check_column(column, states);
  if (failures) {
    std::printf("mfma emulation: %d failures\n", failures);
    return 1;
  }
  std::printf("mfma emulation ok: %d tiles (columns of 64 and 128 outputs)\n", tiles);
  return 0;
}
