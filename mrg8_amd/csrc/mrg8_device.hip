// gfx950 kernels for MRG8 generation.
//
// Replaces the CUDA generation kernels rng_outer_32 and rng_inner
// (cuda/mrg8_cuda.cu:201-399).  Design:
//   * A wave owns a contiguous run (tile) of outputs of one stream.  The
//     8-word state is wave-uniform (scalar registers); lane k holds the
//     centred coefficients of row 0 of A^(PL k + i + 1), i < PL, so one
//     iteration yields 64 * PL consecutive outputs with eight signed
//     32x32->64 multiply-adds each (v_mad_i64_i32, the PL dot products
//     interleaved) and a Mersenne fold.
//   * The next state is the last eight outputs, read back with v_readlane;
//     no LDS and no barrier in the loop.
//   * Outputs leave through coalesced non-temporal vector stores (16 bytes
//     per lane, 1 KiB per wave store); a run first writes the outputs before
//     its first 1 KiB boundary one per lane.
//   * A wave reaches its first position with a jump: each non-zero byte d
//     at position k of the 64-bit offset is one product with A^(d 256^k)
//     (at most 8 steps).  Eight lane groups of eight lanes jump eight states
//     at once, lane r of a group computing row r, and the next digit's
//     record is fetched one step ahead.
//   * k_fill: one stream, one tile per wave; the block's start states are
//     jumped by its first waves and shared through LDS.
//   * k_batch: many streams (segments), a persistent grid striding over
//     tiles, tables and counts read from one device block.
//   * Blocks reach their part of the output through xcd_block: the
//     dispatcher deals blocks to the XCDs in turn, and large grids give each
//     XCD runs of consecutive blocks' output instead.
#include "mrg8_device.hpp"

namespace mrg8 {
namespace {

constexpr double kRnorm = 1.0 / 2147483647.0;  // original/mrg8.cpp:233

typedef double d2_t __attribute__((ext_vector_type(2)));
typedef uint32_t u2_t __attribute__((ext_vector_type(2)));
typedef uint32_t u4_t __attribute__((ext_vector_type(4)));

struct Coef {
  int32_t c[8];
  uint64_t b0;
  uint32_t support;
};

__device__ __forceinline__ Coef load_coef(const CoefRecord* __restrict__ p) {
  const u4_t* q = reinterpret_cast<const u4_t*>(p);
  u4_t a = q[0], b = q[1], c = q[2];
  Coef k;
  k.c[0] = (int32_t)a.x; k.c[1] = (int32_t)a.y; k.c[2] = (int32_t)a.z; k.c[3] = (int32_t)a.w;
  k.c[4] = (int32_t)b.x; k.c[5] = (int32_t)b.y; k.c[6] = (int32_t)b.z; k.c[7] = (int32_t)b.w;
  k.b0 = (uint64_t)c.x | ((uint64_t)c.y << 32);
  k.support = c.z;
  return k;
}

// The 11 used words of a record, without its pad word: a load into a register
// nothing reads lets the compiler reuse that register while the load is in
// flight, which forces a wait for the load where the register is reused --
// in a prefetching loop, before the step the prefetch should overlap.
__device__ __forceinline__ Coef load_coef11(const CoefRecord* __restrict__ p) {
  const u4_t* q = reinterpret_cast<const u4_t*>(p);
  const u4_t a = q[0], b = q[1];
  const u2_t c = *reinterpret_cast<const u2_t*>(q + 2);
  const uint32_t support = reinterpret_cast<const uint32_t*>(q + 2)[2];
  Coef k;
  k.c[0] = (int32_t)a.x; k.c[1] = (int32_t)a.y; k.c[2] = (int32_t)a.z; k.c[3] = (int32_t)a.w;
  k.c[4] = (int32_t)b.x; k.c[5] = (int32_t)b.y; k.c[6] = (int32_t)b.z; k.c[7] = (int32_t)b.w;
  k.b0 = (uint64_t)c.x | ((uint64_t)c.y << 32);
  k.support = support;
  return k;
}

// w = 2 * hi + lo as one 32x32->64 multiply-add (2^32 == 2 mod p).
__device__ __forceinline__ uint64_t fold_hi(uint32_t hi, uint32_t lo) {
  uint64_t w, carry;
  asm("v_mad_u64_u32 %0, %1, %2, 2, %3" : "=v"(w), "=s"(carry) : "v"(hi), "v"((uint64_t)lo));
  return w;
}

// Canonical residue r in [0, p) of V = B0 + sum_j c_j e_j, where e_j = s_j - 2^30;
// the reference's word for the dot product is r + 1 (see CoefRecord).
__device__ __forceinline__ uint32_t dot_residue(const Coef& k, const int32_t e[8]) {
  uint64_t v = k.b0;
#pragma unroll
  for (int j = 0; j < 8; ++j) v += (uint64_t)((int64_t)k.c[j] * (int64_t)e[j]);
  const uint64_t w = fold_hi((uint32_t)(v >> 32), (uint32_t)v);        // < 3 * 2^32
  const uint32_t x = ((uint32_t)w & kP) + (uint32_t)(w >> 31);          // < 2^31 + 6
  return min(x, x - kP);  // x - p wraps above x when x < p
}

// PL residues at once, the eight multiply-adds of the PL dot products
// interleaved so that consecutive 64-bit multiply-adds are independent and
// issue back to back.
template <int PL>
__device__ __forceinline__ void dot_residues(const Coef (&k)[PL], const int32_t e[8], uint32_t r[PL]) {
  uint64_t v[PL];
#pragma unroll
  for (int i = 0; i < PL; ++i) v[i] = k[i].b0;
#pragma unroll
  for (int j = 0; j < 8; ++j)
#pragma unroll
    for (int i = 0; i < PL; ++i) v[i] += (uint64_t)((int64_t)k[i].c[j] * (int64_t)e[j]);
#pragma unroll
  for (int i = 0; i < PL; ++i) {
    const uint64_t w = fold_hi((uint32_t)(v[i] >> 32), (uint32_t)v[i]);
    const uint32_t x = ((uint32_t)w & kP) + (uint32_t)(w >> 31);
    r[i] = min(x, x - kP);
  }
}

// The value of v in each of the eight lanes of this lane's 8-lane group,
// w[j] from lane (lane & ~7) | j: ds_swizzle in bit mode (and-mask 0x18 keeps
// the group bits within the 32-lane half, or-mask j picks the lane), so no
// lane-address registers are needed.
template <int J>
__device__ __forceinline__ uint32_t group_lane(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x18 | (J << 5));
}

__device__ __forceinline__ void gather_group(uint32_t v, uint32_t w[8]) {
  w[0] = group_lane<0>(v);
  w[1] = group_lane<1>(v);
  w[2] = group_lane<2>(v);
  w[3] = group_lane<3>(v);
  w[4] = group_lane<4>(v);
  w[5] = group_lane<5>(v);
  w[6] = group_lane<6>(v);
  w[7] = group_lane<7>(v);
}

__device__ __forceinline__ void centre(const uint32_t s[8], int32_t e[8]) {
#pragma unroll
  for (int j = 0; j < 8; ++j) e[j] = (int32_t)(s[j] - kCenter);
}

// Advances eight states at once: lane group g = lane / 8 holds its own state
// (replicated in the group's lanes, s[] per lane) and its own 64-bit offset;
// for the group's lowest non-zero digit d (kJumpBits bits) at position k,
// lane r of the group computes row r of A^(d 2^(kJumpBits k)), and the
// group's eight new words are gathered back with ds_swizzle.  Groups with no
// digits left idle until every group is done; a 64-bit jump takes at most
// kJumpPositions steps.
__device__ __forceinline__ const CoefRecord* digit_row(const CoefRecord* __restrict__ jrec, uint64_t off, int row) {
  const int k = __builtin_ctzll(off) / kJumpBits;
  const int d = (int)((off >> (kJumpBits * k)) & (uint64_t)kJumpValues);
  return jrec + ((k * kJumpValues + (d - 1)) * 8 + row);
}

__device__ __forceinline__ uint64_t drop_digit(uint64_t off) {
  return off & ~((uint64_t)kJumpValues << (__builtin_ctzll(off) / kJumpBits * kJumpBits));
}

__device__ __forceinline__ void jump_groups(uint32_t s[8], uint64_t off, const CoefRecord* __restrict__ jrec) {
  const int row = threadIdx.x & 7;
  Coef k = {};
  if (off) k = load_coef(digit_row(jrec, off, row));
  while (__ballot(off != 0)) {
    const bool active = off != 0;
    const uint64_t rest = active ? drop_digit(off) : 0;
    Coef next = {};
    if (rest) next = load_coef(digit_row(jrec, rest, row));  // one step ahead
    int32_t e[8];
    uint32_t nz = 0;
    centre(s, e);
#pragma unroll
    for (int j = 0; j < 8; ++j) nz |= (s[j] != 0u ? 1u : 0u) << j;
    const uint32_t y = (k.support & nz) ? dot_residue(k, e) + 1u : 0u;
    uint32_t w[8];
    gather_group(y, w);
#pragma unroll
    for (int j = 0; j < 8; ++j) s[j] = active ? w[j] : s[j];
    off = rest;
    k = next;
  }
}

// The same jump, pipelined, for the persistent batched kernel, whose jumps
// are on the critical path when segments are short and offsets long.  The
// offset's digits are taken position by position, up to the highest position
// where some group has a digit, skipping positions where every group's digit
// is 0: at position k, lane r of each group computes row r of
// A^(d 2^(kJumpBits k)) for its group's digit d (row r of the identity when d
// is 0).  The next position's record is fetched while the current step
// computes (an unconditional load, so that the compiler's wait for it is not
// hoisted above the step), and the zero-product test runs only when the
// state has a zero word.  65536 random 63-bit jumps + 256 words each: 30 us
// against 34 us with jump_groups.  The one-tile-per-wave kernels keep
// jump_groups: with this version their long fills ran 2 % slower (and 3 %
// slower still with no jump at all, so a block's jump phase, which delays its
// first stores, does not cost those kernels time).
__device__ __forceinline__ uint32_t digit_at(uint64_t off, int k) {
  return (uint32_t)(off >> (kJumpBits * k)) & (uint32_t)kJumpValues;
}

__device__ __forceinline__ const CoefRecord* digit_rec(const CoefRecord* __restrict__ jrec, uint64_t off, int k,
                                                       int row) {
  const int d = (int)digit_at(off, k);
  return jrec + (d ? (k * kJumpValues + d - 1) * 8 + row : kIdentityRec + row);
}

// One step of the pipelined jumps below: lane r of group g computes row r of
// the group's matrix (record c) times the group's state, and the group's
// eight new words are gathered back.
template <bool ZERO_WORDS = true>
__device__ __forceinline__ void jump_step(uint32_t s[8], const Coef& c) {
  int32_t e[8];
  centre(s, e);
  uint32_t y = dot_residue(c, e) + 1u;
  // bigDotProd is 0 when every product is 0; with all eight words non-zero
  // that cannot happen (no row of a power of A is zero)
  uint32_t smin = s[0];
#pragma unroll
  for (int j = 1; j < 8; ++j) smin = min(smin, s[j]);
  if (ZERO_WORDS && smin == 0u) {
    uint32_t nz = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) nz |= (s[j] != 0u ? 1u : 0u) << j;
    if (!(c.support & nz)) y = 0u;
  }
  gather_group(y, s);
}

__device__ __forceinline__ void jump_groups_pipelined(uint32_t s[8], uint64_t off,
                                                      const CoefRecord* __restrict__ jrec) {
  const int row = threadIdx.x & 7;
  if (!__ballot(off != 0)) return;
  Coef c = load_coef11(digit_rec(jrec, off, 0, row));
  for (int k = 0;; ++k) {
    // wave-uniform: no group has a digit above position k
    const bool last = k + 1 == kJumpPositions || !__ballot((off >> (kJumpBits * (k + 1))) != 0);
    const Coef next = load_coef11(digit_rec(jrec, last ? 0 : off, last ? 0 : k + 1, row));
    if (__ballot(digit_at(off, k) != 0)) jump_step(s, c);
    if (last) break;
    c = next;
  }
}

template <int FMT>
struct Out;
template <>
struct Out<0> { typedef double T; };
template <>
struct Out<1> { typedef double T; };
template <>
struct Out<2> { typedef uint32_t T; };

// Output value for residue r, the reference's word being x = r + 1.
template <int FMT>
__device__ __forceinline__ typename Out<FMT>::T convert(uint32_t r) {
  if constexpr (FMT == 0) return __fma_rn((double)r, kRnorm, kRnorm);  // double(x) * (1/p), one rounding: original/mrg8.cpp:236
  else if constexpr (FMT == 1) return (double)r * kRnorm;              // double(x - 1) * (1/p): cuda/mrg8_cuda.cu:363
  else return r + 1u;                                                  // x: original/mrg8.cpp:219
}

// Output value of the all-zero state (whose words stay 0).
template <int FMT>
__device__ __forceinline__ typename Out<FMT>::T convert_zero() {
  if constexpr (FMT == 0) return 0.0;
  else if constexpr (FMT == 1) return (double)0xffffffffu * kRnorm;  // (x - 1) wraps in 32 bits
  else return 0u;
}

// Which outputs of an iteration (64 * PL consecutive outputs) a lane makes.
// Doubles with PL = 4 use two halves: lane k makes outputs 2k, 2k+1 and
// 128+2k, 128+2k+1, so each of the lane's two 16-byte stores belongs to a
// fully contiguous 1 KiB store of the wave.  Otherwise lane k makes outputs
// PL k .. PL k + PL - 1 (one store of 4, 8 or 16 bytes).
template <int FMT, int PL>
struct LaneLayout {
  static constexpr bool kSplit = FMT != 2 && PL == 4;
  static __device__ __forceinline__ uint32_t index(uint32_t lane, int i) {
    if constexpr (kSplit) return 128u * (uint32_t)(i / 2) + 2u * lane + (uint32_t)(i % 2);
    else return PL * lane + (uint32_t)i;
  }
  // the slot (0..PL-1) and lane that make output o of an iteration
  static constexpr int slot_of(int o) { return kSplit ? 2 * (o / 128) + o % 2 : o % PL; }
  static constexpr int lane_of(int o) { return kSplit ? (o % 128) / 2 : o / PL; }
};

// Non-temporal stores of one lane's PL outputs of the iteration at `base`.
template <int FMT, int PL>
__device__ __forceinline__ void store_lane(typename Out<FMT>::T* base, uint32_t lane, const uint32_t r[PL]) {
  if constexpr (FMT == 2 && PL == 4) {
    u4_t v = {r[0] + 1u, r[1] + 1u, r[2] + 1u, r[3] + 1u};
    __builtin_nontemporal_store(v, reinterpret_cast<u4_t*>(base + 4 * lane));
  } else if constexpr (FMT == 2 && PL == 2) {
    u2_t v = {r[0] + 1u, r[1] + 1u};
    __builtin_nontemporal_store(v, reinterpret_cast<u2_t*>(base + 2 * lane));
  } else if constexpr (FMT != 2 && PL == 4) {
    d2_t v0 = {convert<FMT>(r[0]), convert<FMT>(r[1])};
    d2_t v1 = {convert<FMT>(r[2]), convert<FMT>(r[3])};
    __builtin_nontemporal_store(v0, reinterpret_cast<d2_t*>(base + 2 * lane));
    __builtin_nontemporal_store(v1, reinterpret_cast<d2_t*>(base + 128 + 2 * lane));
  } else if constexpr (FMT != 2 && PL == 2) {
    d2_t v = {convert<FMT>(r[0]), convert<FMT>(r[1])};
    __builtin_nontemporal_store(v, reinterpret_cast<d2_t*>(base + 2 * lane));
  } else {
#pragma unroll
    for (int i = 0; i < PL; ++i) __builtin_nontemporal_store(convert<FMT>(r[i]), base + PL * lane + i);
  }
}

// This lane's PL records: row 0 of A^(o + 1) for each output o it makes.
template <int FMT, int PL>
__device__ __forceinline__ void load_lane_coefs(Coef (&k)[PL], const CoefRecord* __restrict__ lane_rec) {
  const uint32_t lane = threadIdx.x & 63;
#pragma unroll
  for (int i = 0; i < PL; ++i) k[i] = load_coef(lane_rec + LaneLayout<FMT, PL>::index(lane, i));
}

// Writes the first k (< 64) outputs of a run one per lane and advances the
// wave-uniform state past them.
template <int FMT>
__device__ __forceinline__ void peel_outputs(typename Out<FMT>::T* __restrict__ dst, uint32_t k, uint32_t s[8],
                                             const CoefRecord* __restrict__ lane_rec) {
  const uint32_t lane = threadIdx.x & 63;
  uint32_t any = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) any |= s[j];
  if (any == 0u) {  // the all-zero state stays zero
    if (lane < k) __builtin_nontemporal_store(convert_zero<FMT>(), dst + lane);
    return;
  }
  const Coef c = load_coef(lane_rec + lane);  // row 0 of A^(lane+1)
  int32_t e[8];
  centre(s, e);
  const uint32_t r = dot_residue(c, e);
  if (lane < k) __builtin_nontemporal_store(convert<FMT>(r), dst + lane);
  // new state: the k new words (newest first) followed by the 8-k oldest kept
  uint32_t ns[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const uint32_t src = (uint32_t)j < k ? k - 1 - (uint32_t)j : 0u;
    const uint32_t fresh = __builtin_amdgcn_readlane(r, src) + 1u;
    uint32_t kept = s[0];
#pragma unroll
    for (int q = 1; q < 8; ++q) kept = (uint32_t)(j - q) == k ? s[q] : kept;  // s[j - k]
    kept = (uint32_t)j == k ? s[0] : kept;
    ns[j] = (uint32_t)j < k ? fresh : kept;
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) s[j] = ns[j];
}

// Byte alignment of the wide-store loop's wave stores: 64 lanes x 16 bytes.
constexpr uintptr_t kStoreAlign = 1024;

// Writes the outputs of a run before its first ALIGN-byte boundary (at most
// 63 per round, one per lane) and advances dst, cnt and the state past them.
// With ALIGN = kStoreAlign every 1 KiB wave store of the loop that follows
// covers whole cache lines: on MI355X a 2^28-double fill whose destination
// was 8 or 16 bytes off such a boundary took 471-473 us with a 16-byte peel,
// 346-350 us with this one (aligned: 342 us).
template <int FMT, uintptr_t ALIGN>
__device__ __forceinline__ void peel_to_boundary(typename Out<FMT>::T*& dst, uint64_t& cnt, uint32_t s[8],
                                                 const CoefRecord* __restrict__ lane_rec) {
  typedef typename Out<FMT>::T T;
  uint64_t b = ((ALIGN - (reinterpret_cast<uintptr_t>(dst) & (ALIGN - 1))) & (ALIGN - 1)) / sizeof(T);
  if (b > cnt) b = cnt;
  while (b) {
    const uint32_t k = b < 63 ? (uint32_t)b : 63u;
    peel_outputs<FMT>(dst, k, s, lane_rec);
    dst += k;
    cnt -= k;
    b -= k;
  }
}

// Generates `cnt` outputs of the stream whose wave-uniform state (at the
// first output) is s, into dst[0..cnt), with this lane's records k (and the
// lane records, for the outputs before the first 1 KiB boundary).  `cnt` and
// s must be wave-uniform; dst need only be element-aligned.
template <int FMT, int PL, uintptr_t ALIGN = kStoreAlign>
__device__ __forceinline__ void generate_run(typename Out<FMT>::T* __restrict__ dst, uint64_t cnt, uint32_t s[8],
                                             const Coef (&k)[PL], const CoefRecord* __restrict__ lane_rec) {
  typedef LaneLayout<FMT, PL> L;
  constexpr int kPer = 64 * PL;
  const uint32_t lane = threadIdx.x & 63;
  uint32_t any = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) any |= s[j];
  if (any == 0u) {  // degenerate all-zero state
    for (uint64_t i = lane; i < cnt; i += 64) __builtin_nontemporal_store(convert_zero<FMT>(), dst + i);
    return;
  }
  typename Out<FMT>::T* base = dst;
  peel_to_boundary<FMT, ALIGN>(base, cnt, s, lane_rec);
  const uint64_t full = cnt / kPer;
  const uint64_t rem = cnt - full * kPer;
  for (uint64_t it = 0; it < full; ++it) {
    int32_t e[8];
    centre(s, e);
    uint32_t r[PL];
    dot_residues<PL>(k, e, r);
    store_lane<FMT, PL>(base, lane, r);
    base += kPer;
    // next state: the last eight words, newest first (not needed after the
    // run's last iteration: 8 of 16 readlanes of a 256-output tile at 2 per lane)
    if (it + 1 < full || rem) {
#pragma unroll
      for (int j = 0; j < 8; ++j)
        s[j] = __builtin_amdgcn_readlane(r[L::slot_of(kPer - 1 - j)], L::lane_of(kPer - 1 - j)) + 1u;
    }
  }
  if (rem) {
    int32_t e[8];
    centre(s, e);
    uint32_t r[PL];
    dot_residues<PL>(k, e, r);
#pragma unroll
    for (int i = 0; i < PL; ++i) {
      const uint32_t o = L::index(lane, i);
      if (o < rem) __builtin_nontemporal_store(convert<FMT>(r[i]), base + o);
    }
  }
}

// First output of wave w of a fill: wave 0 also takes the head (the outputs
// before the first tile-aligned address), every later wave starts on one.
__device__ __forceinline__ uint64_t fill_first(const FillArgs& a, uint64_t w) { return w ? a.head + w * a.tile : 0; }

// The part of the output that block b of an nb-block grid writes.  The
// dispatcher deals blocks to the XCDs in turn (block b to XCD (b + r) mod x,
// r where its turn stood after earlier launches), so in dispatch order the
// XCDs interleave block by block.  With group g > 1, the blocks of one XCD
// (b mod x alike) instead take runs of g consecutive blocks' output within
// every span of x g blocks, whatever r is.  Blocks past the last whole span
// keep their own part.  Either way every part is written by exactly one
// block.
__device__ __forceinline__ uint64_t xcd_block(uint32_t b, uint32_t nb, uint32_t x, uint32_t g) {
  if (g <= 1 || x == 0) return b;
  const uint32_t span = x * g, whole = nb / span * span;
  if (b >= whole) return b;
  const uint32_t q = b / x;
  return (uint64_t)(q / g * span + (b - q * x) * g + q % g);
}

template <int FMT, int PL>
__global__ __launch_bounds__(1024) void k_fill(FillArgs a) {
  // Start states of the block's waves: wave j (j < ceil(nw / 8)) jumps the
  // states of waves 8j..8j+7 together (one 8-lane group each) and leaves
  // them in LDS; every wave then reads its own.
  __shared__ uint32_t start_state[16][8];
  const int lane = threadIdx.x & 63;
  const uint32_t wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform wave index
  const uint32_t nw = blockDim.x >> 6;
  const uint64_t wave0 = xcd_block(blockIdx.x, gridDim.x, a.xcds, a.xcd_group) * nw;
  if (wib * 8 < nw) {
    const uint32_t target = wib * 8 + (lane >> 3);
    uint32_t s[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) s[j] = a.start.state[j];
    jump_groups(s, target < nw ? a.start.offset + fill_first(a, wave0 + target) : 0, a.jump);
    if (target < nw && (lane & 7) == 0) {
#pragma unroll
      for (int j = 0; j < 8; ++j) start_state[target][j] = s[j];
    }
  }
  __syncthreads();
  const uint64_t w = wave0 + wib;
  const uint64_t first = fill_first(a, w);
  if (first >= a.n) return;
  const uint64_t cnt = min(a.head + (w + 1) * a.tile, a.n) - first;
  uint32_t s[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) s[j] = __builtin_amdgcn_readfirstlane(start_state[wib][j]);
  Coef k[PL];
  load_lane_coefs<FMT, PL>(k, a.lane);
  generate_run<FMT, PL>(static_cast<typename Out<FMT>::T*>(a.out) + first, cnt, s, k, a.lane);
}

// One tile of a batch segment at dst (generate_run peels the outputs before
// the first ALIGN-byte boundary; only tile 0 of a segment can have any).  The
// persistent kernel passes ALIGN = 16: with the 1 KiB peel it ran 10 % slower
// on the 4096-stream batch.
template <int FMT, int PL, uintptr_t ALIGN = kStoreAlign>
__device__ __forceinline__ void batch_tile(void* out, uint64_t cnt, uint32_t s[8], const Coef (&k)[PL],
                                           const CoefRecord* __restrict__ lane_rec) {
  if (cnt) generate_run<FMT, PL, ALIGN>(static_cast<typename Out<FMT>::T*>(out), cnt, s, k, lane_rec);
}

__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, int lane) {
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, lane);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), lane);
  return (uint64_t)lo | ((uint64_t)hi << 32);
}

// The tables of a batch block (see BatchHeader).
struct BatchTables {
  const BatchSegment* segs;
  const uint64_t* tile_begin;
  const uint32_t* tile_seg;
  uint64_t ntiles;
};

__device__ __forceinline__ BatchTables batch_tables(const uint8_t* block) {
  const BatchHeader* hdr = reinterpret_cast<const BatchHeader*>(block);
  BatchTables t;
  t.segs = reinterpret_cast<const BatchSegment*>(block + sizeof(BatchHeader));
  t.tile_begin = reinterpret_cast<const uint64_t*>(block + hdr->tile_begin_off);
  t.tile_seg = reinterpret_cast<const uint32_t*>(block + hdr->tile_seg_off);
  t.ntiles = hdr->ntiles;
  return t;
}

// Destination, output count, stream position and state source of tile t.
struct TileInfo {
  uint64_t dst;
  uint64_t cnt;
  uint64_t offset;
  const uint32_t* state;
  int32_t fmt;
};

__device__ __forceinline__ TileInfo lookup_tile(const BatchTables& tab, const BatchArgs& a, uint64_t t) {
  const uint32_t seg = tab.tile_seg[t];
  const BatchSegment* sp = tab.segs + seg;
  TileInfo ti;
  ti.fmt = sp->format;
  const uint64_t kt = t - tab.tile_begin[seg];  // tile number within the segment
  uint64_t first, end;
  tile_bounds(segment_head(sp->out, ti.fmt, sp->count, a.tile), sp->count, a.tile, kt, &first, &end);
  ti.cnt = end - first;
  ti.dst = reinterpret_cast<uint64_t>(sp->out) + first * (ti.fmt == 2 ? 4u : 8u);
  ti.offset = sp->offset + first;
  const uint32_t sid = sp->stream;
  ti.state = sid == kInlineState ? sp->state : a.states + 8ull * sid;
  return ti;
}

// Persistent batched kernel.  With W waves in the grid, wave w takes tiles
// w, w + W, w + 2W, ...: at any moment the grid writes a window of about W
// consecutive tiles, as a one-tile-per-wave launch would, and equal tiles
// keep the waves in step without any atomic.  A wave looks up its next
// kGroup tiles together (one 8-lane group each: segment, destination, count,
// start state, which the group then jumps) and generates them one after
// another, so a tile starts storing without waiting on memory.
constexpr uint32_t kGroup = 8;

__global__ __launch_bounds__(1024) void k_batch(BatchArgs a) {
  const int lane = threadIdx.x & 63;
  const int grp = lane >> 3;
  const BatchTables tab = batch_tables(a.block);
  const uint64_t ntiles = tab.ntiles;
  const uint64_t nwaves = (uint64_t)gridDim.x * (blockDim.x >> 6);
  const uint64_t wave =
      xcd_block(blockIdx.x, gridDim.x, a.xcds, a.xcd_group) * (blockDim.x >> 6) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (wave >= ntiles) return;
  const uint64_t mine_total = (ntiles - wave + nwaves - 1) / nwaves;
  for (uint64_t j0 = 0; j0 < mine_total; j0 += kGroup) {
    const uint32_t ngrp = (uint32_t)min<uint64_t>(kGroup, mine_total - j0);
    const bool mine = (uint32_t)grp < ngrp;
    const TileInfo ti = lookup_tile(tab, a, (j0 + (mine ? grp : 0)) * nwaves + wave);
    uint32_t s[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) s[j] = ti.state[j];
    jump_groups_pipelined(s, mine ? ti.offset : 0, a.jump);
    // loaded after the jumps so that they hold no registers during them
    Coef k[2];
    load_lane_coefs<0, 2>(k, a.lane);  // the PL = 2 layout is the same for every format
    for (uint32_t g = 0; g < ngrp; ++g) {
      uint32_t u[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) u[j] = __builtin_amdgcn_readlane(s[j], 8 * g);
      void* out = reinterpret_cast<void*>(readlane_u64(ti.dst, 8 * g));
      const uint64_t ucnt = readlane_u64(ti.cnt, 8 * g);
      switch (__builtin_amdgcn_readlane(ti.fmt, 8 * g)) {
        case 0: batch_tile<0, 2, 16>(out, ucnt, u, k, a.lane); break;
        case 1: batch_tile<1, 2, 16>(out, ucnt, u, k, a.lane); break;
        default: batch_tile<2, 2, 16>(out, ucnt, u, k, a.lane); break;
      }
    }
  }
}

// One tile of a batch with the fill's layout (4 outputs per lane).
template <int FMT>
__device__ __forceinline__ void batch_tile_alone(void* out, uint64_t cnt, uint32_t s[8],
                                                 const CoefRecord* __restrict__ lane_rec) {
  Coef k[4];
  load_lane_coefs<FMT, 4>(k, lane_rec);
  batch_tile<FMT, 4>(out, cnt, s, k, lane_rec);
}

// Batched kernel with one tile per wave (grid = tiles / waves per block), the
// structure of k_fill: waves 0-1 of a block look up and jump the start states
// of the block's tiles (8-lane groups) and leave them in LDS; every wave then
// generates its own tile and exits.
__global__ __launch_bounds__(1024) void k_batch_tiles(BatchArgs a) {
  __shared__ uint32_t start_state[16][8];
  __shared__ uint64_t start_dst[16], start_cnt[16];
  __shared__ int32_t start_fmt[16];
  const int lane = threadIdx.x & 63;
  const BatchTables tab = batch_tables(a.block);
  const uint32_t wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t nw = blockDim.x >> 6;
  // < ntiles: the grid covers the tiles exactly
  const uint64_t tile0 = xcd_block(blockIdx.x, gridDim.x, a.xcds, a.xcd_group) * nw;
  if (wib * 8 < nw) {
    const uint32_t target = wib * 8 + (lane >> 3);
    const bool mine = target < nw && tile0 + target < tab.ntiles;
    const TileInfo ti = lookup_tile(tab, a, mine ? tile0 + target : tile0);
    uint32_t s[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) s[j] = ti.state[j];
    jump_groups(s, mine ? ti.offset : 0, a.jump);
    if (mine && (lane & 7) == 0) {
#pragma unroll
      for (int j = 0; j < 8; ++j) start_state[target][j] = s[j];
      start_dst[target] = ti.dst;
      start_cnt[target] = ti.cnt;
      start_fmt[target] = ti.fmt;
    }
  }
  __syncthreads();
  if (tile0 + wib >= tab.ntiles) return;
  uint32_t u[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) u[j] = __builtin_amdgcn_readfirstlane(start_state[wib][j]);
  const uint64_t d = start_dst[wib], c = start_cnt[wib];
  void* out = reinterpret_cast<void*>(readlane_u64(d, 0));
  const uint64_t cnt = readlane_u64(c, 0);
  switch (__builtin_amdgcn_readfirstlane(start_fmt[wib])) {
    case 0: batch_tile_alone<0>(out, cnt, u, a.lane); break;
    case 1: batch_tile_alone<1>(out, cnt, u, a.lane); break;
    default: batch_tile_alone<2>(out, cnt, u, a.lane); break;
  }
}

// Persistent batched kernel that takes its work in address order: each block
// in turn takes the next group of (waves per block) consecutive tiles from a
// counter in the batch block's header, and then works on it as k_batch_tiles
// works on its block's tiles (waves 0-1 look up and jump the group's start
// states and leave them in LDS, every wave generates one tile).  Blocks stay
// resident, but the grid writes its tiles in the order a one-tile-per-wave
// grid is dispatched in.  The last block to finish sets the counters back to
// zero for the next launch; the counters make launches of one batch block
// that overlap in time incorrect, which the host prevents.
__global__ __launch_bounds__(1024) void k_batch_queue(BatchArgs a) {
  __shared__ uint32_t start_state[16][8];
  __shared__ uint64_t start_dst[16], start_cnt[16];
  __shared__ int32_t start_fmt[16];
  __shared__ uint32_t group;
  const int lane = threadIdx.x & 63;
  const BatchTables tab = batch_tables(a.block);
  BatchHeader* hdr = reinterpret_cast<BatchHeader*>(const_cast<uint8_t*>(a.block));
  const uint32_t wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t nw = blockDim.x >> 6;
  const uint64_t ngroups = (tab.ntiles + nw - 1) / nw;
  for (;;) {
    if (threadIdx.x == 0) group = atomicAdd(&hdr->queue_next, 1u);
    __syncthreads();
    const uint64_t g = group;  // the same value in every wave: the block leaves the loop together
    if (g >= ngroups) break;
    const uint64_t tile0 = g * nw;
    if (wib * 8 < nw) {
      const uint32_t target = wib * 8 + (lane >> 3);
      const bool mine = target < nw && tile0 + target < tab.ntiles;
      const TileInfo ti = lookup_tile(tab, a, mine ? tile0 + target : tile0);
      uint32_t s[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) s[j] = ti.state[j];
      jump_groups(s, mine ? ti.offset : 0, a.jump);
      if (mine && (lane & 7) == 0) {
#pragma unroll
        for (int j = 0; j < 8; ++j) start_state[target][j] = s[j];
        start_dst[target] = ti.dst;
        start_cnt[target] = ti.cnt;
        start_fmt[target] = ti.fmt;
      }
    }
    __syncthreads();
    if (tile0 + wib < tab.ntiles) {
      uint32_t u[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) u[j] = __builtin_amdgcn_readfirstlane(start_state[wib][j]);
      const uint64_t d = start_dst[wib], c = start_cnt[wib];
      void* out = reinterpret_cast<void*>(readlane_u64(d, 0));
      const uint64_t cnt = readlane_u64(c, 0);
      switch (__builtin_amdgcn_readfirstlane(start_fmt[wib])) {
        case 0: batch_tile_alone<0>(out, cnt, u, a.lane); break;
        case 1: batch_tile_alone<1>(out, cnt, u, a.lane); break;
        default: batch_tile_alone<2>(out, cnt, u, a.lane); break;
      }
    }
    __syncthreads();  // LDS and `group` are rewritten by the next group
  }
  if (threadIdx.x == 0) {
    __threadfence();  // this block's last take from the counter comes before its count of done blocks
    if (atomicAdd(&hdr->queue_done, 1u) == gridDim.x - 1) {
      // every block has left the loop: no block takes from the counter again
      atomicExch(&hdr->queue_next, 0u);
      atomicExch(&hdr->queue_done, 0u);
    }
  }
}

// ---- bulk advance: many states, offsets of up to kAdvanceWords words --------

__device__ __forceinline__ const CoefRecord* wide_rec(const AdvanceArgs& a, int k, uint32_t d, int row) {
  if (d == 0u) return a.jump + (kIdentityRec + row);
  return k < kJumpPositions ? a.jump + ((k * kJumpValues + (int)d - 1) * 8 + row)
                            : a.wide + (((k - kJumpPositions) * kJumpValues + (int)d - 1) * 8 + row);
}

// Eight lanes per state (lane r of a group computes row r, as in the jumps
// above); byte positions of the offset in order, each digit extracted once,
// positions where no state of the wave has a digit skipped, the next
// position's records fetched while the current step computes.  Waves whose
// states have no zero word skip the zero-product test and keep the words
// centred between steps (2^20 states, 200-bit offsets: 400 us against 446 us
// for the plain loop).  A state is read before it is written, so `in` and
// `out` may be the same buffer.
// jump_step for states with no zero word, on centred words e = s - 2^30:
// each lane centres its one new word before the gather instead of every lane
// centring eight after it.
__device__ __forceinline__ void jump_step_centred(int32_t e[8], const Coef& c) {
  const uint32_t ey = dot_residue(c, e) + 1u - kCenter;
  uint32_t w[8];
  gather_group(ey, w);
#pragma unroll
  for (int j = 0; j < 8; ++j) e[j] = (int32_t)w[j];
}

template <bool ZERO_WORDS>
__device__ __forceinline__ void advance_digits(uint32_t s[8], const uint64_t (&off)[kAdvanceWords],
                                               const AdvanceArgs& a, int row) {
  const int npos = 8 * (int)a.words;
  int32_t e[8];
  if (!ZERO_WORDS) centre(s, e);
  uint32_t d = (uint32_t)off[0] & (uint32_t)kJumpValues;
  Coef c = load_coef11(wide_rec(a, 0, d, row));
#pragma unroll
  for (int w = 0; w < kAdvanceWords; ++w) {
    if (w >= (int)a.words) break;
    for (int b = 0; b < 8; ++b) {
      const int k = 8 * w + b;
      uint32_t dn;
      if (b < 7)
        dn = (uint32_t)(off[w] >> (8 * (b + 1))) & (uint32_t)kJumpValues;
      else
        dn = w + 1 < kAdvanceWords ? (uint32_t)off[w + 1 < kAdvanceWords ? w + 1 : w] & (uint32_t)kJumpValues : 0u;
      const bool last = k + 1 == npos;
      if (last) dn = 0u;
      const Coef next = load_coef11(wide_rec(a, last ? 0 : k + 1, dn, row));
      if (__ballot(d != 0u)) {
        if (ZERO_WORDS)
          jump_step<true>(s, c);
        else
          jump_step_centred(e, c);
      }
      c = next;
      d = dn;
    }
  }
  if (!ZERO_WORDS) {
#pragma unroll
    for (int j = 0; j < 8; ++j) s[j] = (uint32_t)e[j] + kCenter;
  }
}

__global__ __launch_bounds__(256) void k_advance(AdvanceArgs a) {
  const int row = threadIdx.x & 7;
  const uint64_t item = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 3;
  const bool live = item < a.n;
  uint32_t s[8];
  uint64_t off[kAdvanceWords];
#pragma unroll
  for (int w = 0; w < kAdvanceWords; ++w) off[w] = live && w < (int)a.words ? a.offsets[item * a.words + w] : 0;
  uint32_t smin = 1u;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    s[j] = live ? a.in[8 * item + j] : 1u;
    smin = min(smin, s[j]);
  }
  // a state with no zero word never gets one (no row of a power of A is
  // zero), so the zero-product test is needed only by waves that start with one
  if (__ballot(smin == 0u))
    advance_digits<true>(s, off, a, row);
  else
    advance_digits<false>(s, off, a, row);
  if (live && row == 0) {
#pragma unroll
    for (int j = 0; j < 8; ++j) a.out[8 * item + j] = s[j];
  }
}

// Staging copy of a batch block: the host writes the block into coherent
// pinned memory and this kernel moves it to device memory (16 bytes per
// thread, then 4-byte words for the tail).  On MI355X the runtime's
// host-to-device copy of these few kilobytes occasionally blocked the
// calling thread for 6-7 ms, where a kernel launch returns in microseconds.
__global__ __launch_bounds__(256) void k_stage(const u4_t* __restrict__ src, u4_t* __restrict__ dst, uint64_t n16,
                                                uint64_t tail_words) {
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n16) dst[i] = src[i];
  if (i < tail_words)
    reinterpret_cast<uint32_t*>(dst + n16)[i] = reinterpret_cast<const uint32_t*>(src + n16)[i];
}

// ---- matrix-core fill ---------------------------------------------------------
//
// A tile of 16 COL consecutive outputs is cut into 16 columns of COL (64
// or 128); column c starts from its own state (its 8 preceding words)
// and one step makes its next 16 outputs.  Output j of a step (j < 16) is
// dot(R_j, w) for the column's state w (oldest word first), with R_j the
// block rows.  In exact integers mod p:
//   * each state word is centred, e = w - 2^30, and cut into four balanced
//     base-256 digits e = sum_b 2^(8b) d_b (d_b in [-128, 127]);
//   * M_b = 2^(8b) R_j mod p (centred) is cut the same way,
//     M_b = sum_a 2^(8a) m_(b,a), so that
//     dot(R_j, e) == sum_a 2^(8a) sum_(q,b) m_(b,a)[q] d_b[q]  (mod p):
//     four int8 products of K = 8 words x 4 digits = 32, one per class a,
//     v_mfma_i32_16x16x32_i8 with the step's rows as M, the columns as N;
//   * the accumulators start from per-row constants (the C operand) that
//     carry the 2^30 sum R_j, the reference's -1 and the biases, so that
//     t = C_0 + 2^8 C_1 and u = C_2 + 2^8 C_3 are non-negative and below
//     2^29, and x = t + ((u << 16) & p) + (u >> 15) (2^16 u mod p is a 31-bit
//     rotation) is the residue plus at most one p: eight 32-bit operations
//     per output.
// MFMA layouts (gfx950; the fill's parity tests depend on them): lane
// l = n + 16 g holds the 8 bytes of K-group g for row/column n, and result
// register i of lane l is row 4 g + i of column n.  K-group g, byte i is word W(g) + i / 4, digit
// i % 4, with W = (0, 4, 2, 6); row 4 g + i is step position 2 g + i for
// i < 2 and 8 + W(g) + i - 2 for i >= 2.  So the step's last eight outputs
// (positions 8..15, the column's next state) are registers 2 and 3 of
// exactly the lanes whose K-group needs them: no lane exchange.
// A wave makes a few consecutive tiles (32 KiB of output: the store
// pattern of short-lived waves in address order, which persistent waves
// striding over the tiles did not match: 427-474 against 339 us for 2^28
// doubles).  The 16 column states of the wave's next tile are the current
// ones times A^2048: one more product of the same form, whose rows put the
// new words in registers 0 and 1 of the lanes that need them.  Only a
// wave's first tile computes its column states from the tile state (the
// fan-out records).  Registers stay below 128 and the tile's staging below
// 8.5 KiB, so four waves per SIMD are resident.
// The outputs of a tile are staged in LDS (16 columns x 132 words per wave)
// and leave as the vector kernel's 1 KiB wave stores in address order.

typedef int32_t v4i_t __attribute__((ext_vector_type(4)));


// Balanced base-256 digits of w - 2^30 (w < 2^31), one per byte.
__device__ __forceinline__ uint32_t mfma_digits(uint32_t w) { return (w + 0x40808080u) ^ 0x80808080u; }

__device__ __forceinline__ long mfma_operand(uint32_t w0, uint32_t w1) {
  return (long)((uint64_t)mfma_digits(w0) | ((uint64_t)mfma_digits(w1) << 32));
}

__device__ __forceinline__ uint64_t mad_u64_u32(uint32_t a, uint32_t b, uint64_t c) {
  uint64_t w, carry;
  asm("v_mad_u64_u32 %0, %1, %2, %3, %4" : "=v"(w), "=s"(carry) : "v"(a), "s"(b), "v"(c));
  return w;
}

// Residue in [0, p) of a step output (result register i of the four
// classes, accumulators started from the step constants).
__device__ __forceinline__ uint32_t mfma_step_residue(const v4i_t (&acc)[kMfmaClasses], int i) {
  const uint32_t t = (uint32_t)acc[0][i] + ((uint32_t)acc[1][i] << 8);  // [0, 2 beta + 2^16)
  const uint32_t u = (uint32_t)acc[2][i] + ((uint32_t)acc[3][i] << 8);  // [0, 2 beta + 2^15)
  const uint32_t x = t + ((u << 16) & kP) + (u >> 15);                  // < 2 p
  return min(x, x - kP);
}

// Residue in [0, p) of a stride output (accumulators started from 0, k in
// [kMfmaBeta, kMfmaBeta + p)).
__device__ __forceinline__ uint32_t mfma_stride_residue(const v4i_t (&acc)[kMfmaClasses], int i, uint32_t k) {
  const uint32_t t = (uint32_t)acc[0][i] + ((uint32_t)acc[1][i] << 8);              // [-beta, beta]
  const uint32_t u = (uint32_t)acc[2][i] + ((uint32_t)acc[3][i] << 8) + kMfmaBeta;  // [0, 2 beta]
  const uint64_t y = mad_u64_u32(u, 65536u, (uint64_t)(t + k));                    // < 2^45
  const uint32_t x = ((uint32_t)y & kP) + (uint32_t)(y >> 31);                      // < 2^31 + 2^14
  return min(x, x - kP);
}

// The tiles [t0, t0 + nt) of a matrix-core fill with columns of COL outputs,
// made by one wave with the LDS area st (kMfmaColumns * (COL + 4) words: the
// 4-word pad spreads the columns over the LDS banks).
template <int FMT, int COL>
__device__ __forceinline__ void mfma_tiles(const MfmaFillArgs& a, uint64_t t0, uint32_t nt, uint32_t* st) {
  typedef typename Out<FMT>::T T;
  constexpr int kTile = kMfmaColumns * COL;
  constexpr int kStride = COL + 4;  // LDS words per column
  constexpr int kSteps = COL / 16;
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t n = lane & 15, g = lane >> 4;
  const uint32_t wg = (g & 1) * 4 + (g >> 1) * 2;  // W(g): first word of this K-group
  long A[kMfmaClasses];
  v4i_t C0[kMfmaClasses];
#pragma unroll
  for (int q = 0; q < kMfmaClasses; ++q) {
    A[q] = reinterpret_cast<const long*>(a.mfma_a)[q * 64 + lane];
    C0[q] = reinterpret_cast<const v4i_t*>(a.mfma_c)[q * 4 + g];
  }
  long AJ[kMfmaClasses] = {};
  u2_t KJ = {};
  if (nt > 1) {
#pragma unroll
    for (int q = 0; q < kMfmaClasses; ++q) AJ[q] = reinterpret_cast<const long*>(a.mfma_next_a)[q * 64 + lane];
    KJ = reinterpret_cast<const u2_t*>(a.mfma_next_k)[g];
  }
  // the first tile's column states: words W(g), W(g) + 1 (oldest first) of
  // column n are the tile's outputs 128 n - 8 + q, from the tile state
  long B;
  {
    uint32_t s[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) s[j] = a.start.state[j];
    jump_groups(s, a.start.offset + a.head + t0 * kTile, a.jump);
#pragma unroll
    for (int j = 0; j < 8; ++j) s[j] = __builtin_amdgcn_readfirstlane(s[j]);
    Coef k2[2];
    const CoefRecord* fan = a.mfma_fan + (n ? n - 1 : 0) * 8 + wg;  // outputs COL n - 8 + q
    k2[0] = load_coef(fan);
    k2[1] = load_coef(fan + 1);
    int32_t e[8];
    centre(s, e);
    uint32_t r[2];
    dot_residues<2>(k2, e, r);
    // column 0 starts from the tile state itself: word q is s[7 - q]
    uint32_t s0 = s[7], s1 = s[6];
#pragma unroll
    for (int q = 2; q < 8; q += 2) {
      s0 = wg == (uint32_t)q ? s[7 - q] : s0;
      s1 = wg == (uint32_t)q ? s[6 - q] : s1;
    }
    B = mfma_operand(n ? r[0] + 1u : s0, n ? r[1] + 1u : s1);
  }

  // A tile's outputs go through LDS into this lane's slice of its 1 KiB wave
  // stores (pend: 32 residues, chunk ch = this lane's part of store ch), and
  // the stores leave during the next tile's steps, kChunks / kSteps per step,
  // so that every wave keeps writing while it computes.
  constexpr int kChunks = kTile * (int)sizeof(T) / 1024;  // 1 KiB stores per tile
  constexpr int kPer = 1024 / (64 * (int)sizeof(T));      // outputs per lane per store
  static_assert(kChunks % kSteps == 0 && COL % kPer == 0, "tile shape");
  uint32_t pend[kChunks * kPer];
  T* pend_base = nullptr;
  auto store_chunk = [&](int ch) {
    if constexpr (FMT == 2) {
      const u4_t o = {pend[4 * ch] + 1u, pend[4 * ch + 1] + 1u, pend[4 * ch + 2] + 1u, pend[4 * ch + 3] + 1u};
      __builtin_nontemporal_store(o, reinterpret_cast<u4_t*>(pend_base + 256 * ch + 4 * lane));
    } else {
      const d2_t o = {convert<FMT>(pend[2 * ch]), convert<FMT>(pend[2 * ch + 1])};
      __builtin_nontemporal_store(o, reinterpret_cast<d2_t*>(pend_base + 128 * ch + 2 * lane));
    }
  };
  uint32_t* const col = st + n * kStride;
  for (uint32_t t = 0; t < nt; ++t) {
    const long B0 = B;
#pragma unroll
    for (int step = 0; step < kSteps; ++step) {
      v4i_t acc[kMfmaClasses];
#pragma unroll
      for (int q = 0; q < kMfmaClasses; ++q) acc[q] = __builtin_amdgcn_mfma_i32_16x16x32_i8(A[q], B, C0[q], 0, 0, 0);
      uint32_t r[4];
#pragma unroll
      for (int i = 3; i >= 0; --i) r[i] = mfma_step_residue(acc, i);  // registers 2, 3 first: the next operand
      if (step + 1 < kSteps) B = mfma_operand(r[2] + 1u, r[3] + 1u);
      *reinterpret_cast<u2_t*>(col + 16 * step + 2 * g) = u2_t{r[0], r[1]};
      *reinterpret_cast<u2_t*>(col + 16 * step + 8 + wg) = u2_t{r[2], r[3]};
      if (t) {
#pragma unroll
        for (int k = 0; k < kChunks / kSteps; ++k) store_chunk(step * (kChunks / kSteps) + k);
      }
    }
    // the next tile's column states: A^(16 COL) times this tile's (register
    // i < 2 of this lane: word W(g) + i)
    if (t + 1 < nt) {
      v4i_t acc[kMfmaClasses];
#pragma unroll
      for (int q = 0; q < kMfmaClasses; ++q)
        acc[q] = __builtin_amdgcn_mfma_i32_16x16x32_i8(AJ[q], B0, v4i_t{}, 0, 0, 0);
      B = mfma_operand(mfma_stride_residue(acc, 0, KJ.x) + 1u, mfma_stride_residue(acc, 1, KJ.y) + 1u);
    }
    // this tile's slice of every store: tile outputs ch E + kPer lane .. + kPer - 1
    // (E = 1024 / sizeof(T) per store), within one column
#pragma unroll
    for (int ch = 0; ch < kChunks; ++ch) {
      const uint32_t pos = ch * (1024 / (int)sizeof(T)) + kPer * lane;
      const uint32_t* src = st + (pos / COL) * kStride + pos % COL;
      if constexpr (FMT == 2) {
        const u4_t v = *reinterpret_cast<const u4_t*>(src);
        pend[4 * ch] = v.x; pend[4 * ch + 1] = v.y; pend[4 * ch + 2] = v.z; pend[4 * ch + 3] = v.w;
      } else {
        const u2_t v = *reinterpret_cast<const u2_t*>(src);
        pend[2 * ch] = v.x; pend[2 * ch + 1] = v.y;
      }
    }
    pend_base = static_cast<T*>(a.out) + a.head + (t0 + t) * kTile;
  }
#pragma unroll
  for (int ch = 0; ch < kChunks; ++ch) store_chunk(ch);
}

// Matrix-core fill: waves [0, a.waves) make tiles, the next a.extra waves
// the short runs before the first tile and after the last with the vector
// kernel's loop.
template <int FMT, int COL, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_fill_mfma(MfmaFillArgs a) {
  __shared__ uint32_t stage[WAVES][kMfmaColumns * (COL + 4)];
  const uint32_t wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint64_t w = (uint64_t)blockIdx.x * WAVES + wib;
  if (w < a.waves) {
    const uint64_t t0 = w * a.tiles_per_wave;
    mfma_tiles<FMT, COL>(a, t0, (uint32_t)min((uint64_t)a.tiles_per_wave, a.tiles - t0), stage[wib]);
    return;
  }
  const uint64_t e = w - a.waves;
  if (e >= a.extra) return;
  const uint64_t first = a.extra_first[e], cnt = a.extra_count[e];
  uint32_t s[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) s[j] = a.start.state[j];
  jump_groups(s, a.start.offset + first, a.jump);
#pragma unroll
  for (int j = 0; j < 8; ++j) s[j] = __builtin_amdgcn_readfirstlane(s[j]);
  Coef k[4];
  load_lane_coefs<FMT, 4>(k, a.lane);
  generate_run<FMT, 4>(static_cast<typename Out<FMT>::T*>(a.out) + first, cnt, s, k, a.lane);
}

template <int FMT, int PL>
hipError_t fill_one(const FillArgs& a, int block_threads, hipStream_t stream) {
  if (a.n == 0) return hipSuccess;
  const uint64_t waves = fill_waves(a.n, a.head, a.tile);
  const uint64_t wpb = (uint64_t)block_threads / 64;
  const uint64_t blocks = (waves + wpb - 1) / wpb;
  if (blocks * (uint64_t)block_threads > 0xffffffffull) return hipErrorInvalidConfiguration;  // 32-bit grid
  hipLaunchKernelGGL((k_fill<FMT, PL>), dim3((uint32_t)blocks), dim3(block_threads), 0, stream, a);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_fill(const FillArgs& a, int format, int per_lane, int block_threads, hipStream_t stream) {
  switch (format * 8 + per_lane) {
    case 0 * 8 + 1: return fill_one<0, 1>(a, block_threads, stream);
    case 0 * 8 + 2: return fill_one<0, 2>(a, block_threads, stream);
    case 0 * 8 + 4: return fill_one<0, 4>(a, block_threads, stream);
    case 1 * 8 + 1: return fill_one<1, 1>(a, block_threads, stream);
    case 1 * 8 + 2: return fill_one<1, 2>(a, block_threads, stream);
    case 1 * 8 + 4: return fill_one<1, 4>(a, block_threads, stream);
    case 2 * 8 + 1: return fill_one<2, 1>(a, block_threads, stream);
    case 2 * 8 + 2: return fill_one<2, 2>(a, block_threads, stream);
    case 2 * 8 + 4: return fill_one<2, 4>(a, block_threads, stream);
    default: return hipErrorInvalidValue;
  }
}

namespace {

template <int FMT, int COL, int WAVES>
hipError_t fill_mfma_one(const MfmaFillArgs& a, uint64_t blocks, hipStream_t stream) {
  hipLaunchKernelGGL((k_fill_mfma<FMT, COL, WAVES>), dim3((uint32_t)blocks), dim3(64 * WAVES), 0, stream, a);
  return hipGetLastError();
}

template <int FMT, int COL>
hipError_t fill_mfma_waves(const MfmaFillArgs& a, uint64_t blocks, hipStream_t stream) {
  switch (a.waves_per_block) {
    case 4: return fill_mfma_one<FMT, COL, 4>(a, blocks, stream);
    case 8: return fill_mfma_one<FMT, COL, 8>(a, blocks, stream);
    case 16: return fill_mfma_one<FMT, COL, 16>(a, blocks, stream);
    default: return hipErrorInvalidValue;
  }
}

template <int FMT>
hipError_t fill_mfma_column(const MfmaFillArgs& a, uint64_t blocks, hipStream_t stream) {
  return a.column == 64 ? fill_mfma_waves<FMT, 64>(a, blocks, stream) : fill_mfma_waves<FMT, 128>(a, blocks, stream);
}

}  // namespace

hipError_t launch_fill_mfma(const MfmaFillArgs& a, int format, hipStream_t stream) {
  const uint64_t waves = a.waves + a.extra;
  if (waves == 0) return hipSuccess;
  if (a.extra > 2 || a.tiles_per_wave == 0 || a.waves != (a.tiles + a.tiles_per_wave - 1) / a.tiles_per_wave ||
      (a.column != 64 && a.column != 128) || (a.waves_per_block != 4 && a.waves_per_block != 8 && a.waves_per_block != 16))
    return hipErrorInvalidValue;
  const uint64_t blocks = (waves + a.waves_per_block - 1) / a.waves_per_block;
  if (blocks * 64 * a.waves_per_block > 0xffffffffull) return hipErrorInvalidConfiguration;  // 32-bit grid
  switch (format) {
    case 0: return fill_mfma_column<0>(a, blocks, stream);
    case 1: return fill_mfma_column<1>(a, blocks, stream);
    case 2: return fill_mfma_column<2>(a, blocks, stream);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_batch(const BatchArgs& a, int grid, int block_threads, hipStream_t stream) {
  hipLaunchKernelGGL(k_batch, dim3(grid), dim3(block_threads), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_batch_queue(const BatchArgs& a, int grid, int block_threads, hipStream_t stream) {
  if (grid <= 0) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL(k_batch_queue, dim3(grid), dim3(block_threads), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_advance(const AdvanceArgs& a, hipStream_t stream) {
  if (a.n == 0) return hipSuccess;
  const uint64_t threads = a.n * 8;
  if (a.n > (0xffffffffull >> 3) || a.words < 1 || a.words > (uint32_t)kAdvanceWords) return hipErrorInvalidValue;
  const uint64_t blocks = (threads + 255) / 256;
  hipLaunchKernelGGL(k_advance, dim3((uint32_t)blocks), dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_stage(const void* src, void* dst, size_t bytes, hipStream_t stream) {
  if (bytes == 0) return hipSuccess;
  if (bytes % 4 || (reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) % 16)
    return hipErrorInvalidValue;
  const uint64_t n16 = bytes / 16, tail = (bytes % 16) / 4;
  const uint64_t blocks = (n16 + 255) / 256 + (n16 == 0 ? 1 : 0);
  if (blocks > 0xffffffffull) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_stage, dim3((uint32_t)blocks), dim3(256), 0, stream, static_cast<const u4_t*>(src),
                     static_cast<u4_t*>(dst), n16, tail);
  return hipGetLastError();
}

hipError_t launch_batch_tiles(const BatchArgs& a, uint64_t ntiles, int block_threads, hipStream_t stream) {
  const uint64_t wpb = (uint64_t)block_threads / 64;
  const uint64_t blocks = (ntiles + wpb - 1) / wpb;
  if (blocks == 0) return hipSuccess;
  if (blocks * (uint64_t)block_threads > 0xffffffffull) return hipErrorInvalidConfiguration;  // 32-bit grid
  hipLaunchKernelGGL(k_batch_tiles, dim3((uint32_t)blocks), dim3(block_threads), 0, stream, a);
  return hipGetLastError();
}

}  // namespace mrg8
