// Split-precision building blocks of the mixed form: complex products on the
// BF16 matrix cores (v_mfma_f32_16x16x32_bf16, FP32 accumulation) with each
// FP32 operand split into a BF16 pair, x = hi + lo,
//   hi = RN_bf16(x),  lo = RN_bf16(x - hi)   (x - hi is exact in FP32),
// and a product taken as hi*hi + hi*lo + lo*hi (lo*lo, below 2^-16 of the
// product, is dropped). The pair carries about 16 significant bits, so a
// product's relative error is a few units of 2^-17.
//
// Complex layout for the 16x16x32 instruction. Lane l holds 8 K values
// k = 8*(l/16) + v of row (A) or column (B) l%16. K runs over (index, part)
// pairs: v = 2*a + p is element a = 0..3 of the lane's four consecutive
// K indices (antennas or users) and part p = 0 (re), 1 (im). For C = X Z with
// complex X (rows) and Z (columns), output rows interleaved as (r re, r im):
//   A[2r][(k,0)] = Re X, A[2r][(k,1)] = -Im X, A[2r+1][(k,0)] = Im X,
//   A[2r+1][(k,1)] = Re X;   B[(k,0)][n] = Re Z, B[(k,1)][n] = Im Z,
// so D[2r][n] = Re C and D[2r+1][n] = Im C, and each lane ends with whole
// complex results (rows 4*(l/16) .. +3 = two complex rows) as in the FP32
// layout of eq_core.h. One instruction covers 16 complex K indices.
//
// A BF16 pair (re, im) travels as one 32-bit word, re in the low half. The
// A operand of row 2r is that word with the sign of the high half flipped,
// of row 2r+1 the word with its halves swapped.
#pragma once

#include "eq_core.h"

namespace mega {
namespace detail {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef uint32_t v4u __attribute__((ext_vector_type(4)));
typedef float v2f __attribute__((ext_vector_type(2)));
typedef __bf16 v2bf __attribute__((ext_vector_type(2)));

// Row length (words) of the split weight planes: B rounded up to whole
// 16-antenna MFMA steps (the pad holds zeros).
__host__ __device__ constexpr uint32_t split_row(uint32_t B) { return (B + 15) / 16 * 16; }

__device__ __forceinline__ v4f mfma_bf16(v4u a, v4u b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a),
                                                 __builtin_bit_cast(v8bf, b), c, 0, 0, 0);
}

// (re, im) -> packed BF16 words of the hi and lo parts.
struct SplitC {
  uint32_t hi, lo;
};

__device__ __forceinline__ SplitC split_c(float2 z) {
  const v2bf h = __builtin_convertvector((v2f){z.x, z.y}, v2bf);
  const uint32_t hw = __builtin_bit_cast(uint32_t, h);
  const float hre = __uint_as_float(hw << 16), him = __uint_as_float(hw & 0xffff0000u);
  const v2bf l = __builtin_convertvector((v2f){z.x - hre, z.y - him}, v2bf);
  return SplitC{hw, __builtin_bit_cast(uint32_t, l)};
}

// A-operand word of an interleaved output row: h = 0 (re row) -> (re, -im),
// h = 1 (im row) -> (im, re). Branch-free for a lane-varying h: a rotate by
// 16*h then a sign flip of the high half when h = 0 (ARow holds both
// per-lane constants).
struct ARow {
  uint32_t rot, flip;
  __device__ __forceinline__ explicit ARow(bool h)
      : rot(h ? 16u : 0u), flip(h ? 0u : 0x80000000u) {}
};

__device__ __forceinline__ uint32_t a_word(uint32_t w, ARow r) {
  return __builtin_amdgcn_alignbit(w, w, r.rot) ^ r.flip;
}

// acc += X Z over one K step of 16 complex indices, both split.
__device__ __forceinline__ v4f mfma_x3(const v4u& ah, const v4u& al, const v4u& bh,
                                       const v4u& bl, v4f acc) {
  acc = mfma_bf16(al, bh, acc);
  acc = mfma_bf16(ah, bl, acc);
  return mfma_bf16(ah, bh, acc);
}

// Three-piece split, for products whose errors a condition number amplifies
// (the Gram matrix that the inverse is refined against, W = X A): x = hi + lo
// + rest with rest = RN_bf16((x - hi) - lo), both differences exact in FP32.
// The pieces carry about 24 significant bits; a product taken as
//   hi*hi + (lo*hi + hi*lo) + (lo*lo + rest*hi + hi*rest)
// drops terms below 2^-25 of it, under the FP32 accumulation's own rounding.
struct SplitC3 {
  uint32_t hi, lo, rest;
};

__device__ __forceinline__ SplitC3 split_c3(float2 z) {
  const SplitC s = split_c(z);
  const float hre = __uint_as_float(s.hi << 16), him = __uint_as_float(s.hi & 0xffff0000u);
  const float lre = __uint_as_float(s.lo << 16), lim = __uint_as_float(s.lo & 0xffff0000u);
  const v2bf r = __builtin_convertvector((v2f){(z.x - hre) - lre, (z.y - him) - lim}, v2bf);
  return SplitC3{s.hi, s.lo, __builtin_bit_cast(uint32_t, r)};
}

// acc += X Z over one K step of 16 complex indices, both split in three;
// the small terms first.
__device__ __forceinline__ v4f mfma_x6(const v4u& ah, const v4u& al, const v4u& ar,
                                       const v4u& bh, const v4u& bl, const v4u& br, v4f acc) {
  acc = mfma_bf16(al, bl, acc);
  acc = mfma_bf16(ar, bh, acc);
  acc = mfma_bf16(ah, br, acc);
  return mfma_x3(ah, al, bh, bl, acc);
}

}  // namespace detail
}  // namespace mega
