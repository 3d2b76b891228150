// Device building blocks of the fused equalize + demod GEMM, shared by the
// stand-alone equalize kernel and the fused ZF + equalize kernel.
//
// Per subcarrier group the work of a slot is one complex GEMM,
//   X (U x sg*S) = W_g (U x B) * Y (B x sg*S),
// on the FP32 matrix cores (v_mfma_f32_16x16x4_f32: FP32 products, FP32
// accumulation). The complex product is split into two real MFMAs per step of
// 4 antennas with the output rows interleaved as (u re, u im):
//   MFMA1: A1[2u][k] = Re W[u,b], A1[2u+1][k] = Im W[u,b], B1[k][n] = Re y[b,n]
//   MFMA2: A2[2u][k] = -Im W[u,b], A2[2u+1][k] = Re W[u,b], B2[k][n] = Im y[b,n]
// so D[2u][n] = Re x, D[2u+1][n] = Im x: every lane ends up holding whole
// equalized symbols (users 2q, 2q+1 of its 16-row tile at column n) for the
// fused demod. A 16-row M tile covers 8 users; a 16-column N tile covers 16
// (symbol, subcarrier) columns of the group.
//
// Each wave keeps its slice of W_g (MT M tiles x KS antenna steps) in
// registers and streams y straight from HBM (lane = column: 16 lanes read 128
// contiguous bytes of an antenna row; one float2 feeds both MFMAs), with the
// next tile's y in flight during the current tile's MFMAs. When B is wider
// than one wave's KS steps, the waves of a workgroup split the antennas and
// their partial sums are reduced through LDS (double-buffered, one barrier per
// tile).
#pragma once

#include <type_traits>

#include "mega/demod_core.h"
#include "internal.h"

namespace mega {
namespace detail {

// f(integral_constant<int, I>) for I = B .. E-1, unrolled at compile time.
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4f mfma4(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

struct EqParams {
  const float2* csi;  // W (equalize) or A = H^H (fused detect), [slots][G][B][U]
  const float2* y;
  _Float16* llr;
  float2* xhat;
  float2* w_out;      // fused detect: optional W output, same layout as csi
  int32_t* info;      // fused detect: optional per-group status [slots*G]
  uint32_t U, B, N, sg, S, G, slots;
  uint32_t nmg;              // workgroups per tile: groups of 8*MT users
  uint32_t nwk;              // waves per workgroup: antenna splits
  uint32_t col_tiles;        // 16-column tiles of a full group's sg*S columns
  uint32_t tiles_per_wg, chunks;
  uint64_t ue_stride, cb_size, cb_stride, zeros, valid_bits;
  uint64_t csi_slot, y_sym, y_slot, llr_sym, llr_slot;
  uint32_t cb_size32;  // cb_size (the host checks N*Q and cb_size fit 32 bits)
  uint32_t sg_inv, cb_inv;  // div_inv(sg), div_inv(cb_size32)
  uint32_t last_inv;        // div_inv of the last group's subcarriers
  uint32_t store32;    // every LLR offset even and the base 4-byte aligned
  uint32_t Q;          // modulation order (used by the runtime-Q epilogue, Q = 0)
};

// n / d for any 32-bit n by a multiply: inv = div_inv(d) = ceil(2^32 / d)
// (0 for d = 1). n * inv / 2^32 = n / d + n * e / 2^32 with 0 <= e < 1, so
// its integer part is the quotient or one above it, and one comparison
// corrects it. The compiler's expansion of a division by a value it does not
// know is about 18 vector instructions.
__host__ __device__ constexpr uint32_t div_inv(uint32_t d) {
  return d <= 1 ? 0u : (uint32_t)(((1ull << 32) + d - 1) / d);
}
__device__ __forceinline__ uint32_t div_by(uint32_t n, uint32_t d, uint32_t inv) {
  if (inv == 0) return n;  // d = 1
  const uint32_t q = __umulhi(n, inv);
  return (uint64_t)q * d > n ? q - 1 : q;
}
// Symbol of column n of a group of nk subcarriers: nk is sg except in a
// partial last group.
__device__ __forceinline__ uint32_t group_sym(const EqParams& p, uint32_t n, uint32_t nk) {
  return div_by(n, nk, nk == p.sg ? p.sg_inv : p.last_inv);
}

// Where the Q LLRs of subcarrier k go within a user row
// (equalize_kernel.cuh:461-478), and whether they can be written as Q/2
// aligned 32-bit words.
struct LlrSlot {
  uint64_t off;
  uint32_t bit;
  bool full;
};

// Q = 0 selects the runtime-Q form (p.Q), used where one kernel serves cells
// of different modulation orders.
template <int Q>
__device__ __forceinline__ LlrSlot llr_slot(const EqParams& p, uint32_t k) {
  const uint32_t q = Q > 0 ? (uint32_t)Q : p.Q;
  LlrSlot s;
  s.bit = k * q;
  const uint32_t cb = div_by(s.bit, p.cb_size32, p.cb_inv);
  s.off = (uint64_t)cb * p.cb_stride + p.zeros + (s.bit - cb * p.cb_size32);
  s.full = p.store32 && s.bit + q <= p.valid_bits && s.off + q <= p.ue_stride;
  return s;
}

template <int Q>
__device__ __forceinline__ void store_llrs(const EqParams& p, _Float16* row,
                                           const LlrSlot& s, float re, float im) {
  if constexpr (Q == 0) {
    switch (p.Q) {
      case 2: store_llrs<2>(p, row, s, re, im); break;
      case 4: store_llrs<4>(p, row, s, re, im); break;
      case 6: store_llrs<6>(p, row, s, re, im); break;
      default: store_llrs<8>(p, row, s, re, im); break;
    }
  } else {
    _Float16 v[Q];
    demod::demod<Q>(re, im, v);
    _Float16* dst = row + s.off;
    if (s.full) {
#pragma unroll
      for (int j = 0; j < Q / 2; ++j) {
        uint32_t lo = __builtin_bit_cast(uint16_t, v[2 * j]);
        uint32_t hi = __builtin_bit_cast(uint16_t, v[2 * j + 1]);
        reinterpret_cast<uint32_t*>(dst)[j] = lo | (hi << 16);
      }
    } else {
#pragma unroll
      for (int j = 0; j < Q; ++j)
        if (s.bit + j < p.valid_bits && s.off + j < p.ue_stride) dst[j] = v[j];
    }
  }
}

// Output side of one column (lane%16) of column tile t of group g: which
// (symbol, subcarrier) it is, where its LLRs go, and the x_hat row.
struct EqCol {
  bool valid;
  LlrSlot ls;
  _Float16* rows;  // user 0's LLR row of this symbol
  float2* xrow;    // user 0's x_hat of this (symbol, subcarrier), or null
};

template <int Q>
__device__ __forceinline__ EqCol eq_col(const EqParams& p, uint32_t slot, uint32_t g,
                                        uint32_t t) {
  const uint32_t k0 = g * p.sg;
  const uint32_t nk = min(p.sg, p.N - k0);
  const uint32_t n = 16 * t + (threadIdx.x & 15);
  EqCol c;
  c.valid = n < nk * p.S;
  const uint32_t sym = c.valid ? group_sym(p, n, nk) : 0;
  const uint32_t k = k0 + (c.valid ? n - sym * nk : 0);
  c.ls = llr_slot<Q>(p, k);
  c.rows = p.llr + slot * p.llr_slot + sym * p.llr_sym;
  c.xrow = p.xhat ? p.xhat + ((uint64_t)slot * p.S + sym) * p.U * p.N + k : nullptr;
  return c;
}

// Demod and store the two equalized symbols of users u0, u0+1 held in an
// MFMA result (d = re u0, im u0, re u0+1, im u0+1).
template <int Q>
__device__ __forceinline__ void eq_store(const EqParams& p, const EqCol& c,
                                         uint32_t u0, v4f d) {
  if (!c.valid) return;
  if (u0 < p.U) {
    store_llrs<Q>(p, c.rows + u0 * p.ue_stride, c.ls, d[0], d[1]);
    if (c.xrow) c.xrow[(uint64_t)u0 * p.N] = make_float2(d[0], d[1]);
  }
  if (u0 + 1 < p.U) {
    store_llrs<Q>(p, c.rows + (u0 + 1) * p.ue_stride, c.ls, d[2], d[3]);
    if (c.xrow) c.xrow[(uint64_t)(u0 + 1) * p.N] = make_float2(d[2], d[3]);
  }
}

// Maps the hardware workgroup id to a logical work index so that consecutive
// logical indices share an XCD (ids are dealt to the 8 XCDs round-robin).
__device__ __forceinline__ uint32_t xcd_linear(uint32_t w, uint32_t T) {
  const uint32_t q = T / 8, r = T % 8, x = w % 8, j = w / 8;
  return x * q + min(x, r) + j;
}

// Waves per workgroup the register budget allows for an (MT, KS) tile.
constexpr int max_waves(int MT, int KS) { return MT * KS <= 8 ? 16 : 8; }

// MFMA A operands of wave (mg, wk)'s slice of W_g, for users
// u = 8*(mg*MT + m) + (lane%16)/2 and antennas b = 4*(wk*KS + s) + lane/16:
//   a1 = lane odd ? Im W[u,b] : Re W[u,b],   a2 = lane odd ? Re W[u,b] : -Im W[u,b]
// (zero outside U x B). Two sources with the same interface:
//   WRegs  - the slice held in 2*MT*KS registers, loaded once;
//   WLds   - read per use from a copy of W in LDS (W[u + ld*b]), which frees
//            those registers for occupancy.
template <int MT, int KS>
struct WRegs {
  float a1[MT][KS], a2[MT][KS];
  __device__ __forceinline__ void load(const float2* W, uint32_t ld, uint32_t U,
                                       uint32_t B, uint32_t mg, uint32_t wk) {
    const int lane = threadIdx.x & 63;
    const int i16 = lane & 15, q4 = lane >> 4;
    const bool odd = i16 & 1;
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      const uint32_t u = 8 * (mg * MT + m) + (i16 >> 1);
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        const uint32_t b = 4 * (wk * KS + s) + q4;
        float2 v = make_float2(0.f, 0.f);
        if (u < U && b < B) v = W[u + (uint64_t)ld * b];
        a1[m][s] = odd ? v.y : v.x;
        a2[m][s] = odd ? v.x : -v.y;
      }
    }
  }
  __device__ __forceinline__ void get(int m, int s, float& x1, float& x2) const {
    x1 = a1[m][s];
    x2 = a2[m][s];
  }
};

template <int MT, int KS>
struct WLds {
  const float2* w;
  uint32_t ld, U, B, mg, wk;
  __device__ __forceinline__ void get(int m, int s, float& x1, float& x2) const {
    const int lane = threadIdx.x & 63;
    const int i16 = lane & 15, q4 = lane >> 4;
    const uint32_t u = 8 * (mg * MT + m) + (i16 >> 1);
    const uint32_t b = 4 * (wk * KS + s) + q4;
    const float2 v = (u < U && b < B) ? w[u + ld * b] : make_float2(0.f, 0.f);
    const bool odd = i16 & 1;
    x1 = odd ? v.y : v.x;
    x2 = odd ? v.x : -v.y;
  }
};

// y operands of column tile t of group g (zeros outside the group's columns
// and beyond B): lane = column, 16 lanes read 128 contiguous bytes of a row.
template <int KS>
__device__ __forceinline__ void eq_load_tile(const EqParams& p, uint32_t slot,
                                             uint32_t g, uint32_t wk, uint32_t t,
                                             float2 (&yv)[KS]) {
  const int lane = threadIdx.x & 63;
  const int i16 = lane & 15, q4 = lane >> 4;
  const uint32_t k0 = g * p.sg;
  const uint32_t nk = min(p.sg, p.N - k0);
  const uint32_t n = 16 * t + i16;
  const bool valid_n = n < nk * p.S;
  const uint32_t sym = valid_n ? group_sym(p, n, nk) : 0;
  const uint32_t kk = valid_n ? n - sym * nk : 0;
  const float2* yc = p.y + slot * p.y_slot + k0 + sym * p.y_sym + kk;
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    const uint32_t b = 4 * (wk * KS + s) + q4;
    yv[s] = (valid_n && b < p.B) ? yc[(uint64_t)b * p.N] : make_float2(0.f, 0.f);
  }
}

// Column tiles [t0, t1) of group g of `slot`, with this wave's W slice from
// ws and tile t0's y operands already issued into ynext: GEMM, demod,
// LLR (and optional x_hat) stores. All waves of the workgroup must call it
// with the same tile range (barriers when nwk > 1).
template <int MT, int KS, int Q, class WSrc, bool PREFETCH = true>
__device__ __forceinline__ void eq_tiles(const EqParams& p, const WSrc& ws,
                                         uint32_t slot, uint32_t g,
                                         uint32_t mg, uint32_t wk, uint32_t t0,
                                         uint32_t t1, float* sRed,
                                         float2 (&ynext)[KS]) {
  const int lane = threadIdx.x & 63;
  const int q4 = lane >> 4;
  const int nw = blockDim.x >> 6;
  int buf = 0;

  for (uint32_t t = t0; t < t1; ++t) {
    float2 yv[KS];
    if (PREFETCH || t == t0) {
#pragma unroll
      for (int s = 0; s < KS; ++s) yv[s] = ynext[s];
    } else {
      eq_load_tile<KS>(p, slot, g, wk, t, yv);
    }
    if (PREFETCH && t + 1 < t1)
      eq_load_tile<KS>(p, slot, g, wk, t + 1, ynext);  // in flight during the MFMAs

    v4f acc1[MT], acc2[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      acc1[m] = v4f{0.f, 0.f, 0.f, 0.f};
      acc2[m] = v4f{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int s = 0; s < KS; ++s) {
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        float x1, x2;
        ws.get(m, s, x1, x2);
        acc1[m] = mfma4(x1, yv[s].x, acc1[m]);
        acc2[m] = mfma4(x2, yv[s].y, acc2[m]);
      }
    }

    const EqCol col = eq_col<Q>(p, slot, g, t);
    auto epilogue = [&](int m, v4f d) { eq_store<Q>(p, col, 8 * (mg * MT + m) + 2 * q4, d); };

    if (p.nwk == 1) {
#pragma unroll
      for (int m = 0; m < MT; ++m) epilogue(m, acc1[m] + acc2[m]);
    } else {
      float* mine = sRed + ((size_t)(buf * nw + wk) * MT) * 256;
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        const v4f d = acc1[m] + acc2[m];
#pragma unroll
        for (int r = 0; r < 4; ++r) mine[m * 256 + r * 64 + lane] = d[r];
      }
      __syncthreads();
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        if (m % (int)p.nwk != (int)wk) continue;
        v4f d = v4f{0.f, 0.f, 0.f, 0.f};
        for (uint32_t w2 = 0; w2 < p.nwk; ++w2) {
          const float* src = sRed + ((size_t)(buf * nw + w2) * MT + m) * 256;
#pragma unroll
          for (int r = 0; r < 4; ++r) d[r] += src[r * 64 + lane];
        }
        epilogue(m, d);
      }
      buf ^= 1;
    }
  }
}

// One column tile t of group g for users 8*MT*mg ... 8*MT*(mg+1)-1, by one
// wave over all B antennas, with W streamed from memory instead of held: for
// groups whose W slice does not fit a wave's registers. The antennas go in
// chunks of KC steps through a ring of R register sets: y (the HBM stream)
// and W (re-read by every tile of the group, so mostly L2 hits) are both
// issued R-1 chunks ahead of their use. No LDS and no barriers: each wave is
// independent.
//
// Requires B % (4*KC*R) == 0 (eq_stream_ok). Every load is then in bounds
// without a mask: columns past the group's end re-read its last column and
// users past U re-read user U-1 (their results are never stored), and the
// look-ahead past the last chunk re-reads that chunk. With no branches or
// selects around the loads the compiler waits for exactly the chunk in use.
template <int MT, int KC, int R, int NCH, int Q>
__device__ __forceinline__ void eq_unit_stream(const EqParams& p, uint32_t slot,
                                               uint32_t g, uint32_t t, uint32_t mg) {
  const int lane = threadIdx.x & 63;
  const int i16 = lane & 15, q4 = lane >> 4;
  const bool odd = i16 & 1;
  const uint32_t U = p.U, B = p.B, N = p.N;
  const uint32_t k0 = g * p.sg;
  const uint32_t nk = min(p.sg, N - k0);
  const uint32_t n = min(16 * t + i16, nk * p.S - 1);
  const uint32_t sym = group_sym(p, n, nk);
  const float2* ycol = p.y + slot * p.y_slot + sym * p.y_sym + k0 + (n - sym * nk) + q4 * N;
  const float2* W = p.csi + slot * p.csi_slot + (uint64_t)g * U * B;
  const float2* wrow[MT];
#pragma unroll
  for (int m = 0; m < MT; ++m)
    wrow[m] = W + min(8 * (mg * MT + m) + (i16 >> 1), U - 1) + (uint64_t)q4 * U;
  // NCH > 0: the chunk count is known at compile time and the chunk loop is
  // unrolled completely, so the wait before each chunk counts exactly the
  // loads issued after it (a loop back edge makes the compiler wait for all
  // but the newest loads, shortening the look-ahead to one chunk).
  const uint32_t nch = NCH > 0 ? (uint32_t)NCH : B / (4 * KC);

  auto load = [&](uint32_t c, float2 (&yd)[KC], float2 (&wd)[MT][KC]) {
    c = min(c, nch - 1);
#pragma unroll
    for (int s = 0; s < KC; ++s) yd[s] = ycol[(uint64_t)(4 * (c * KC + s)) * N];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int s = 0; s < KC; ++s) wd[m][s] = wrow[m][(uint64_t)(4 * (c * KC + s)) * U];
  };

  float2 yr[R][KC];
  float2 wr[R][MT][KC];
#pragma unroll
  for (int r = 0; r + 1 < R; ++r) load(r, yr[r], wr[r]);

  v4f acc1[MT], acc2[MT];
#pragma unroll
  for (int m = 0; m < MT; ++m) {
    acc1[m] = v4f{0.f, 0.f, 0.f, 0.f};
    acc2[m] = v4f{0.f, 0.f, 0.f, 0.f};
  }
  auto chunk = [&](uint32_t c, int r) {
    load(c + R - 1, yr[(r + R - 1) % R], wr[(r + R - 1) % R]);
    // keep the look-ahead loads ahead of this chunk's MFMAs (the scheduler
    // would otherwise sink them next to their uses)
    __builtin_amdgcn_sched_barrier(0);
    // W values -> the two A operands, in place and all at once: the MFMAs
    // then read registers no VALU instruction is waiting to overwrite.
    float2 a[MT][KC];
#pragma unroll
    for (int s = 0; s < KC; ++s)
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        const float2 v = wr[r][m][s];
        a[m][s] = make_float2(odd ? v.y : v.x, odd ? v.x : -v.y);
      }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int s = 0; s < KC; ++s)
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        acc1[m] = mfma4(a[m][s].x, yr[r][s].x, acc1[m]);
        acc2[m] = mfma4(a[m][s].y, yr[r][s].y, acc2[m]);
      }
  };
  if constexpr (NCH > 0) {
    static_assert(NCH % R == 0, "whole rings");
#pragma unroll
    for (int c = 0; c < NCH; ++c) chunk(c, c % R);
  } else {
    for (uint32_t c0 = 0; c0 < nch; c0 += R) {
#pragma unroll
      for (int r = 0; r < R; ++r) chunk(c0 + r, r);
    }
  }
  const EqCol col = eq_col<Q>(p, slot, g, t);
#pragma unroll
  for (int m = 0; m < MT; ++m) eq_store<Q>(p, col, 8 * (mg * MT + m) + 2 * q4, acc1[m] + acc2[m]);
}

// Shapes eq_unit_stream<MT, KC, R, Q> takes.
inline bool eq_stream_ok(uint32_t B, int KC, int R) { return B % (4 * KC * R) == 0; }

// Fills the launch-independent fields of EqParams from the C-ABI arguments
// (equalize.hip).
hipError_t fill_eq_params(const EqualizeArgs& a, EqParams& p, const char** why);

// Wide-cell FP32 form (equalize_wide.hip): antennas split over the four
// waves of a workgroup, W slices in registers, three-product complex MFMAs.
// Shapes it takes: up to 32 users, 33 to 256 antennas.
bool eq_wide_ok(uint32_t U, uint32_t B);
hipError_t launch_equalize_wide(const EqParams& p, uint32_t tc_override, uint32_t nw_override,
                                uint32_t persistent, hipStream_t st,
                                const char** why);

// Host-side tile plan shared by the equalize and detect launchers.
struct EqPlan {
  uint32_t MT, KS, nmg, nwk;
  size_t red_lds;  // bytes of the LDS reduction buffer (0 when nwk == 1)
};

inline uint32_t pow2ceil_u32(uint32_t v) {
  uint32_t r = 1;
  while (r < v) r <<= 1;
  return r;
}

// MT M-tiles (8 users each) per wave, KS 4-antenna steps per wave; the W slice
// per wave is 2*MT*KS registers (<= 64, so that two tiles of y operands, the
// accumulators and addresses fit without scratch at 2 waves per SIMD).
inline EqPlan plan_equalize(uint32_t U, uint32_t B, uint32_t ks_override) {
  EqPlan pl;
  pl.MT = U <= 8 ? 1 : U <= 16 ? 2 : 4;
  pl.nmg = (U + 8 * pl.MT - 1) / (8 * pl.MT);
  const uint32_t cap = pl.MT == 4 ? 8u : 16u;
  uint32_t KS = pow2ceil_u32((B + 3) / 4);
  KS = KS < 8 ? 8 : KS;
  KS = KS > cap ? cap : KS;
  if (ks_override == 8 || (ks_override == 16 && pl.MT <= 2)) KS = ks_override;
  pl.KS = KS;
  pl.nwk = (B + 4 * KS - 1) / (4 * KS);
  pl.red_lds = pl.nwk > 1 ? (size_t)2 * pl.nwk * pl.MT * 256 * sizeof(float) : 0;
  return pl;
}

}  // namespace detail
}  // namespace mega
