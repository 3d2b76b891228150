// Fused equalize + max-log demod: x_k = W_{k/sg} y_k, then binary16 LLRs.
//
// Specification: the naive kernel of src/operators/sgemv/equalize_kernel.cuh
// :415-480 (one dot product per (user, subcarrier), demod, store bit j while
// kQ + j < valid_bits), plus the tiled kernel's row bound (offset < ue_stride,
// :286,319). Every (user, subcarrier) is stored exactly once by the lane that
// computed it, so a subcarrier group that is not a multiple of a tile width
// never touches a neighbouring group's outputs.
//
// Per subcarrier group the work of a slot is one complex GEMM,
//   X (U x sg*S) = W_g (U x B) * Y (B x sg*S),
// computed on the FP32 matrix cores (v_mfma_f32_16x16x4_f32, FP32 products,
// FP32 accumulation). The complex product is split into two real MFMAs per
// step of 4 antennas with the output rows interleaved as (u re, u im):
//   MFMA1: A1[2u][k] = Re W[u,b], A1[2u+1][k] = Im W[u,b], B1[k][n] = Re y[b,n]
//   MFMA2: A2[2u][k] = -Im W[u,b], A2[2u+1][k] = Re W[u,b], B2[k][n] = Im y[b,n]
// so D[2u][n] = Re x, D[2u+1][n] = Im x and every lane ends up holding whole
// equalized symbols (users 2q, 2q+1 of its 16-row tile at column n) for the
// fused demod. A 16-row M tile covers 8 users; a 16-column N tile covers 16
// (symbol, subcarrier) columns of the group.
//
// Data movement: each wave keeps its slice of W_g in registers (loaded once
// per workgroup) and streams y straight from HBM (lane = column, 16 lanes read
// 128 contiguous bytes of an antenna row; one float2 feeds both MFMAs). When
// B is wider than one wave's K range, waves split the antennas and their
// partial sums are reduced through LDS (double-buffered, one barrier per tile).
// More than 8*MT users are split over adjacent workgroups (each reads the same
// y tile). Workgroups are numbered so that the column chunks and user groups
// of one subcarrier group land on one XCD (dispatch is round-robin over the 8
// XCDs), keeping re-reads of W_g and y in that XCD's L2.
#include "demod.h"
#include "internal.h"

namespace mega {
namespace detail {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

struct EqParams {
  const float2* csi;
  const float2* y;
  _Float16* llr;
  float2* xhat;
  uint32_t U, B, N, sg, S, G, slots;
  uint32_t nmg;              // workgroups per tile: groups of 8*MT users
  uint32_t nwk;              // waves per workgroup: antenna splits
  uint32_t col_tiles;        // 16-column tiles of a full group's sg*S columns
  uint32_t tiles_per_wg, chunks;
  uint64_t ue_stride, cb_size, cb_stride, zeros, valid_bits;
  uint64_t csi_slot, y_sym, y_slot, llr_sym, llr_slot;
  uint32_t store32;  // every LLR offset even and the base 4-byte aligned
};

template <int Q>
__device__ __forceinline__ void store_llrs(const EqParams& p, _Float16* row,
                                           uint64_t k, float re, float im) {
  _Float16 v[Q];
  demod::demod<Q>(re, im, v);
  const uint64_t bit = k * Q;
  const uint64_t off = (bit / p.cb_size) * p.cb_stride + p.zeros + bit % p.cb_size;
  _Float16* dst = row + off;
  if (p.store32 && bit + Q <= p.valid_bits && off + Q <= p.ue_stride) {
#pragma unroll
    for (int j = 0; j < Q / 2; ++j) {
      uint32_t lo = __builtin_bit_cast(uint16_t, v[2 * j]);
      uint32_t hi = __builtin_bit_cast(uint16_t, v[2 * j + 1]);
      reinterpret_cast<uint32_t*>(dst)[j] = lo | (hi << 16);
    }
  } else {
#pragma unroll
    for (int j = 0; j < Q; ++j)
      if (bit + j < p.valid_bits && off + j < p.ue_stride) dst[j] = v[j];
  }
}

// Maps the hardware workgroup id to a logical work index so that consecutive
// logical indices share an XCD (ids are dealt to XCDs round-robin).
__device__ __forceinline__ uint32_t xcd_linear(uint32_t w, uint32_t T) {
  const uint32_t q = T / 8, r = T % 8, x = w % 8, j = w / 8;
  return x * q + min(x, r) + j;
}

// Waves per workgroup the register budget allows for a (MT, KS) tile.
constexpr int max_waves(int MT, int KS) { return MT * KS <= 16 ? 16 : 8; }

template <int MT, int KS, int Q>
__global__ __launch_bounds__(64 * max_waves(MT, KS)) void equalize_mfma(const EqParams p) {
  extern __shared__ float sRed[];  // [2][nwk][MT][4][64] when nwk > 1
  const int lane = threadIdx.x & 63;
  const int wk = threadIdx.x >> 6;  // antenna split of this wave
  const int nw = blockDim.x >> 6;   // == nwk
  const int i16 = lane & 15, q4 = lane >> 4;

  // logical = ((slot * G + g) * chunks + chunk) * nmg + mg
  const uint32_t logical = xcd_linear(blockIdx.x, gridDim.x);
  const uint32_t mg = logical % p.nmg;
  const uint32_t chunk = (logical / p.nmg) % p.chunks;
  const uint32_t rest = logical / p.nmg / p.chunks;
  const uint32_t g = rest % p.G;
  const uint32_t slot = rest / p.G;

  const uint32_t k0 = g * p.sg;
  const uint32_t nk = min(p.sg, p.N - k0);
  const uint32_t ncols = nk * p.S;
  const uint32_t ntiles = (ncols + 15) / 16;
  const uint32_t t0 = chunk * p.tiles_per_wg;
  const uint32_t t1 = min(t0 + p.tiles_per_wg, ntiles);
  if (t0 >= t1) return;  // uniform over the workgroup

  const uint32_t U = p.U, B = p.B;

  // ---- W_g slice -> MFMA A operands (registers) ----------------------------
  const float2* W = p.csi + slot * p.csi_slot + (uint64_t)g * U * B;
  float wa1[MT][KS], wa2[MT][KS];
  const bool odd = i16 & 1;
#pragma unroll
  for (int m = 0; m < MT; ++m) {
    const uint32_t u = 8 * (mg * MT + m) + (i16 >> 1);
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const uint32_t b = 4 * (wk * KS + s) + q4;
      float2 v = make_float2(0.f, 0.f);
      if (u < U && b < B) v = W[u + (uint64_t)U * b];
      wa1[m][s] = odd ? v.y : v.x;
      wa2[m][s] = odd ? v.x : -v.y;
    }
  }

  const float2* Y = p.y + slot * p.y_slot + k0;
  _Float16* L = p.llr + slot * p.llr_slot;
  int buf = 0;

  for (uint32_t t = t0; t < t1; ++t) {
    const uint32_t n = 16 * t + i16;
    const bool valid_n = n < ncols;
    const uint32_t sym = valid_n ? n / nk : 0;
    const uint32_t kk = valid_n ? n - sym * nk : 0;
    const float2* yc = Y + sym * p.y_sym + kk;

    float2 yv[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const uint32_t b = 4 * (wk * KS + s) + q4;
      yv[s] = (valid_n && b < B) ? yc[(uint64_t)b * p.N] : make_float2(0.f, 0.f);
    }
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
        acc1[m] = __builtin_amdgcn_mfma_f32_16x16x4f32(wa1[m][s], yv[s].x, acc1[m], 0, 0, 0);
        acc2[m] = __builtin_amdgcn_mfma_f32_16x16x4f32(wa2[m][s], yv[s].y, acc2[m], 0, 0, 0);
      }
    }

    const uint64_t k = k0 + kk;
    _Float16* rows = L + sym * p.llr_sym;
    auto epilogue = [&](int m, v4f d) {
      if (!valid_n) return;
      const uint32_t u0 = 8 * (mg * MT + m) + 2 * q4;
      if (u0 < U) {
        store_llrs<Q>(p, rows + u0 * p.ue_stride, k, d[0], d[1]);
        if (p.xhat)
          p.xhat[(((uint64_t)slot * p.S + sym) * U + u0) * p.N + k] = make_float2(d[0], d[1]);
      }
      if (u0 + 1 < U) {
        store_llrs<Q>(p, rows + (u0 + 1) * p.ue_stride, k, d[2], d[3]);
        if (p.xhat)
          p.xhat[(((uint64_t)slot * p.S + sym) * U + u0 + 1) * p.N + k] = make_float2(d[2], d[3]);
      }
    };

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
        if (m % (int)p.nwk != wk) continue;
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

template <int MT, int KS, int Q>
hipError_t launch_one(const EqParams& p, dim3 grid, dim3 block, size_t lds,
                      hipStream_t st) {
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute(
        reinterpret_cast<const void*>(&equalize_mfma<MT, KS, Q>),
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((equalize_mfma<MT, KS, Q>), grid, block, lds, st, p);
  return hipGetLastError();
}

template <int MT, int KS>
hipError_t launch_q(const EqParams& p, dim3 grid, dim3 block, size_t lds,
                    uint32_t Q, hipStream_t st) {
  switch (Q) {
    case 2: return launch_one<MT, KS, 2>(p, grid, block, lds, st);
    case 4: return launch_one<MT, KS, 4>(p, grid, block, lds, st);
    case 6: return launch_one<MT, KS, 6>(p, grid, block, lds, st);
    case 8: return launch_one<MT, KS, 8>(p, grid, block, lds, st);
    default: return hipErrorInvalidValue;
  }
}

uint32_t pow2ceil(uint32_t v) {
  uint32_t r = 1;
  while (r < v) r <<= 1;
  return r;
}

}  // namespace

hipError_t launch_equalize(const EqualizeArgs& a, hipStream_t st,
                           const char** why) {
  if (a.U == 0 || a.B == 0 || a.N == 0 || a.sg == 0 || a.S == 0 || a.slots == 0)
    return hipSuccess;
  // Tile selection: MT M-tiles (8 users each) per wave, KS 4-antenna steps per
  // wave; W slice per wave = 2*MT*KS registers (<= 64, so that the y operands,
  // accumulators and addresses fit without scratch at 2 waves per SIMD).
  const uint32_t MT = a.U <= 8 ? 1 : a.U <= 16 ? 2 : 4;
  const uint32_t nmg = (a.U + 8 * MT - 1) / (8 * MT);
  uint32_t KS = pow2ceil((a.B + 3) / 4);
  KS = KS < 8 ? 8 : KS;
  KS = KS > 32 / MT ? 32 / MT : KS;
  const uint32_t nwk = (a.B + 4 * KS - 1) / (4 * KS);
  const uint32_t nw = nwk;
  EqParams p;
  p.csi = a.csi;
  p.y = a.y;
  p.llr = a.llr;
  p.xhat = a.xhat;
  p.U = a.U;
  p.B = a.B;
  p.N = a.N;
  p.sg = a.sg;
  p.S = a.S;
  p.G = a.G;
  p.slots = a.slots;
  p.nmg = nmg;
  p.nwk = nwk;
  p.ue_stride = a.ue_stride;
  p.cb_size = a.cb_size;
  p.cb_stride = a.cb_stride;
  p.zeros = a.zeros;
  p.valid_bits = a.valid_bits;
  p.csi_slot = a.csi_slot;
  p.y_sym = a.y_sym;
  p.y_slot = a.y_slot;
  p.llr_sym = a.llr_sym;
  p.llr_slot = a.llr_slot;
  p.store32 = ((reinterpret_cast<uintptr_t>(a.llr) & 3) == 0) &&
              (a.ue_stride % 2 == 0) && (a.cb_stride % 2 == 0) &&
              (a.zeros % 2 == 0) && (a.cb_size % 2 == 0) &&
              (a.llr_sym % 2 == 0) && (a.llr_slot % 2 == 0);
  // Column tiles of a group, chunked over workgroups until the grid has at
  // least 2048 workgroups (8 per CU).
  p.col_tiles = (a.sg * a.S + 15) / 16;
  const uint64_t units = (uint64_t)a.slots * a.G;
  uint32_t chunks = 1;
  while (units * chunks < 2048 && chunks < p.col_tiles) ++chunks;
  p.tiles_per_wg = (p.col_tiles + chunks - 1) / chunks;
  p.chunks = (p.col_tiles + p.tiles_per_wg - 1) / p.tiles_per_wg;
  const uint64_t nblocks = units * p.chunks * nmg;
  if (nblocks > 0x7fffffffull) {
    *why = "grid too large";
    return hipErrorInvalidValue;
  }
  if (nw > (uint32_t)max_waves(MT, KS)) {
    *why = "too many waves per workgroup for the register budget (U*B too large)";
    return hipErrorInvalidValue;
  }
  const size_t lds = nwk > 1 ? (size_t)2 * nw * MT * 256 * sizeof(float) : 0;
  dim3 grid((uint32_t)nblocks), block(64 * nw);
  switch (MT * 100 + KS) {
    case 108: return launch_q<1, 8>(p, grid, block, lds, a.Q, st);
    case 116: return launch_q<1, 16>(p, grid, block, lds, a.Q, st);
    case 132: return launch_q<1, 32>(p, grid, block, lds, a.Q, st);
    case 208: return launch_q<2, 8>(p, grid, block, lds, a.Q, st);
    case 216: return launch_q<2, 16>(p, grid, block, lds, a.Q, st);
    case 408: return launch_q<4, 8>(p, grid, block, lds, a.Q, st);
    default:
      *why = "no tile configuration";
      return hipErrorInvalidValue;
  }
}

}  // namespace detail
}  // namespace mega
