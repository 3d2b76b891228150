// Fused equalize + max-log demod: x_k = W_{k/sg} y_k, then binary16 LLRs.
//
// Specification: the naive kernel of src/operators/sgemv/equalize_kernel.cuh
// :415-480 (one dot product per (user, subcarrier), demod, store bit j while
// kQ + j < valid_bits), plus the tiled kernel's row bound (offset < ue_stride,
// :286,319). Every (user, subcarrier) is stored exactly once by the lane that
// computed it, so a subcarrier group that is not a multiple of a tile width
// never touches a neighbouring group's outputs.
//
// One workgroup per (slot, group, column chunk, user group): W_g is read once
// per workgroup into registers and applied to every (symbol, subcarrier)
// column of the chunk (the GEMM and its data movement are in eq_core.h). More
// than 8*MT users are split over adjacent workgroups (each reads the same y
// tile). Workgroups are numbered so that the column chunks and user groups of
// one subcarrier group land on one XCD (dispatch is round-robin over the 8
// XCDs), keeping re-reads of W_g and y in that XCD's L2.
#include <cstdlib>

#include "eq_core.h"

namespace mega {
namespace detail {
namespace {

template <int MT, int KS, int Q>
__global__ __launch_bounds__(64 * max_waves(MT, KS)) void equalize_mfma(const EqParams p) {
  extern __shared__ float sRed[];  // [2][nwk][MT][4][64] when nwk > 1
  const uint32_t wk = threadIdx.x >> 6;  // antenna split of this wave

  // logical = ((slot * G + g) * chunks + chunk) * nmg + mg
  const uint32_t logical = xcd_linear(blockIdx.x, gridDim.x);
  const uint32_t mg = logical % p.nmg;
  const uint32_t chunk = (logical / p.nmg) % p.chunks;
  const uint32_t rest = logical / p.nmg / p.chunks;
  const uint32_t g = rest % p.G;
  const uint32_t slot = rest / p.G;

  const uint32_t nk = min(p.sg, p.N - g * p.sg);
  const uint32_t ntiles = (nk * p.S + 15) / 16;
  const uint32_t t0 = chunk * p.tiles_per_wg;
  const uint32_t t1 = min(t0 + p.tiles_per_wg, ntiles);
  if (t0 >= t1) return;  // uniform over the workgroup

  const float2* W = p.csi + slot * p.csi_slot + (uint64_t)g * p.U * p.B;
  float2 ynext[KS];
  eq_load_tile<KS>(p, slot, g, wk, t0, ynext);
  WRegs<MT, KS> ws;
  ws.load(W, p.U, p.U, p.B, mg, wk);
  eq_tiles<MT, KS, Q>(p, ws, slot, g, mg, wk, t0, t1, sRed, ynext);
}

// Streamed-W form: one wave per (slot, group, column tile, user group), four
// independent waves per workgroup, units numbered so that a group's tiles
// share an XCD (and its L2 copy of W_g).
template <int MT, int KC, int R, int NCH, int Q>
__global__ __launch_bounds__(256) void equalize_stream(const EqParams p, uint32_t units) {
  const uint32_t unit = xcd_linear(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
  if (unit >= units) return;
  const uint32_t mg = unit % p.nmg;
  const uint32_t t = (unit / p.nmg) % p.col_tiles;
  const uint32_t rest = unit / p.nmg / p.col_tiles;
  const uint32_t g = rest % p.G, slot = rest / p.G;
  if (16 * t >= min(p.sg, p.N - g * p.sg) * p.S) return;  // past a partial group
  eq_unit_stream<MT, KC, R, NCH, Q>(p, slot, g, t, mg);
}

template <int MT, int KC, int R, int NCH>
hipError_t launch_stream_q(const EqParams& p, uint32_t units, uint32_t Q, hipStream_t st) {
  const dim3 grid((units + 3) / 4), block(256);
  switch (Q) {
    case 2: hipLaunchKernelGGL((equalize_stream<MT, KC, R, NCH, 2>), grid, block, 0, st, p, units); break;
    case 4: hipLaunchKernelGGL((equalize_stream<MT, KC, R, NCH, 4>), grid, block, 0, st, p, units); break;
    case 6: hipLaunchKernelGGL((equalize_stream<MT, KC, R, NCH, 6>), grid, block, 0, st, p, units); break;
    case 8: hipLaunchKernelGGL((equalize_stream<MT, KC, R, NCH, 8>), grid, block, 0, st, p, units); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Chunk shapes (KC antenna steps x ring of R chunks); all need B % 32 == 0.
// 0: 2x4 (default), 1: 4x2, 2: 1x8 (MEGA_EQ_SCH selects, for tuning sweeps).
// B = 64, 128, 256 get fully unrolled chunk loops (shape 0).
constexpr int kStreamKC[3] = {2, 4, 1};
constexpr int kStreamR[3] = {4, 2, 8};

template <int MT>
hipError_t launch_stream(const EqParams& p, uint32_t units, uint32_t Q, int shape,
                         hipStream_t st) {
  if (shape == 1) return launch_stream_q<MT, 4, 2, 0>(p, units, Q, st);
  if (shape == 2) return launch_stream_q<MT, 1, 8, 0>(p, units, Q, st);
  switch (p.B) {
    case 64: return launch_stream_q<MT, 2, 4, 8>(p, units, Q, st);
    case 128: return launch_stream_q<MT, 2, 4, 16>(p, units, Q, st);
    case 256: return launch_stream_q<MT, 2, 4, 32>(p, units, Q, st);
    default: return launch_stream_q<MT, 2, 4, 0>(p, units, Q, st);
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

}  // namespace

hipError_t fill_eq_params(const EqualizeArgs& a, EqParams& p, const char** why) {
  if ((uint64_t)a.N * a.Q > 0xffffffffull || a.cb_size > 0xffffffffull) {
    *why = "N*Q or cb_size exceeds 32 bits";
    return hipErrorInvalidValue;
  }
  p.csi = a.csi;
  p.y = a.y;
  p.llr = a.llr;
  p.xhat = a.xhat;
  p.w_out = nullptr;
  p.info = nullptr;
  p.U = a.U;
  p.B = a.B;
  p.N = a.N;
  p.sg = a.sg;
  p.S = a.S;
  p.G = a.G;
  p.slots = a.slots;
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
  p.cb_size32 = (uint32_t)a.cb_size;
  p.sg_inv = div_inv(a.sg);
  p.last_inv = div_inv(a.N - (a.G - 1) * a.sg);
  p.cb_inv = div_inv(p.cb_size32);
  p.Q = a.Q;
  p.store32 = ((reinterpret_cast<uintptr_t>(a.llr) & 3) == 0) &&
              (a.ue_stride % 2 == 0) && (a.cb_stride % 2 == 0) &&
              (a.zeros % 2 == 0) && (a.cb_size % 2 == 0) &&
              (a.llr_sym % 2 == 0) && (a.llr_slot % 2 == 0);
  p.col_tiles = (a.sg * a.S + 15) / 16;
  return hipSuccess;
}

hipError_t launch_equalize(const EqualizeArgs& a, hipStream_t st,
                           const char** why) {
  if (a.U == 0 || a.B == 0 || a.N == 0 || a.sg == 0 || a.S == 0 || a.slots == 0)
    return hipSuccess;
  EqParams p;
  hipError_t e = fill_eq_params(a, p, why);
  if (e != hipSuccess) return e;
  // Wide cells (17 to 64 users, 128 to 256 antennas, e.g. 256 x 32): the
  // antennas split over a workgroup's waves with W in registers
  // (equalize_wide.hip); MEGA_EQ_MODE=3 forces that form wherever it applies.
  const int mode = tune(kEqMode);
  if (eq_wide_ok(a.U, a.B) &&
      (mode == 3 || (mode == 0 && a.U > 16 && a.B >= 128)))
    return launch_equalize_wide(p, (uint32_t)tune(kEqWtc), (uint32_t)tune(kEqWnw),
                                (uint32_t)tune(kEqWp), st, why);
  // Tile plan; MEGA_EQ_KS=8|16 overrides the antenna step count for sweeps.
  const EqPlan pl = plan_equalize(a.U, a.B, (uint32_t)tune(kEqKs));
  // W held in registers while one wave covers all antennas; otherwise W is
  // streamed (MEGA_EQ_MODE=1 forces streaming, 2 forbids it).
  const int sh = tune(kEqSch);  // 0..2
  if (mode != 2 && (mode == 1 || pl.nwk > 1) && eq_stream_ok(a.B, kStreamKC[sh], kStreamR[sh])) {
    uint32_t MT = a.U <= 8 ? 1 : a.U <= 16 ? 2 : 4;
    const int smt = tune(kEqSmt);  // users per wave / 8 (sweeps)
    if (smt == 1 || smt == 2 || smt == 4) MT = (uint32_t)smt;
    p.nmg = (a.U + 8 * MT - 1) / (8 * MT);
    p.nwk = 1;
    p.chunks = 1;
    p.tiles_per_wg = p.col_tiles;
    const uint64_t units = (uint64_t)a.slots * a.G * p.col_tiles * p.nmg;
    if (units > 0xffffffffull - 3) {
      *why = "grid too large";
      return hipErrorInvalidValue;
    }
    switch (MT) {
      case 1: return launch_stream<1>(p, (uint32_t)units, a.Q, sh, st);
      case 2: return launch_stream<2>(p, (uint32_t)units, a.Q, sh, st);
      default: return launch_stream<4>(p, (uint32_t)units, a.Q, sh, st);
    }
  }
  if (pl.nwk > (uint32_t)max_waves(pl.MT, pl.KS)) {
    *why = "too many waves per workgroup for the register budget (U*B too large)";
    return hipErrorInvalidValue;
  }
  p.nmg = pl.nmg;
  p.nwk = pl.nwk;
  // Column tiles of a group, chunked over workgroups until the grid has at
  // least 2048 workgroups (8 per CU); MEGA_EQ_TPW overrides tiles per chunk.
  const uint64_t units = (uint64_t)a.slots * a.G;
  uint32_t chunks = 1;
  while (units * chunks < 2048 && chunks < p.col_tiles) ++chunks;
  p.tiles_per_wg = (p.col_tiles + chunks - 1) / chunks;
  const int tpw = tune(kEqTpw);
  if (tpw >= 1) p.tiles_per_wg = (uint32_t)tpw < p.col_tiles ? (uint32_t)tpw : p.col_tiles;
  p.chunks = (p.col_tiles + p.tiles_per_wg - 1) / p.tiles_per_wg;
  const uint64_t nblocks = units * p.chunks * pl.nmg;
  if (nblocks > 0x7fffffffull) {
    *why = "grid too large";
    return hipErrorInvalidValue;
  }
  dim3 grid((uint32_t)nblocks), block(64 * pl.nwk);
  switch (pl.MT * 100 + pl.KS) {
    case 108: return launch_q<1, 8>(p, grid, block, pl.red_lds, a.Q, st);
    case 116: return launch_q<1, 16>(p, grid, block, pl.red_lds, a.Q, st);
    case 208: return launch_q<2, 8>(p, grid, block, pl.red_lds, a.Q, st);
    case 216: return launch_q<2, 16>(p, grid, block, pl.red_lds, a.Q, st);
    case 408: return launch_q<4, 8>(p, grid, block, pl.red_lds, a.Q, st);
    default:
      *why = "no tile configuration";
      return hipErrorInvalidValue;
  }
}

}  // namespace detail
}  // namespace mega
