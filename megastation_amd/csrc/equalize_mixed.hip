// Equalize + demod of the mixed form: x = W y on the BF16 matrix cores with
// split operands (prec_core.h; FP32 accumulation), W read as the split BF16
// planes the mixed ZF kernel writes, y split on the fly. One wave per (slot,
// group, T 16-column tiles, group of 16*MT users), four independent waves per
// workgroup, as the streamed FP32 kernel (eq_core.h); each W load serves T
// column tiles. Demod and LLR placement are the shared epilogue.
#include <cstdlib>

#include "prec_core.h"

namespace mega {
namespace detail {
namespace {

// Rows of the A operand are 16 users with W's natural (re, im) words, so W
// is used as loaded; the complex product takes two accumulators per tile,
//   D_re = W * [Re y; -Im y],   D_im = W * [Im y; Re y]   (K = (antenna, part)),
// whose B operands are the y words with the high half negated or the halves
// swapped (shared by all user tiles). Lane l of D_re / D_im then holds users
// 4*(l/16) .. +3 of its tile at column l%16.
//
// Antenna steps of 16 go through a ring of R register sets: y (the HBM
// stream) and W (split planes, mostly L2 hits) both issued R-1 steps ahead
// (one depth for both: the load counter is in order, so a shallower W ring
// would cap the y prefetch at its depth). Loads past the last antenna are
// clamped (always in bounds) and their y is zeroed (the W planes hold zeros
// past B), so the ring needs no early exit. NSW > 0: the wave sums steps
// s_off .. s_off+NSW-1, a count known at compile time, with the loop unrolled
// completely and no loads past its last step; NSW = 0: all split_row(B)/16
// steps, counted at run time (loads past the last step re-read it).
template <int MT, int T, int R, int NSW>
__device__ __forceinline__ void eq_mixed_acc(const EqParams& p, const uint32_t* __restrict__ Ws,
                                             uint32_t slot, uint32_t g, uint32_t t0,
                                             uint32_t mg, uint32_t s_off, v4f (&are)[T][MT],
                                             v4f (&aim)[T][MT]) {
  const int lane = threadIdx.x & 63;
  const int i16 = lane & 15, q4 = lane >> 4;
  const uint32_t U = p.U, B = p.B, N = p.N;
  const uint32_t Bp = split_row(B);
  const uint32_t k0 = g * p.sg;
  const uint32_t nk = min(p.sg, N - k0);
  const float2* ycol[T];
#pragma unroll
  for (int c = 0; c < T; ++c) {
    // columns past the group re-read its last one (their results are dropped)
    const uint32_t n = min(16 * (t0 + c) + i16, nk * p.S - 1);
    const uint32_t sym = group_sym(p, n, nk);
    ycol[c] = p.y + slot * p.y_slot + sym * p.y_sym + k0 + (n - sym * nk);
  }
  const uint32_t* Wg = Ws + 2 * (slot * p.csi_slot + (uint64_t)g * U * Bp);
  const uint32_t* wrow[MT];
#pragma unroll
  for (int m = 0; m < MT; ++m)
    wrow[m] = Wg + (uint64_t)min(16 * (mg * MT + m) + i16, U - 1) * Bp + 4 * q4;
  const uint64_t lo_off = (uint64_t)U * Bp;
  const uint32_t nst = Bp / 16;

  auto load = [&](uint32_t st, float2 (&yd)[T][4], v4u (&wh)[MT], v4u (&wl)[MT]) {
    st = min(st, nst - 1);
#pragma unroll
    for (int c = 0; c < T; ++c)
#pragma unroll
      for (int a = 0; a < 4; ++a)
        yd[c][a] = ycol[c][(uint64_t)min(16 * st + 4 * q4 + a, B - 1) * N];
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      wh[m] = *reinterpret_cast<const v4u*>(wrow[m] + 16 * st);
      wl[m] = *reinterpret_cast<const v4u*>(wrow[m] + lo_off + 16 * st);
    }
  };

  float2 yr[R][T][4];
  v4u wh[R][MT], wl[R][MT];
#pragma unroll
  for (int r = 0; r + 1 < R; ++r)
    if (NSW == 0 || r < NSW) load(s_off + r, yr[r], wh[r], wl[r]);
#pragma unroll
  for (int c = 0; c < T; ++c)
#pragma unroll
    for (int m = 0; m < MT; ++m) are[c][m] = aim[c][m] = v4f{0.f, 0.f, 0.f, 0.f};

  // local step i of the wave (global step s_off + i), ring slot r; pf: the
  // load R-1 steps ahead is issued
  auto step = [&](uint32_t i, int r, bool pf) {
    const uint32_t st = s_off + i;
    if (pf) load(st + R - 1, yr[(r + R - 1) % R], wh[(r + R - 1) % R], wl[(r + R - 1) % R]);
    __builtin_amdgcn_sched_barrier(0);
    v4u b1h[T], b1l[T], b2h[T], b2l[T];
#pragma unroll
    for (int c = 0; c < T; ++c)
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const bool ok = 16 * st + 4 * q4 + a < B;
        const SplitC sp = split_c(ok ? yr[r][c][a] : make_float2(0.f, 0.f));
        b1h[c][a] = sp.hi ^ 0x80000000u;
        b1l[c][a] = sp.lo ^ 0x80000000u;
        b2h[c][a] = __builtin_rotateleft32(sp.hi, 16);
        b2l[c][a] = __builtin_rotateleft32(sp.lo, 16);
      }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int c = 0; c < T; ++c) {
        are[c][m] = mfma_x3(wh[r][m], wl[r][m], b1h[c], b1l[c], are[c][m]);
        aim[c][m] = mfma_x3(wh[r][m], wl[r][m], b2h[c], b2l[c], aim[c][m]);
      }
  };
  if constexpr (NSW > 0) {
    static_for<0, NSW>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      step(i, i % R, i + R - 1 < NSW);
    });
  } else {
    for (uint32_t s0 = 0; s0 < nst; s0 += R) {
#pragma unroll
      for (int r = 0; r < R; ++r) step(s0 + r, r, true);
    }
  }
}

// Demod and store of column tile t0+c, users of tile m of group mg.
template <int MT, int Q>
__device__ __forceinline__ void eq_mixed_store(const EqParams& p, uint32_t slot, uint32_t g,
                                               uint32_t t, uint32_t mg, int m, v4f re, v4f im) {
  const EqCol col = eq_col<Q>(p, slot, g, t);
  const uint32_t u0 = 16 * (mg * MT + m) + 4 * ((threadIdx.x & 63) >> 4);
  eq_store<Q>(p, col, u0, v4f{re[0], im[0], re[1], im[1]});
  eq_store<Q>(p, col, u0 + 2, v4f{re[2], im[2], re[3], im[3]});
}

template <int MT, int T, int R, int NST, int Q>
__device__ __forceinline__ void eq_unit_mixed(const EqParams& p, const uint32_t* __restrict__ Ws,
                                              uint32_t slot, uint32_t g, uint32_t t0,
                                              uint32_t mg) {
  v4f are[T][MT], aim[T][MT];
  eq_mixed_acc<MT, T, R, NST>(p, Ws, slot, g, t0, mg, 0, are, aim);
  const uint32_t ncol = min(p.sg, p.N - g * p.sg) * p.S;
#pragma unroll
  for (int c = 0; c < T; ++c) {
    if (16 * (t0 + c) >= ncol) break;  // uniform: past the group
#pragma unroll
    for (int m = 0; m < MT; ++m) eq_mixed_store<MT, Q>(p, slot, g, t0 + c, mg, m, are[c][m], aim[c][m]);
  }
}

// Decodes a unit index: (slot, group, T-tile unit t, user group mg).
struct MixedUnit {
  uint32_t slot, g, t, mg;
  __device__ __forceinline__ MixedUnit(const EqParams& p, uint32_t unit) {
    mg = unit % p.nmg;
    t = (unit / p.nmg) % p.col_tiles;
    const uint32_t rest = unit / p.nmg / p.col_tiles;
    g = rest % p.G;
    slot = rest / p.G;
  }
};

// One wave per (slot, group, T column tiles, user group); p.col_tiles counts
// units of T tiles here.
template <int MT, int T, int R, int NST, int Q>
__global__ __launch_bounds__(256) void equalize_mixed(const EqParams p, const uint32_t* Ws,
                                                      uint32_t units) {
  const uint32_t unit = xcd_linear(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
  if (unit >= units) return;
  const MixedUnit u(p, unit);
  if (16 * T * u.t >= min(p.sg, p.N - u.g * p.sg) * p.S) return;  // past a partial group
  eq_unit_mixed<MT, T, R, NST, Q>(p, Ws, u.slot, u.g, T * u.t, u.mg);
}

// Split-K form for grids too small to hide HBM latency (one slot of a wide
// cell: a few waves per SIMD, each with a long chain of antenna steps): KS
// waves of a workgroup share a unit, wave k summing antenna steps
// k*NST/KS .. (k+1)*NST/KS - 1, and their partial sums meet in LDS, where
// wave k adds up (in wave order) and stores the output tiles k, k+KS, ...
// of the unit's T*MT. 4/KS units per workgroup.
template <int MT, int T, int R, int NST, int KS, int Q>
__global__ __launch_bounds__(256) void equalize_mixed_ks(const EqParams p, const uint32_t* Ws,
                                                         uint32_t units) {
  static_assert(NST % KS == 0 && (KS == 2 || KS == 4), "whole steps per wave");
  constexpr int NSW = NST / KS, NP = T * MT;
  __shared__ v4f red[4][NP][2][64];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, kp = w % KS;
  const uint32_t unit = xcd_linear(blockIdx.x, gridDim.x) * (4 / KS) + w / KS;
  const MixedUnit u(p, unit < units ? unit : 0);
  const uint32_t ncol = min(p.sg, p.N - u.g * p.sg) * p.S;
  // a unit past the grid or a partial group skips its work but still meets
  // the barrier with the other waves of the workgroup
  const bool live = unit < units && 16 * T * u.t < ncol;
  if (live) {
    v4f are[T][MT], aim[T][MT];
    eq_mixed_acc<MT, T, R, NSW>(p, Ws, u.slot, u.g, T * u.t, u.mg, kp * NSW, are, aim);
#pragma unroll
    for (int c = 0; c < T; ++c)
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        red[w][c * MT + m][0][lane] = are[c][m];
        red[w][c * MT + m][1][lane] = aim[c][m];
      }
  }
  __syncthreads();
  if (!live) return;
  const int w0 = w - kp;
#pragma unroll
  for (int pi = kp; pi < NP; pi += KS) {
    const int c = pi / MT, m = pi % MT;
    if (16 * (T * u.t + c) >= ncol) continue;  // uniform: past the group
    v4f re = red[w0][pi][0][lane], im = red[w0][pi][1][lane];
#pragma unroll
    for (int j = 1; j < KS; ++j) {
      re += red[w0 + j][pi][0][lane];
      im += red[w0 + j][pi][1][lane];
    }
    eq_mixed_store<MT, Q>(p, u.slot, u.g, T * u.t + c, u.mg, m, re, im);
  }
}

// W-in-registers form for small cells (MT * NSW <= 8: W of the wave's users
// and all antennas is at most 64 registers): one wave per (slot, group,
// chunk of column tiles, group of 16*MT users) loads its W planes once and
// streams y for every (tile, antenna step) of the chunk through a ring of
// R register sets (R divides NSW, so every ring index is static; the
// prefetch runs across tile boundaries). Same products, order and epilogue
// as eq_unit_mixed.
template <int MT, int NSW, int R, int Q>
__device__ __forceinline__ void eq_unit_mixed_reg(const EqParams& p,
                                                  const uint32_t* __restrict__ Ws,
                                                  uint32_t slot, uint32_t g, uint32_t t0,
                                                  uint32_t t1, uint32_t mg) {
  static_assert(NSW % R == 0, "ring");
  const int lane = threadIdx.x & 63;
  const int i16 = lane & 15, q4 = lane >> 4;
  const uint32_t U = p.U, B = p.B, N = p.N;
  const uint32_t Bp = split_row(B);
  const uint32_t k0 = g * p.sg;
  const uint32_t nk = min(p.sg, N - k0);
  const uint32_t* Wg = Ws + 2 * (slot * p.csi_slot + (uint64_t)g * U * Bp);
  const uint64_t lo_off = (uint64_t)U * Bp;
  v4u wh[MT][NSW], wl[MT][NSW];
#pragma unroll
  for (int m = 0; m < MT; ++m) {
    const uint32_t* wrow = Wg + (uint64_t)min(16 * (mg * MT + m) + i16, U - 1) * Bp + 4 * q4;
#pragma unroll
    for (int st = 0; st < NSW; ++st) {
      wh[m][st] = *reinterpret_cast<const v4u*>(wrow + 16 * st);
      wl[m][st] = *reinterpret_cast<const v4u*>(wrow + lo_off + 16 * st);
    }
  }
  auto ycol = [&](uint32_t t) {  // column pointer of tile t (clamped into the chunk)
    const uint32_t n = min(16 * min(t, t1 - 1) + i16, nk * p.S - 1);
    const uint32_t sym = group_sym(p, n, nk);
    return p.y + slot * p.y_slot + sym * p.y_sym + k0 + (n - sym * nk);
  };
  float2 yr[R][4];
  auto load = [&](const float2* yc, int st, float2 (&yd)[4]) {
#pragma unroll
    for (int a = 0; a < 4; ++a) yd[a] = yc[(uint64_t)min(16 * st + 4 * q4 + a, (int)B - 1) * N];
  };
  const float2* ycur = ycol(t0);
  const float2* ynext = ycol(t0 + 1);
#pragma unroll
  for (int r = 0; r + 1 < R; ++r) load(ycur, r, yr[r]);
  for (uint32_t t = t0; t < t1; ++t) {
    v4f are[MT], aim[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m) are[m] = aim[m] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int st = 0; st < NSW; ++st) {
      const int ps = st + R - 1;  // step to prefetch (may be in the next tile)
      load(ps < NSW ? ycur : ynext, ps % NSW, yr[ps % R]);
      __builtin_amdgcn_sched_barrier(0);
      v4u b1h, b1l, b2h, b2l;
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const bool ok = 16 * st + 4 * q4 + a < (int)B;
        const SplitC sp = split_c(ok ? yr[st % R][a] : make_float2(0.f, 0.f));
        b1h[a] = sp.hi ^ 0x80000000u;
        b1l[a] = sp.lo ^ 0x80000000u;
        b2h[a] = __builtin_rotateleft32(sp.hi, 16);
        b2l[a] = __builtin_rotateleft32(sp.lo, 16);
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        are[m] = mfma_x3(wh[m][st], wl[m][st], b1h, b1l, are[m]);
        aim[m] = mfma_x3(wh[m][st], wl[m][st], b2h, b2l, aim[m]);
      }
    }
    ycur = ynext;
    ynext = ycol(t + 2);
    if (16 * t < nk * p.S) {  // uniform: a tile of the group
      const EqCol col = eq_col<Q>(p, slot, g, t);
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        const uint32_t u0 = 16 * (mg * MT + m) + 4 * q4;
        eq_store<Q>(p, col, u0, v4f{are[m][0], aim[m][0], are[m][1], aim[m][1]});
        eq_store<Q>(p, col, u0 + 2, v4f{are[m][2], aim[m][2], are[m][3], aim[m][3]});
      }
    }
  }
}

// units = slots * G * chunks * nmg, chunk c of a group = tiles
// [c * tiles_per_wg, min((c+1) * tiles_per_wg, col_tiles)).
template <int MT, int NSW, int R, int Q>
__global__ __launch_bounds__(256) void equalize_mixed_reg(const EqParams p, const uint32_t* Ws,
                                                          uint32_t units) {
  const uint32_t unit = xcd_linear(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
  if (unit >= units) return;
  const uint32_t mg = unit % p.nmg;
  const uint32_t c = (unit / p.nmg) % p.chunks;
  const uint32_t rest = unit / p.nmg / p.chunks;
  const uint32_t g = rest % p.G, slot = rest / p.G;
  const uint32_t ntiles = (min(p.sg, p.N - g * p.sg) * p.S + 15) / 16;
  const uint32_t t0 = c * p.tiles_per_wg, t1 = min(t0 + p.tiles_per_wg, ntiles);
  if (t0 >= t1) return;  // past a partial group
  eq_unit_mixed_reg<MT, NSW, R, Q>(p, Ws, slot, g, t0, t1, mg);
}

template <int MT, int NSW>
hipError_t launch_reg(const EqParams& p, const uint32_t* wp, uint32_t units, uint32_t Q,
                      hipStream_t st) {
  const dim3 grid((units + 3) / 4), block(256);
  switch (Q) {
    case 2: hipLaunchKernelGGL((equalize_mixed_reg<MT, NSW, 4, 2>), grid, block, 0, st, p, wp, units); break;
    case 4: hipLaunchKernelGGL((equalize_mixed_reg<MT, NSW, 4, 4>), grid, block, 0, st, p, wp, units); break;
    case 6: hipLaunchKernelGGL((equalize_mixed_reg<MT, NSW, 4, 6>), grid, block, 0, st, p, wp, units); break;
    case 8: hipLaunchKernelGGL((equalize_mixed_reg<MT, NSW, 4, 8>), grid, block, 0, st, p, wp, units); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int MT, int T, int R, int NST>
hipError_t launch_q(const EqParams& p, const uint32_t* wp, uint32_t units, uint32_t Q,
                    hipStream_t st) {
  const dim3 grid((units + 3) / 4), block(256);
  switch (Q) {
    case 2: hipLaunchKernelGGL((equalize_mixed<MT, T, R, NST, 2>), grid, block, 0, st, p, wp, units); break;
    case 4: hipLaunchKernelGGL((equalize_mixed<MT, T, R, NST, 4>), grid, block, 0, st, p, wp, units); break;
    case 6: hipLaunchKernelGGL((equalize_mixed<MT, T, R, NST, 6>), grid, block, 0, st, p, wp, units); break;
    case 8: hipLaunchKernelGGL((equalize_mixed<MT, T, R, NST, 8>), grid, block, 0, st, p, wp, units); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int MT, int T, int R, int NST, int KS>
hipError_t launch_ks_q(const EqParams& p, const uint32_t* wp, uint32_t units, uint32_t Q,
                       hipStream_t st) {
  constexpr int UPW = 4 / KS;
  const dim3 grid((units + UPW - 1) / UPW), block(256);
  switch (Q) {
    case 2: hipLaunchKernelGGL((equalize_mixed_ks<MT, T, R, NST, KS, 2>), grid, block, 0, st, p, wp, units); break;
    case 4: hipLaunchKernelGGL((equalize_mixed_ks<MT, T, R, NST, KS, 4>), grid, block, 0, st, p, wp, units); break;
    case 6: hipLaunchKernelGGL((equalize_mixed_ks<MT, T, R, NST, KS, 6>), grid, block, 0, st, p, wp, units); break;
    case 8: hipLaunchKernelGGL((equalize_mixed_ks<MT, T, R, NST, KS, 8>), grid, block, 0, st, p, wp, units); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Ring depth 4 (one 16-user tile per wave) or 3 (two tiles); B up to 256 in
// whole steps gets a fully unrolled step loop. ks = 2 or 4 takes the split-K
// form (split rows of 128 or 256 antennas).
template <int MT, int T>
hipError_t launch_mt(const EqParams& p, const uint32_t* wp, uint32_t units, uint32_t Q,
                     uint32_t ks, hipStream_t st) {
  constexpr int R = MT * T >= 2 ? 3 : 4;
  switch (split_row(p.B)) {
    case 64: return launch_q<MT, T, R, 4>(p, wp, units, Q, st);
    case 128:
      if (ks == 2) return launch_ks_q<MT, T, R, 8, 2>(p, wp, units, Q, st);
      return launch_q<MT, T, R, 8>(p, wp, units, Q, st);
    case 256:
      if (ks == 2) return launch_ks_q<MT, T, R, 16, 2>(p, wp, units, Q, st);
      if (ks == 4) return launch_ks_q<MT, T, R, 16, 4>(p, wp, units, Q, st);
      return launch_q<MT, T, R, 16>(p, wp, units, Q, st);
    default: return launch_q<MT, T, R, 0>(p, wp, units, Q, st);
  }
}

}  // namespace

// a.csi is ignored: the weights come from wp (split planes of mega_zf_split;
// a.csi_slot in 8-byte units, G*U*split_row(B) for a dense buffer).
hipError_t launch_equalize_mixed(const EqualizeArgs& a, const void* wp, hipStream_t st,
                                 const char** why) {
  if (a.U == 0 || a.B == 0 || a.N == 0 || a.sg == 0 || a.S == 0 || a.slots == 0)
    return hipSuccess;
  EqParams p;
  hipError_t e = fill_eq_params(a, p, why);
  if (e != hipSuccess) return e;
  // 16-user tiles per wave: 2 when there are more than 16 users (one y read
  // for 32 users), else 1; MEGA_EQ_XMT=1|2 overrides (tuning sweeps).
  uint32_t MT = a.U <= 16 ? 1 : 2;
  const int smt = tune(kEqXmt);
  if (smt == 1 || smt == 2) MT = (uint32_t)smt;
  // W in registers when the wave's W fits in 64 registers (split rows of at
  // most 128 antennas; MEGA_EQ_XREG=1 forces streaming): W is read once
  // per wave for a chunk of the group's tiles, chunks sized so the grid has
  // at least 4096 waves.
  const uint32_t nsw = split_row(a.B) / 16;
  const int xreg = tune(kEqXreg);
  if (xreg != 1 && (nsw == 4 || nsw == 8) && MT * nsw <= 8) {
    p.nmg = (a.U + 16 * MT - 1) / (16 * MT);
    p.nwk = 1;
    const uint64_t base = (uint64_t)a.slots * a.G * p.nmg;
    uint32_t chunks = 1;
    while (base * chunks < 4096 && chunks < p.col_tiles) ++chunks;
    p.tiles_per_wg = (p.col_tiles + chunks - 1) / chunks;
    p.chunks = (p.col_tiles + p.tiles_per_wg - 1) / p.tiles_per_wg;
    const uint64_t units = base * p.chunks;
    if (units > 0xffffffffull - 3) {
      *why = "grid too large";
      return hipErrorInvalidValue;
    }
    const uint32_t* w = static_cast<const uint32_t*>(wp);
    const uint32_t n = (uint32_t)units;
    if (MT == 1)
      return nsw == 4 ? launch_reg<1, 4>(p, w, n, a.Q, st) : launch_reg<1, 8>(p, w, n, a.Q, st);
    return launch_reg<2, 4>(p, w, n, a.Q, st);
  }
  // Column tiles per wave: 2 (one W stream for 32 columns) when a group has
  // more than one tile, else 1; MEGA_EQ_XT=1|2 overrides.
  uint32_t T = p.col_tiles > 1 ? 2 : 1;
  const int xt = tune(kEqXt);
  if (xt == 1 || xt == 2) T = (uint32_t)xt;
  p.nmg = (a.U + 16 * MT - 1) / (16 * MT);
  p.nwk = 1;
  p.chunks = 1;
  p.tiles_per_wg = p.col_tiles;
  p.col_tiles = (p.col_tiles + T - 1) / T;  // units of T tiles
  const uint64_t units = (uint64_t)a.slots * a.G * p.col_tiles * p.nmg;
  if (units > 0xffffffffull - 3) {
    *why = "grid too large";
    return hipErrorInvalidValue;
  }
  // Antenna split across the waves of a workgroup: two waves per unit for
  // 256-antenna rows when the grid is too small to hide HBM latency (under 8
  // waves per CU, e.g. one slot of a 256x32 cell: 30 us instead of 36), none
  // otherwise; MEGA_EQ_XKS=1|2|4 overrides.
  uint32_t ks = split_row(a.B) == 256 && units < 2048 ? 2 : 1;
  const int xks = tune(kEqXks);
  if (xks == 1 || xks == 2 || xks == 4) ks = (uint32_t)xks;
  const uint32_t* w = static_cast<const uint32_t*>(wp);
  const uint32_t n = (uint32_t)units;
  if (T == 2)
    return MT == 1 ? launch_mt<1, 2>(p, w, n, a.Q, ks, st) : launch_mt<2, 2>(p, w, n, a.Q, ks, st);
  return MT == 1 ? launch_mt<1, 1>(p, w, n, a.Q, ks, st) : launch_mt<2, 1>(p, w, n, a.Q, ks, st);
}

}  // namespace detail
}  // namespace mega
