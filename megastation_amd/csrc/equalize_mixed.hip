// Equalize + demod of the mixed form: x = W y on the BF16 matrix cores with
// split operands (prec_core.h; FP32 accumulation), W read as the split BF16
// words the mixed ZF kernel writes (Wp[u + U b] = {hi pair, lo pair}), y
// split on the fly. One wave per (slot, group, 16-column tile, group of 8*MT
// users), four independent waves per workgroup, as the streamed FP32 kernel
// (eq_core.h); demod and LLR placement are the shared epilogue.
#include "prec_core.h"

namespace mega {
namespace detail {
namespace {

// Antenna steps of 16: y (the HBM stream) in a ring of RY register sets,
// RY-1 steps ahead; W one step ahead. Loads past the last step or antenna
// are clamped (always in bounds) and their y is zeroed, so the ring needs
// no early exit.
template <int MT, int RY, int Q>
__device__ __forceinline__ void eq_unit_mixed(const EqParams& p, const uint2* __restrict__ Wp,
                                              uint32_t slot, uint32_t g, uint32_t t,
                                              uint32_t mg) {
  static_assert(RY % 2 == 0, "the W ring (2) must divide the y ring");
  const int lane = threadIdx.x & 63;
  const int i16 = lane & 15, q4 = lane >> 4;
  const bool h = i16 & 1;
  const uint32_t U = p.U, B = p.B, N = p.N;
  const uint32_t k0 = g * p.sg;
  const uint32_t nk = min(p.sg, N - k0);
  const uint32_t n = min(16 * t + i16, nk * p.S - 1);
  const uint32_t sym = n / nk;
  const float2* ycol = p.y + slot * p.y_slot + sym * p.y_sym + k0 + (n - sym * nk);
  const uint2* W = Wp + slot * p.csi_slot + (uint64_t)g * U * B;
  const uint2* wrow[MT];
#pragma unroll
  for (int m = 0; m < MT; ++m) wrow[m] = W + min(8 * (mg * MT + m) + (i16 >> 1), U - 1);
  const uint32_t nst = (B + 15) / 16;

  auto load_y = [&](uint32_t st, float2 (&d)[4]) {
    st = min(st, nst - 1);
#pragma unroll
    for (int a = 0; a < 4; ++a) d[a] = ycol[(uint64_t)min(16 * st + 4 * q4 + a, B - 1) * N];
  };
  auto load_w = [&](uint32_t st, uint2 (&d)[MT][4]) {
    st = min(st, nst - 1);
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int a = 0; a < 4; ++a) d[m][a] = wrow[m][(uint64_t)U * min(16 * st + 4 * q4 + a, B - 1)];
  };

  float2 yr[RY][4];
  uint2 wr[2][MT][4];
#pragma unroll
  for (int r = 0; r + 1 < RY; ++r) load_y(r, yr[r]);
  load_w(0, wr[0]);
  v4f acc[MT];
#pragma unroll
  for (int m = 0; m < MT; ++m) acc[m] = v4f{0.f, 0.f, 0.f, 0.f};

  for (uint32_t s0 = 0; s0 < nst; s0 += RY) {
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      const uint32_t st = s0 + r;
      load_y(st + RY - 1, yr[(r + RY - 1) % RY]);
      load_w(st + 1, wr[(r + 1) % 2]);
      __builtin_amdgcn_sched_barrier(0);
      v4u yh, yl;
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const bool ok = 16 * st + 4 * q4 + a < B;
        const SplitC sp = split_c(ok ? yr[r][a] : make_float2(0.f, 0.f));
        yh[a] = sp.hi;
        yl[a] = sp.lo;
      }
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        v4u ah, al;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          ah[a] = a_word(wr[r % 2][m][a].x, h);
          al[a] = a_word(wr[r % 2][m][a].y, h);
        }
        acc[m] = mfma_x3(ah, al, yh, yl, acc[m]);
      }
    }
  }
  const EqCol col = eq_col<Q>(p, slot, g, t);
#pragma unroll
  for (int m = 0; m < MT; ++m) eq_store<Q>(p, col, 8 * (mg * MT + m) + 2 * q4, acc[m]);
}

template <int MT, int Q>
__global__ __launch_bounds__(256) void equalize_mixed(const EqParams p, const uint2* Wp,
                                                      uint32_t units) {
  const uint32_t unit = xcd_linear(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
  if (unit >= units) return;
  const uint32_t mg = unit % p.nmg;
  const uint32_t t = (unit / p.nmg) % p.col_tiles;
  const uint32_t rest = unit / p.nmg / p.col_tiles;
  const uint32_t g = rest % p.G, slot = rest / p.G;
  if (16 * t >= min(p.sg, p.N - g * p.sg) * p.S) return;  // past a partial group
  eq_unit_mixed<MT, 4, Q>(p, Wp, slot, g, t, mg);
}

template <int MT>
hipError_t launch_mt(const EqParams& p, const uint2* wp, uint32_t units, uint32_t Q,
                     hipStream_t st) {
  const dim3 grid((units + 3) / 4), block(256);
  switch (Q) {
    case 2: hipLaunchKernelGGL((equalize_mixed<MT, 2>), grid, block, 0, st, p, wp, units); break;
    case 4: hipLaunchKernelGGL((equalize_mixed<MT, 4>), grid, block, 0, st, p, wp, units); break;
    case 6: hipLaunchKernelGGL((equalize_mixed<MT, 6>), grid, block, 0, st, p, wp, units); break;
    case 8: hipLaunchKernelGGL((equalize_mixed<MT, 8>), grid, block, 0, st, p, wp, units); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace

// a.csi is ignored: the weights come from wp (split words, CSI indexing and
// a.csi_slot as the slot stride).
hipError_t launch_equalize_mixed(const EqualizeArgs& a, const void* wp, hipStream_t st,
                                 const char** why) {
  if (a.U == 0 || a.B == 0 || a.N == 0 || a.sg == 0 || a.S == 0 || a.slots == 0)
    return hipSuccess;
  EqParams p;
  hipError_t e = fill_eq_params(a, p, why);
  if (e != hipSuccess) return e;
  const uint32_t MT = a.U <= 8 ? 1 : a.U <= 16 ? 2 : 4;
  p.nmg = (a.U + 8 * MT - 1) / (8 * MT);
  p.nwk = 1;
  p.chunks = 1;
  p.tiles_per_wg = p.col_tiles;
  const uint64_t units = (uint64_t)a.slots * a.G * p.col_tiles * p.nmg;
  if (units > 0xffffffffull - 3) {
    *why = "grid too large";
    return hipErrorInvalidValue;
  }
  const uint2* w = static_cast<const uint2*>(wp);
  switch (MT) {
    case 1: return launch_mt<1>(p, w, (uint32_t)units, a.Q, st);
    case 2: return launch_mt<2>(p, w, (uint32_t)units, a.Q, st);
    default: return launch_mt<4>(p, w, (uint32_t)units, a.Q, st);
  }
}

}  // namespace detail
}  // namespace mega
