// FP32 equalize + demod for wide cells (up to 64 users, 128 or 256
// antennas, e.g. 256 x 32): the antennas are split over the four or eight
// waves of a workgroup, each wave keeps its slice of W_g in registers for
// the whole unit (16, 32 or 64 users: one, two or four 16-user tiles), and
// the complex products take three real MFMAs instead of four.
//
// Unit = (slot, group, chunk of TC column tiles of 16 (symbol, subcarrier)
// columns). Wave w owns antennas 4*KS*w .. 4*KS*(w+1)-1 (KS steps of 4) and
// computes, for every column tile of the unit, the partial products over its
// antennas for all 16*MT users; the four partial tiles meet in LDS (double
// buffered, one barrier per tile) and each wave adds up (in wave order),
// demodulates and stores a quarter of the tile's outputs.
//
// Products (Gauss): with W = Wr + i Wi and y = yr + i yi, per 4-antenna step
// and 16-user x 16-column tile
//   K1 += (Wr + Wi) yr,   K2 += Wr (yi - yr),   K3 += Wi (yr + yi)
// on v_mfma_f32_16x16x4_f32 (FP32 products, FP32 accumulation), and at the
// end Re x = K1 - K3, Im x = K1 + K2. Three MFMAs per step where the
// interleaved form (eq_core.h) takes four: the FP32 form at 256 x 32 is bound
// by the matrix cores' issue rate as much as by HBM. Rounding: every term is
// an FP32 product of sums of two FP32 values, so the error bound is that of
// the four-product form with |W| and |y| replaced by |Wr| + |Wi| and
// |yr| + |yi| (at most twice as large); parity is judged against the FP64
// oracle like every FP32 form.
//
// Operand layout of the 16x16x4 instruction: lane l supplies A[l%16][l/16]
// (user 16*mt + l%16, antenna 4*s + l/16 of the wave's slice) and
// B[l/16][l%16] (that antenna, column l%16 of the tile); the result lane l
// holds rows 4*(l/16) + r, column l%16.
#include "eq_core.h"

namespace mega {
namespace detail {
namespace {

template <int MT, int KS, int NW, int Q>
__global__ __launch_bounds__(64 * NW, NW == 8 && MT <= 2 ? 4 : 2) void equalize_wide(const EqParams p,
                                                                           uint32_t chunks,
                                                            uint32_t tc, uint32_t units) {
  // partial x tiles: [buffer][wave][16-user tile][re, im][row r][lane]
  __shared__ float red[2][NW][MT][2][4][64];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i16 = lane & 15, q4 = lane >> 4;
  const uint32_t unit = xcd_linear(blockIdx.x, gridDim.x);
  if (unit >= units) return;  // uniform over the workgroup
  const uint32_t c = unit % chunks;
  const uint32_t rest = unit / chunks;
  const uint32_t g = rest % p.G, slot = rest / p.G;
  const uint32_t U = p.U, B = p.B, N = p.N;
  const uint32_t k0 = g * p.sg;
  const uint32_t nk = min(p.sg, N - k0);
  const uint32_t ncol = nk * p.S;
  const uint32_t t0 = c * tc;
  const uint32_t t1 = min(t0 + tc, (ncol + 15) / 16);
  if (t0 >= t1) return;  // a chunk past a partial group: uniform

  // ---- this wave's W slice: Wr, Wi per (user tile, step); Wr + Wi is formed
  // at each use (one add per MFMA triple saves a third of the slice's
  // registers) -----------------------------------------------------------------
  const float2* W = p.csi + slot * p.csi_slot + (uint64_t)g * U * B;
  const uint32_t bw = 4 * KS * w;  // first antenna of the slice
  float wre[MT][KS], wim[MT][KS];
#pragma unroll
  for (int m = 0; m < MT; ++m) {
    const uint32_t u = 16 * m + i16;
    const float2* wrow = W + min(u, U - 1);
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const uint32_t b = bw + 4 * s + q4;
      const float2 v = wrow[(uint64_t)min(b, B - 1) * U];
      const bool ok = u < U && b < B;
      wre[m][s] = ok ? v.x : 0.f;
      wim[m][s] = ok ? v.y : 0.f;
    }
  }

  // ---- y of a column tile: this lane's column, the slice's antennas --------
  // (columns past the group's end and antennas past B re-read valid
  // addresses; their products are discarded or multiplied by W = 0)
  auto load_y = [&](uint32_t t, float2 (&yd)[KS]) {
    const uint32_t n = min(16 * t + i16, ncol - 1);
    const uint32_t sym = group_sym(p, n, nk);
    const float2* yc = p.y + slot * p.y_slot + sym * p.y_sym + k0 + (n - sym * nk);
#pragma unroll
    for (int s = 0; s < KS; ++s) yd[s] = yc[(uint64_t)min(bw + 4 * s + q4, B - 1) * N];
  };

  // outputs of column tile t from partial buffer b: wave w adds up, demods and
  // stores outputs (m, r) = pi / 4, pi % 4 for pi = w, w + NW, ...
  auto finish = [&](uint32_t t, int b) {
    const EqCol col = eq_col<Q>(p, slot, g, t);
#pragma unroll
    for (int pi = w; pi < 4 * MT; pi += NW) {
      const int m = pi / 4, r = pi % 4;
      float re = red[b][0][m][0][r][lane], im = red[b][0][m][1][r][lane];
#pragma unroll
      for (int j = 1; j < NW; ++j) {
        re += red[b][j][m][0][r][lane];
        im += red[b][j][m][1][r][lane];
      }
      const uint32_t u = 16 * m + 4 * q4 + r;
      if (col.valid && u < U) {
        store_llrs<Q>(p, col.rows + u * p.ue_stride, col.ls, re, im);
        if (col.xrow) col.xrow[(uint64_t)u * N] = make_float2(re, im);
      }
    }
  };

  // Software pipeline: iteration t issues tile t + 1's y loads, then tile
  // t - 1's outputs, then tile t's MFMAs. Loads and stores share one
  // completion counter, so the wait for y at the head of an iteration also
  // waits for every store issued before it; issued ahead of the MFMAs, the
  // stores complete while the matrix cores work instead of stalling the next
  // tile.
  float2 ycur[KS], ynext[KS];
  load_y(t0, ynext);
  int buf = 0;
  for (uint32_t t = t0; t < t1; ++t) {
#pragma unroll
    for (int s = 0; s < KS; ++s) ycur[s] = ynext[s];
    if (t + 1 < t1) load_y(t + 1, ynext);  // in flight during this tile's MFMAs
    if (t > t0) finish(t - 1, buf ^ 1);
    __builtin_amdgcn_sched_barrier(0);
    v4f k1[MT], k2[MT], k3[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m) k1[m] = k2[m] = k3[m] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const float yr = ycur[s].x;
      const float yd = ycur[s].y - ycur[s].x;
      const float yp = ycur[s].x + ycur[s].y;
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        k1[m] = mfma4(wre[m][s] + wim[m][s], yr, k1[m]);
        k2[m] = mfma4(wre[m][s], yd, k2[m]);
        k3[m] = mfma4(wim[m][s], yp, k3[m]);
      }
    }
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        red[buf][w][m][0][r][lane] = k1[m][r] - k3[m][r];
        red[buf][w][m][1][r][lane] = k1[m][r] + k2[m][r];
      }
    // after this barrier buffer buf holds tile t's partials; buffer buf ^ 1
    // (tile t - 1's, read above) is free for tile t + 1's
    __syncthreads();
    buf ^= 1;
  }
  finish(t1 - 1, buf ^ 1);
}

// Persistent form of the same arithmetic: P workgroups, workgroup i sweeping
// the contiguous range [i T / P, (i + 1) T / P) of the T = slots * G *
// col_tiles column tiles in (slot, group, tile) order, so every workgroup
// gets the same number of tiles to within one and reloads its W slices only
// where its range crosses into the next group (at most about two groups per
// workgroup, against one reload per unit of tc tiles above). The next tile's
// y, of whatever group, is in flight during the current tile's MFMAs. A
// tile's arithmetic is the unit kernel's, so the results are the same bits.
template <int MT, int KS, int NW, int Q>
__global__ __launch_bounds__(64 * NW, NW == 8 ? 4 : 2) void equalize_wide_p(const EqParams p,
                                                                             uint32_t total) {
  __shared__ float red[2][NW][MT][2][4][64];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i16 = lane & 15, q4 = lane >> 4;
  const uint32_t wg = xcd_linear(blockIdx.x, gridDim.x);
  const uint32_t r0 = (uint32_t)((uint64_t)wg * total / gridDim.x);
  const uint32_t r1 = (uint32_t)((uint64_t)(wg + 1) * total / gridDim.x);
  const uint32_t U = p.U, B = p.B, N = p.N, CT = p.col_tiles;
  const uint32_t bw = 4 * KS * w;  // first antenna of the wave's slice
  // (slot, group, tile) of linear tile index i; tiles past a partial group's
  // end have no columns (n0 >= ncol) and are skipped
  struct Where {
    uint32_t slot, g, t, k0, nk, ncol;
  };
  auto where = [&](uint32_t i) {
    Where x;
    x.t = i % CT;
    const uint32_t sg_ = i / CT;
    x.g = sg_ % p.G;
    x.slot = sg_ / p.G;
    x.k0 = x.g * p.sg;
    x.nk = min(p.sg, N - x.k0);
    x.ncol = x.nk * p.S;
    return x;
  };
  float wre[MT][KS], wim[MT][KS];
  auto load_w = [&](const Where& x) {
    const float2* Wg = p.csi + x.slot * p.csi_slot + (uint64_t)x.g * U * B;
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      const uint32_t u = 16 * m + i16;
      const float2* wrow = Wg + min(u, U - 1);
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        const uint32_t b = bw + 4 * s + q4;
        const float2 v = wrow[(uint64_t)min(b, B - 1) * U];
        const bool ok = u < U && b < B;
        wre[m][s] = ok ? v.x : 0.f;
        wim[m][s] = ok ? v.y : 0.f;
      }
    }
  };
  auto load_y = [&](const Where& x, float2 (&yd)[KS]) {
    const uint32_t n = min(16 * x.t + i16, x.ncol - 1);
    const uint32_t sym = group_sym(p, n, x.nk);
    const float2* yc = p.y + x.slot * p.y_slot + sym * p.y_sym + x.k0 + (n - sym * x.nk);
#pragma unroll
    for (int s = 0; s < KS; ++s) yd[s] = yc[(uint64_t)min(bw + 4 * s + q4, B - 1) * N];
  };
  if (r0 >= r1) return;  // uniform over the workgroup
  Where cur = where(r0);
  uint32_t cur_key = cur.slot * p.G + cur.g;
  load_w(cur);
  float2 ycur[KS], ynext[KS];
  load_y(cur, ynext);
  int buf = 0;
  for (uint32_t i = r0; i < r1; ++i) {
#pragma unroll
    for (int s = 0; s < KS; ++s) ycur[s] = ynext[s];
    const Where nx = where(i + 1 < r1 ? i + 1 : i);
    if (i + 1 < r1) load_y(nx, ynext);  // in flight during this tile's MFMAs
    __builtin_amdgcn_sched_barrier(0);
    v4f k1[MT], k2[MT], k3[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m) k1[m] = k2[m] = k3[m] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const float yr = ycur[s].x;
      const float yd = ycur[s].y - ycur[s].x;
      const float yp = ycur[s].x + ycur[s].y;
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        k1[m] = mfma4(wre[m][s] + wim[m][s], yr, k1[m]);
        k2[m] = mfma4(wre[m][s], yd, k2[m]);
        k3[m] = mfma4(wim[m][s], yp, k3[m]);
      }
    }
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        red[buf][w][m][0][r][lane] = k1[m][r] - k3[m][r];
        red[buf][w][m][1][r][lane] = k1[m][r] + k2[m][r];
      }
    __syncthreads();
    const uint32_t nkey = nx.slot * p.G + nx.g;
    if (i + 1 < r1 && nkey != cur_key) load_w(nx);  // the MFMAs above have read the slice
    if (16 * cur.t < cur.ncol) {  // uniform: a tile with columns
      const EqCol col = eq_col<Q>(p, cur.slot, cur.g, cur.t);
#pragma unroll
      for (int pi = w; pi < 4 * MT; pi += NW) {
        const int m = pi / 4, r = pi % 4;
        float re = red[buf][0][m][0][r][lane], im = red[buf][0][m][1][r][lane];
#pragma unroll
        for (int j = 1; j < NW; ++j) {
          re += red[buf][j][m][0][r][lane];
          im += red[buf][j][m][1][r][lane];
        }
        const uint32_t u = 16 * m + 4 * q4 + r;
        if (col.valid && u < U) {
          store_llrs<Q>(p, col.rows + u * p.ue_stride, col.ls, re, im);
          if (col.xrow) col.xrow[(uint64_t)u * N] = make_float2(re, im);
        }
      }
    }
    cur = nx;
    cur_key = nkey;
    buf ^= 1;
  }
}

template <int MT, int KS, int NW>
hipError_t launch_p(const EqParams& p, uint32_t wgs, uint32_t total, uint32_t Q,
                    hipStream_t st) {
  const dim3 grid(wgs), block(64 * NW);
  switch (Q) {
    case 2: hipLaunchKernelGGL((equalize_wide_p<MT, KS, NW, 2>), grid, block, 0, st, p, total); break;
    case 4: hipLaunchKernelGGL((equalize_wide_p<MT, KS, NW, 4>), grid, block, 0, st, p, total); break;
    case 6: hipLaunchKernelGGL((equalize_wide_p<MT, KS, NW, 6>), grid, block, 0, st, p, total); break;
    case 8: hipLaunchKernelGGL((equalize_wide_p<MT, KS, NW, 8>), grid, block, 0, st, p, total); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int MT, int KS, int NW>
hipError_t launch_q(const EqParams& p, uint32_t chunks, uint32_t tc, uint32_t units,
                    uint32_t Q, hipStream_t st) {
  const dim3 grid(units), block(64 * NW);
  switch (Q) {
    case 2: hipLaunchKernelGGL((equalize_wide<MT, KS, NW, 2>), grid, block, 0, st, p, chunks, tc, units); break;
    case 4: hipLaunchKernelGGL((equalize_wide<MT, KS, NW, 4>), grid, block, 0, st, p, chunks, tc, units); break;
    case 6: hipLaunchKernelGGL((equalize_wide<MT, KS, NW, 6>), grid, block, 0, st, p, chunks, tc, units); break;
    case 8: hipLaunchKernelGGL((equalize_wide<MT, KS, NW, 8>), grid, block, 0, st, p, chunks, tc, units); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace

bool eq_wide_ok(uint32_t U, uint32_t B) { return U >= 1 && U <= 64 && B > 32 && B <= 256; }

// Column tiles per workgroup: the most that still give every CU five
// workgroups (1280 in the grid), so each wave's W slice serves as many tiles
// as the grid's size allows; tc_override (>= 1) replaces it for sweeps.
// Measured at 256 x 32 (273 groups of 10 tiles), four waves: 2 tiles (1365
// workgroups) 48.9 us, 1 tile 50.2, 3 tiles 54.4, 4 tiles 51.5, 10 tiles
// 61.8; eight waves: 44.8, 45.6, 48.9, 48.6 (5 tiles), 56.0; the persistent
// form (persistent = k: k * 256 workgroups over balanced tile ranges) 45.8
// (k = 1), 46.1 (2), 49.4 (4), 52.7 (8). A wall-clock probe of the unit
// kernel at 2 tiles (four waves): each workgroup's first tile takes 9.4 us
// (W and y loads in flight, about 1.3 us of MFMAs), later tiles 2.8 us, the
// grid runs in about three rounds; without the y loads the launch takes
// 30 us, without the W loads 40, without the LLR stores 38. Probes of the
// eight-wave kernel at 2 tiles (event times, each including about 4 us of
// launch): y loads alone 21.9 us, y loads and MFMAs 31.2, with the W loads
// 34.1, everything 45.1; MFMAs alone (no loads) 23.0 at a 2.3 GHz shader
// clock (1.9 GHz under the whole kernel). Within a tile, FP32 MFMAs and VALU
// work serialize on a SIMD: a wave's VALU work waits out the MFMA burst of
// another wave on its SIMD (the FP32 matrix rate equals the FP32 vector
// rate); in aggregate no unit is saturated (matrix cores busy 0.28 of the
// SIMD-cycles, VALU 0.08), the time being each tile's chain of loads, MFMAs,
// reduction, demod and stores. Two schedules that overlap more of that chain
// measured no faster and were dropped: one 16-wave workgroup per CU
// with y three tiles ahead and the next group's W prefetched (48.0 us), and
// a split into 8 multiplying waves (y three tiles ahead, no other global
// access) and 4 output waves that reduce, demodulate, store and copy the
// next group's W into LDS (47.7 us, bit-identical; instrumented, a tile took
// about 5700 shader cycles of which 2700 were the multiplying waves' MFMAs,
// the rest the output waves' VALU work on the same SIMDs).
hipError_t launch_equalize_wide(const EqParams& p0, uint32_t tc_override, uint32_t nw_override,
                                uint32_t persistent, hipStream_t st, const char** why) {
  EqParams p = p0;
  const uint32_t MT = p.U <= 16 ? 1 : p.U <= 32 ? 2 : 4;
  // waves per workgroup: 4 (antennas 32 or 64 per wave) or 8 (16 or 32);
  // eight by default for rows of more than 128 antennas (four waves per SIMD
  // at 126 VGPRs against two at 234: 44.8 against 49.2 us at 256 x 32)
  const uint32_t NW = nw_override == 8 || persistent >= 1 ? 8
                      : nw_override == 4                  ? 4
                      : p.B > 128                         ? 8
                                                          : 4;
  const uint32_t KS = (p.B <= 128 ? 8 : 16) * 4 / NW;
  if (persistent >= 1 && MT <= 2) {
    // persistent[k]: k * 256 workgroups (k = 2: two per CU at four waves per
    // SIMD), never more than the tiles
    const uint64_t total = (uint64_t)p.slots * p.G * p.col_tiles;
    if (total > 0xffffffffull) {
      *why = "grid too large";
      return hipErrorInvalidValue;
    }
    const uint32_t wgs = (uint32_t)(total < 256ull * persistent ? total : 256ull * persistent);
    if (MT == 1) return KS == 4 ? launch_p<1, 4, 8>(p, wgs, (uint32_t)total, p.Q, st)
                                : launch_p<1, 8, 8>(p, wgs, (uint32_t)total, p.Q, st);
    return KS == 4 ? launch_p<2, 4, 8>(p, wgs, (uint32_t)total, p.Q, st)
                   : launch_p<2, 8, 8>(p, wgs, (uint32_t)total, p.Q, st);
  }
  const uint64_t groups = (uint64_t)p.slots * p.G;
  // 1280 units: five workgroups per CU; 33 to 64 users (one workgroup per
  // CU, 128 KB of partials) 640. Measured with four user tiles at 256 x 64
  // (273 groups of 10 tiles): 4 tiles a unit 82.8 us, 2 85.2, 1 96.0, 10
  // 108.6, the streamed plan 87.6; at 256 x 48 79.8 against 83.1, at
  // 128 x 64 49.5 against 53.7.
  const uint64_t min_units = MT == 4 ? 640 : 1280;
  uint32_t tc = 1;
  while (tc < p.col_tiles && groups * ((p.col_tiles + tc) / (tc + 1)) >= min_units) ++tc;
  if (tc_override >= 1) tc = tc_override < p.col_tiles ? tc_override : p.col_tiles;
  const uint32_t chunks = (p.col_tiles + tc - 1) / tc;
  const uint64_t units = groups * chunks;
  if (units > 0x7fffffffull) {
    *why = "grid too large";
    return hipErrorInvalidValue;
  }
  const uint32_t n = (uint32_t)units;
  if (MT == 4)  // 33 to 64 users: eight waves, two a SIMD (a 64-register W slice)
    return p.B <= 128 ? launch_q<4, 4, 8>(p, chunks, tc, n, p.Q, st)
                      : launch_q<4, 8, 8>(p, chunks, tc, n, p.Q, st);
  if (NW == 8) {
    if (MT == 1) return KS == 4 ? launch_q<1, 4, 8>(p, chunks, tc, n, p.Q, st)
                                : launch_q<1, 8, 8>(p, chunks, tc, n, p.Q, st);
    return KS == 4 ? launch_q<2, 4, 8>(p, chunks, tc, n, p.Q, st)
                   : launch_q<2, 8, 8>(p, chunks, tc, n, p.Q, st);
  }
  if (MT == 1) return KS == 8 ? launch_q<1, 8, 4>(p, chunks, tc, n, p.Q, st)
                              : launch_q<1, 16, 4>(p, chunks, tc, n, p.Q, st);
  return KS == 8 ? launch_q<2, 8, 4>(p, chunks, tc, n, p.Q, st)
                 : launch_q<2, 16, 4>(p, chunks, tc, n, p.Q, st);
}

}  // namespace detail
}  // namespace mega
