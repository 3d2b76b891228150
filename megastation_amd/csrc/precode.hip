// Downlink precode (the reference's Precode operator,
// src/operators/sgemv/precode.cu:20-57 and precode_kernel.cuh): per OFDM
// symbol and subcarrier k of group g = k / sg,
//   out[b][(k + bs_start + row/2) % row] = sum_u P_g[b][u] ue[u][k],
// P_g the B x U precoder (column-major, [G][U][B]), ue the user symbols
// ([U][N] per symbol), out the antenna rows of the IFFT input ([B][row] per
// symbol; the cyclic placement is the FFT shift). Like the reference, only the
// N / sg full groups are computed; other output entries are left untouched.
//
// The products run on the FP32 matrix cores (v_mfma_f32_16x16x4f32, 157
// TFLOP/s dense: at 256 x 32 a slot's precode is 2.8 GFLOP against 116 MB,
// beyond the vector ALUs' reach in the time HBM needs). The columns of a
// group are its (symbol, subcarrier) pairs, sg * S of them, since P_g serves
// every symbol. One workgroup per (group, block of 64 antennas) stages that
// block of P_g in LDS once ([U][64], at most 32 KB, coalesced loads); its four
// waves take the group's 16-column tiles in turn, each computing 32 antennas
// (half the block; the two halves on alternate waves) x 16 columns with A operands from LDS and B operands (user symbols) gathered
// from HBM. A complex product
// takes real 16x16x4 steps with rows interleaved (antenna r re, r im) and K
// interleaved (user u re, u im):
//   A[2r][2u] = Re P, A[2r][2u+1] = -Im P, A[2r+1][2u] = Im P,
//   A[2r+1][2u+1] = Re P;   B[2u][n] = Re x, B[2u+1][n] = Im x,
// so D[2r][n] = Re out and D[2r+1][n] = Im out, and each lane ends with
// whole complex outputs (rows 4*(l/16) .. +3 = two antennas) at column l%16.
// FP32 accumulation, users in increasing order (two per step).
#include <cstdlib>

#include "eq_core.h"

namespace mega {
namespace detail {
namespace {

// 8-antenna accumulator tiles per wave: NACC = 4 (32 antennas; the two
// halves of a workgroup's 64 antennas on alternate waves) or 8 (64).

// Workgroup = (group g, block of 64 antennas ab), ab fastest, in the logical
// order of xcd_linear: consecutive workgroups share an XCD, so a group's user
// symbols are fetched into one L2 for all its antenna blocks, and adjacent
// groups' partial 128-byte output lines meet in the same L2. Wave w takes the
// 16-column tiles w, w + 4, ... A tile's user symbols (at most UMAX users,
// two per step) are all loaded before its products, so the HBM latency is
// paid once per tile, not once per step.
template <int UMAX, int NACC>
__global__ __launch_bounds__(256) void precode_mfma(const PrecodeArgs a, uint32_t nab,
                                                   uint32_t ntile) {
  constexpr int NKS = UMAX / 2;
  constexpr int WPT = 8 / NACC;  // waves sharing a tile (antenna halves)
  const int wv = threadIdx.x >> 6;
  const uint32_t bh = (uint32_t)(wv % WPT) * 8 * NACC;  // the wave's antenna offset
  __shared__ float2 sP[kMaxUsers * 64];
  const uint32_t wg = xcd_linear(blockIdx.x, gridDim.x);
  const uint32_t ab = wg % nab, g = wg / nab;
  const uint32_t U = a.U, B = a.B, N = a.N, sg = a.sg;
  const uint32_t b0 = ab * 64;
  const int l = threadIdx.x & 63, i16 = l & 15, q = l >> 4;
  const int up = q >> 1, p = q & 1, h = i16 & 1;
  const uint32_t ncol = sg * a.S;
  // B operands of tile t: column n of the tile (past the group: its last
  // column, results dropped), K index q = 2 * (user within the step) + part
  auto load_x = [&](uint32_t t, float (&bv)[NKS]) {
    const uint32_t n = min(16 * t + i16, ncol - 1);
    const uint32_t s = n / sg, k = g * sg + (n - s * sg);
    const float2* xcol = a.ue + (uint64_t)s * a.ue_sym + k;
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks) {
      const uint32_t u = 2 * ks + up;
      const float2 x = u < U ? xcol[(uint64_t)u * N] : make_float2(0.f, 0.f);
      bv[ks] = p ? x.y : x.x;
    }
  };
  const float2* P = a.prec + (uint64_t)g * U * B;
  for (uint32_t e = threadIdx.x; e < U * 64; e += 256) {
    const uint32_t u = e / 64, bb = e % 64;
    sP[e] = b0 + bb < B ? P[(uint64_t)u * B + b0 + bb] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  // A operand of row 2r + h and K (user, part p): Re P when h == p, else
  // -Im P (h = 0) or Im P (h = 1)
  const bool take_re = h == p;
  const float sgn = h ? 1.f : -1.f;
  for (uint32_t t = wv / WPT; t < ntile; t += 4 / WPT) {
    float bv[NKS];
    load_x(t, bv);
    v4f acc[NACC];
#pragma unroll
    for (int m = 0; m < NACC; ++m) acc[m] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks) {
      if (2 * ks >= (int)U) break;  // uniform
      const uint32_t u = min(2 * ks + up, U - 1);  // u >= U: bv = 0 makes it inert
      const float2* prow = sP + u * 64 + bh + (i16 >> 1);
#pragma unroll
      for (int m = 0; m < NACC; ++m) {
        const float2 pv = prow[8 * m];
        const float av = take_re ? pv.x : sgn * pv.y;
        acc[m] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv[ks], acc[m], 0, 0, 0);
      }
    }
    const uint32_t n_raw = 16 * t + i16;
    if (n_raw < ncol) {
      const uint32_t s = n_raw / sg, k = g * sg + (n_raw - s * sg);
      const uint64_t col = ((uint64_t)k + a.bs_start + a.row / 2) % a.row;
      float2* O = a.out + (uint64_t)s * a.out_sym + col;
#pragma unroll
      for (int m = 0; m < NACC; ++m) {
        const uint32_t b = b0 + bh + 8 * m + 2 * q;
        if (b < B) O[(uint64_t)b * a.row] = make_float2(acc[m][0], acc[m][1]);
        if (b + 1 < B) O[(uint64_t)(b + 1) * a.row] = make_float2(acc[m][2], acc[m][3]);
      }
    }
  }
}

}  // namespace

hipError_t launch_precode(const PrecodeArgs& a, hipStream_t st, const char** why) {
  if (a.U == 0 || a.B == 0 || a.S == 0 || a.sg == 0 || a.N < a.sg) return hipSuccess;
  if (a.U > kMaxUsers) {
    *why = "more than 64 users";
    return hipErrorInvalidValue;
  }
  const uint64_t G = a.N / a.sg;  // full groups only, as the reference
  const uint64_t nab = (a.B + 63) / 64;
  const uint64_t ntile = ((uint64_t)a.sg * a.S + 15) / 16;
  if ((uint64_t)a.sg * a.S > 0xffffffffull || G * nab > 0xffffffffull) {
    *why = "grid too large";
    return hipErrorInvalidValue;
  }
  const dim3 grid((uint32_t)(G * nab)), block(256);
  // 32 antennas per wave (fewer registers, one more wave per SIMD at 32
  // users: 3% faster at 256x32); MEGA_PREC_ACC=8 selects 64
  const bool half = tune(kPrecAcc) != 8;
#define MEGA_PREC(UM)                                                                       \
  if (half)                                                                                  \
    hipLaunchKernelGGL((precode_mfma<UM, 4>), grid, block, 0, st, a, (uint32_t)nab,          \
                       (uint32_t)ntile);                                                     \
  else                                                                                       \
    hipLaunchKernelGGL((precode_mfma<UM, 8>), grid, block, 0, st, a, (uint32_t)nab,          \
                       (uint32_t)ntile);
  if (a.U <= 16) {
    MEGA_PREC(16)
  } else if (a.U <= 32) {
    MEGA_PREC(32)
  } else {
    MEGA_PREC(64)
  }
#undef MEGA_PREC
  return hipGetLastError();
}

}  // namespace detail
}  // namespace mega
