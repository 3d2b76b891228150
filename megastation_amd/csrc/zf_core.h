// Zero-forcing weights of one subcarrier group, W = (A A^H)^-1 A with
// A = H^H (U x B, column-major), computed by one workgroup of NTH threads.
//
// Reference: src/operators/beamform/beamform.cu:74-122 runs a batched Cgemm
// (Gram), a batched potrf (lower Cholesky) and G*B single-right-hand-side
// potrs solves through cuBLAS/cuSOLVER. Here the three steps are fused per
// group, both products on the FP32 matrix cores (v_mfma_f32_16x16x4_f32):
//   1. Gram G = A A^H, users padded to 16-row tiles, K = 4 antennas per step:
//        Re G += Re A Re A^T + Im A Im A^T,  Im G += Im A Re A^T - Re A Im A^T.
//      For the 16x16x4 instruction the A operand of lane l is X[l%16][l/16]
//      and the B operand is Y[l%16][l/16] for C = X Y^T, so one load of
//      A[u = l%16][b = 4s + l/16] (64 contiguous complex for U = 16) feeds
//      both operands of a diagonal tile.
//   2. G^-1 by in-place Gauss-Jordan elimination in LDS (no pivoting: G is
//      Hermitian positive definite, and its pivots are exactly the squared
//      Cholesky diagonals). A pivot not above U * 2^-23 * G_kk marks the group
//      as numerically singular: status k+1 and W = 0 (the reference never
//      reads cuSOLVER's status and would emit Inf/NaN).
//   3. W = G^-1 A, M = users, N = 16 antennas, K = 4 users per step:
//        Re W = Re Gi Re A - Im Gi Im A,  Im W = Re Gi Im A + Im Gi Re A.
//      A wave owns whole antenna columns (all user tiles), loads them once
//      and only then stores, so writing W over A (in place) is safe.
#pragma once

#include "eq_core.h"

namespace mega {
namespace detail {

constexpr float kPivotTol = 1.1920929e-7f;  // 2^-23, times U

template <int UT>
struct ZfShared {
  static constexpr int UP = 16 * UT;
  static constexpr int LDG = UP + 1;  // padded row (complex elements)
  float2 G[UP * LDG];
  float diag[UP];
};

// Returns the group status (0 = ok, k+1 = pivot k failed), identical in every
// thread. W is written as W[u + ld_w * b] (ld_w = U for the CSI layout, or a
// padded LDS row); w2 (optional) receives a second copy with ld = U.
template <int UT, int NTH>
__device__ __forceinline__ int zf_group(const float2* A, uint32_t U, uint32_t B,
                                        ZfShared<UT>& sh, float2* w,
                                        uint32_t ld_w, float2* w2) {
  constexpr int UP = ZfShared<UT>::UP;
  constexpr int LDG = ZfShared<UT>::LDG;
  constexpr int NW = NTH / 64;
  constexpr int EPT = (UP * UP + NTH - 1) / NTH;
  float2* sG = sh.G;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int i16 = lane & 15, q4 = lane >> 4;
  const uint32_t ks = (B + 3) / 4;

  // ---- 1. Gram tiles (mi >= mj) ----------------------------------------------
  constexpr int NT = UT * (UT + 1) / 2;
  for (int tile = wave; tile < NT; tile += NW) {
    int mi = 0, mj = tile;
    while (mj > mi) {
      mj -= mi + 1;
      ++mi;
    }
    const uint32_t ui = 16 * mi + i16, uj = 16 * mj + i16;
    v4f gr1 = {0, 0, 0, 0}, gr2 = {0, 0, 0, 0}, gi1 = {0, 0, 0, 0}, gi2 = {0, 0, 0, 0};
#pragma unroll 8
    for (uint32_t s = 0; s < ks; ++s) {
      const uint32_t b = 4 * s + q4;
      const bool bok = b < B;
      const float2 ai = (ui < U && bok) ? A[ui + (uint64_t)U * b] : make_float2(0.f, 0.f);
      const float2 aj = (mi == mj) ? ai
                        : ((uj < U && bok) ? A[uj + (uint64_t)U * b] : make_float2(0.f, 0.f));
      gr1 = mfma4(ai.x, aj.x, gr1);
      gr2 = mfma4(ai.y, aj.y, gr2);
      gi1 = mfma4(ai.y, aj.x, gi1);
      gi2 = mfma4(-ai.x, aj.y, gi2);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = 16 * mi + 4 * q4 + r, col = 16 * mj + i16;
      const float2 v = make_float2(gr1[r] + gr2[r], gi1[r] + gi2[r]);
      sG[row * LDG + col] = v;
      if (mi != mj) sG[col * LDG + row] = make_float2(v.x, -v.y);
    }
  }
  __syncthreads();
  for (int i = tid; i < UP; i += NTH) {
    if (i >= (int)U) sG[i * LDG + i] = make_float2(1.f, 0.f);
    sh.diag[i] = sG[i * LDG + i].x;
  }
  __syncthreads();

  // ---- 2. Gauss-Jordan inverse in place ----------------------------------------
  int fail = 0;
  const float tol = kPivotTol * (float)U;
  for (uint32_t k = 0; k < U; ++k) {
    const float piv = sG[k * LDG + k].x;
    if (!(piv > tol * sh.diag[k]) || !isfinite(piv)) {
      fail = (int)k + 1;  // uniform: every thread read the same value
      break;
    }
    const float ip = 1.0f / piv;
    float2 nv[EPT];
#pragma unroll
    for (int j = 0; j < EPT; ++j) {
      const int e = tid + j * NTH;
      if (e < UP * UP) {
        const int i = e / UP, c = e % UP;
        const float2 aik = sG[i * LDG + k], akj = sG[k * LDG + c];
        float2 v;
        if (i == (int)k && c == (int)k) {
          v = make_float2(ip, 0.f);
        } else if (i == (int)k) {
          v = make_float2(akj.x * ip, akj.y * ip);
        } else if (c == (int)k) {
          v = make_float2(-aik.x * ip, -aik.y * ip);
        } else {
          const float2 aij = sG[i * LDG + c];
          const float pr = aik.x * akj.x - aik.y * akj.y;
          const float pi = aik.x * akj.y + aik.y * akj.x;
          v = make_float2(aij.x - pr * ip, aij.y - pi * ip);
        }
        nv[j] = v;
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < EPT; ++j) {
      const int e = tid + j * NTH;
      if (e < UP * UP) sG[(e / UP) * LDG + e % UP] = nv[j];
    }
    __syncthreads();
  }
  if (fail) {
    for (uint32_t e = tid; e < U * B; e += NTH) {
      const uint32_t u = e % U, b = e / U;
      w[u + (uint64_t)ld_w * b] = make_float2(0.f, 0.f);
      if (w2) w2[e] = make_float2(0.f, 0.f);
    }
    return fail;
  }

  // ---- 3. W = G^-1 A -----------------------------------------------------------
  constexpr int KSU = UP / 4;  // K steps over users
  const uint32_t nb_tiles = (B + 15) / 16;
  for (uint32_t nb = wave; nb < nb_tiles; nb += NW) {
    const uint32_t b = 16 * nb + i16;
    float2 av[KSU];
#pragma unroll
    for (int s = 0; s < KSU; ++s) {
      const uint32_t u = 4 * s + q4;
      av[s] = (u < U && b < B) ? A[u + (uint64_t)U * b] : make_float2(0.f, 0.f);
    }
    v4f res_r[UT], res_i[UT];
#pragma unroll
    for (int mi = 0; mi < UT; ++mi) {
      v4f wr1 = {0, 0, 0, 0}, wr2 = {0, 0, 0, 0}, wi1 = {0, 0, 0, 0}, wi2 = {0, 0, 0, 0};
#pragma unroll
      for (int s = 0; s < KSU; ++s) {
        const float2 gv = sG[(16 * mi + i16) * LDG + 4 * s + q4];
        wr1 = mfma4(gv.x, av[s].x, wr1);
        wr2 = mfma4(-gv.y, av[s].y, wr2);
        wi1 = mfma4(gv.x, av[s].y, wi1);
        wi2 = mfma4(gv.y, av[s].x, wi2);
      }
      res_r[mi] = wr1 + wr2;
      res_i[mi] = wi1 + wi2;
    }
    if (b < B) {
#pragma unroll
      for (int mi = 0; mi < UT; ++mi)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const uint32_t u = 16 * mi + 4 * q4 + r;
          if (u < U) {
            const float2 v = make_float2(res_r[mi][r], res_i[mi][r]);
            w[u + (uint64_t)ld_w * b] = v;
            if (w2) w2[u + (uint64_t)U * b] = v;
          }
        }
    }
  }
  return 0;
}

}  // namespace detail
}  // namespace mega
