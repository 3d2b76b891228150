// Zero-forcing weights of one subcarrier group, W = (A A^H)^-1 A with
// A = H^H (U x B, column-major), computed by one workgroup of NTH threads.
//
// Reference: src/operators/beamform/beamform.cu:74-122 runs a batched Cgemm
// (Gram), a batched potrf (lower Cholesky) and G*B single-right-hand-side
// potrs solves through cuBLAS/cuSOLVER. Here the steps are fused per group,
// both products on the FP32 matrix cores (v_mfma_f32_16x16x4_f32):
//   0. For up to 16 users A is staged into LDS once (rows of U+1 complex, all
//      loads in flight) and both products read it from there. Larger groups
//      read A from global memory (L2 for the second product), which keeps the
//      workgroup's LDS small enough for several groups per CU.
//   1. Gram G = A A^H, users padded to 16-row tiles, K = 4 antennas per step:
//        Re G += Re A Re A^T + Im A Im A^T,  Im G += Im A Re A^T - Re A Im A^T.
//      For the 16x16x4 instruction the A operand of lane l is X[l%16][l/16]
//      and the B operand is Y[l%16][l/16] for C = X Y^T, so one load of
//      A[u = l%16][b = 4s + l/16] feeds both operands of a diagonal tile.
//      The antennas are summed in fixed chunks of 64 (partial sums added in
//      chunk order), so the result does not depend on how many waves share
//      the work.
//   2. -G^-1 by U sweeps of the (Hermitian-preserving) sweep operator on the
//      Gram matrix: for pivot k, with a'_ik = -1 when i = k and a'_kc = -1
//      when c = k,
//        new[i][c] = (i == k or c == k ? 0 : a[i][c]) - a'_ik * a'_kc / a_kk,
//      which after all pivots leaves -G^-1 (no pivoting: G is Hermitian
//      positive definite and the pivots are exactly the squared Cholesky
//      diagonals). Entries stay in registers; the pivot row goes through
//      LDS (its column is its conjugate), one barrier per pivot. A pivot
//      not above U * 2^-23 * G_kk marks the
//      group numerically singular: status k+1 and W = 0 (the reference never
//      reads cuSOLVER's status and would emit Inf/NaN).
//   3. W = G^-1 A = -(sweep result) A, M = users, N = 16 antennas, K = 4 users
//      per step. A wave owns whole antenna columns (all user tiles), loads
//      them before storing, so writing W over A (in place) is safe.
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
  float2 prow[2][UP];
  float diag[UP];
};

// Gram work split: NT lower-triangle tiles x NCH antenna chunks of 64. When a
// workgroup of up to 8 waves has more waves than tiles, the chunks of a tile
// are spread over the waves and their partial sums go through LDS; otherwise
// a wave sums its tile's chunks in registers, in the same order (16-wave
// workgroups, at 128 registers a lane, have no room for the split's two
// chunk buffers).
template <int UT, int NW>
struct ZfPlan {
  static constexpr int NT = UT * (UT + 1) / 2;
  static constexpr bool SPLIT = NW > NT && NW <= 8;
};

constexpr uint32_t kGramChunk = 16;  // K steps (4 antennas each) per chunk

__host__ __device__ constexpr uint32_t gram_chunks(uint32_t B) {
  return ((B + 3) / 4 + kGramChunk - 1) / kGramChunk;
}

// LDS bytes zf_group needs beyond ZfShared: A (when staged) and Gram partials.
template <int UT, int NTH>
__host__ __device__ constexpr size_t zf_dyn_lds(uint32_t U, uint32_t B, bool a_lds) {
  return (a_lds ? (size_t)(U + 1) * B * sizeof(float2) : 0) +
         (ZfPlan<UT, NTH / 64>::SPLIT
              ? (size_t)ZfPlan<UT, NTH / 64>::NT * gram_chunks(B) * 256 * sizeof(float2)
              : 0);
}

// Copies A (U x B, column-major) into LDS rows of U+1 complex, with every
// thread keeping up to 16 loads in flight.
template <int NTH>
__device__ __forceinline__ void stage_a(const float2* A, uint32_t U, uint32_t B,
                                        float2* sA) {
  const uint32_t total = U * B;
  const int tid = threadIdx.x;
  for (uint32_t base = 0; base < total; base += NTH * 16) {
    float2 tmp[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const uint32_t e = base + j * NTH + tid;
      tmp[j] = e < total ? A[e] : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const uint32_t e = base + j * NTH + tid;
      if (e < total) {
        const uint32_t b = e / U, u = e - b * U;
        sA[b * (U + 1) + u] = tmp[j];
      }
    }
  }
}

// One Gram item from global A (ld = U): tile (mi, mj) over antenna chunk ch.
// All loads of the chunk are issued up front (gram_item_fetch) with clamped,
// always valid addresses and no branches; rows past U carry clamped data (the
// caller clears them), antennas past B are multiplied by 0 (the steps past
// the last antenna then add exact zeros). gram_item_mfma accumulates the
// four real products of every step; the item's value is
// (gr1 + gr2, gi1 + gi2) per result register.
struct GramBuf {
  float2 ai[kGramChunk], aj[kGramChunk];
};

__device__ __forceinline__ void gram_item_fetch(const float2* A, uint32_t U, uint32_t B, int mi,
                                                int mj, uint32_t ch, GramBuf& L) {
  const int lane = threadIdx.x & 63, i16 = lane & 15, q4 = lane >> 4;
  const uint32_t ui = min(16u * mi + i16, U - 1), uj = min(16u * mj + i16, U - 1);
#pragma unroll
  for (uint32_t s = 0; s < kGramChunk; ++s) {
    const uint32_t b = min(4 * (ch * kGramChunk + s) + q4, B - 1);
    L.ai[s] = A[ui + (uint64_t)U * b];
    L.aj[s] = A[uj + (uint64_t)U * b];
  }
}

__device__ __forceinline__ void gram_item_mfma(const GramBuf& L, uint32_t B, uint32_t ch,
                                               v4f& gr1, v4f& gr2, v4f& gi1, v4f& gi2) {
  const int q4 = (threadIdx.x & 63) >> 4;
  __builtin_amdgcn_sched_barrier(0);  // the loads stay ahead of the MFMAs
#pragma unroll
  for (uint32_t s = 0; s < kGramChunk; ++s) {
    const float keep = 4 * (ch * kGramChunk + s) + q4 < B ? 1.f : 0.f;
    const float2 ai = make_float2(L.ai[s].x * keep, L.ai[s].y * keep), aj = L.aj[s];
    gr1 = mfma4(ai.x, aj.x, gr1);
    gr2 = mfma4(ai.y, aj.y, gr2);
    gi1 = mfma4(ai.y, aj.x, gi1);
    gi2 = mfma4(-ai.x, aj.y, gi2);
  }
}

// Entry (lane l, register r) of Gram tile (mi, mj) into G (rows of LDG) and
// its Hermitian mirror; rows or columns past U are written as 0.
template <int LDG>
__device__ __forceinline__ void gram_put_tile(float2* G, uint32_t U, int mi, int mj, int l,
                                              int r, float2 v) {
  const int row = 16 * mi + 4 * (l >> 4) + r, col = 16 * mj + (l & 15);
  if (row >= (int)U || col >= (int)U) v = make_float2(0.f, 0.f);
  G[row * LDG + col] = v;
  if (mi != mj) G[col * LDG + row] = make_float2(v.x, -v.y);
}

__device__ __forceinline__ void tile_of(int tile, int& mi, int& mj) {
  mi = 0;
  mj = tile;
  while (mj > mi) {
    mj -= mi + 1;
    ++mi;
  }
}

__device__ __forceinline__ float2 cmul2(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ float2 shfl2(float2 v, int src) {
  return make_float2(__shfl(v.x, src, 64), __shfl(v.y, src, 64));
}

// Reciprocal of a pivot: the hardware estimate and one Newton step.
__device__ __forceinline__ float pivot_rcp(float p) {
  float r = __builtin_amdgcn_rcpf(p);
  const float e = __builtin_fmaf(-p, r, 1.0f);
  return __builtin_fmaf(r, e, r);
}

// Overwrites the Gram matrix in sh.G (Hermitian, users padded to UP with an
// identity block) with -G^-1 by U pivots of the sweep operator; returns 0 or
// k+1 for the first pivot k at or below U * 2^-23 * G_kk (sh.G then holds
// unspecified values). All NTH threads of the workgroup call it after a
// barrier that follows the writes of sh.G; it ends with a barrier.
//
// NS threads sweep (one wave for up to 16 users, else the workgroup): each
// keeps EPT entries in registers, column c = tid % UP of rows
// i = tid / UP + RS*j. The sweep keeps the matrix Hermitian, so column k is
// the conjugate of row k and only row k is exchanged: its owners publish it
// (double-buffered) right after updating it for pivot k-1, then one barrier,
// then every thread issues its reads of row k and finishes the rest of its
// pivot-(k-1) updates while they are in flight. The pivot loop is unrolled,
// so the register index of every entry is static.
template <int UT, int NTH>
__device__ __forceinline__ int zf_sweep(ZfShared<UT>& sh, uint32_t U) {
  constexpr int UP = ZfShared<UT>::UP;
  constexpr int LDG = ZfShared<UT>::LDG;
  constexpr int NS = UT == 1 ? 64 : NTH;
  static_assert(NS <= NTH && (UP * UP) % NS == 0 && NS % UP == 0, "sweep layout");
  constexpr int EPT = UP * UP / NS;
  constexpr int RS = NS / UP;
  const int tid = threadIdx.x;
  float2* sG = sh.G;
  for (int i = tid; i < UP; i += NTH) {
    if (i >= (int)U) sG[i * LDG + i] = make_float2(1.f, 0.f);
    sh.diag[i] = sG[i * LDG + i].x;
  }
  __syncthreads();

  const float tol = kPivotTol * (float)U;
  int fail = 0;
  if (tid < NS) {  // whole waves
    const int c = tid % UP, r0 = tid / UP;
    float2 a[EPT];
#pragma unroll
    for (int j = 0; j < EPT; ++j) a[j] = sG[(r0 + RS * j) * LDG + c];
    // row k of the current matrix: entry c, entries r0 + RS*j, the pivot
    float2 rc, ri[EPT];
    float piv, dk;
    auto read_row = [&](const float2* row, int k) {
      rc = row[c];
#pragma unroll
      for (int j = 0; j < EPT; ++j) ri[j] = row[r0 + RS * j];
      piv = row[k].x;
      dk = sh.diag[k];
    };
    read_row(sG, 0);  // row 0 is still the Gram matrix's
    static_for<0, UP>([&](auto kc) {
      constexpr int k = decltype(kc)::value;
      if (k >= (int)U) return;  // uniform: pivots past U are not swept
      const bool bad = !(piv > tol * dk) || !isfinite(piv);
      fail = (fail == 0 && bad) ? k + 1 : fail;
      const float ip = pivot_rcp(piv);
      const bool ck = c == k;
      const float2 yip = ck ? make_float2(-ip, 0.f) : make_float2(rc.x * ip, rc.y * ip);
      float2 xk[EPT];
#pragma unroll
      for (int j = 0; j < EPT; ++j) xk[j] = make_float2(ri[j].x, -ri[j].y);  // a[i][k]
      auto update = [&](auto jc) {
        constexpr int j = decltype(jc)::value;
        const bool ik = (j == k / RS) && (r0 == k % RS);
        const float2 x = ik ? make_float2(-1.f, 0.f) : xk[j];
        const float2 base = (ik || ck) ? make_float2(0.f, 0.f) : a[j];
        a[j] = make_float2(base.x - x.x * yip.x + x.y * yip.y,
                           base.y - x.x * yip.y - x.y * yip.x);
      };
      constexpr int jn = (k + 1) / RS;  // register of row k+1 (if owned)
      if constexpr (k + 1 < UP) {
        update(std::integral_constant<int, jn>{});
        if (r0 == (k + 1) % RS) sh.prow[(k + 1) & 1][c] = a[jn];
      }
      if constexpr (UT == 1) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
      } else {
        __syncthreads();
      }
      if constexpr (k + 1 < UP) read_row(sh.prow[(k + 1) & 1], k + 1);
      __builtin_amdgcn_sched_barrier(0);  // the reads go out before the updates
      static_for<0, EPT>([&](auto jc) {
        if constexpr (k + 1 >= UP || decltype(jc)::value != jn) update(jc);
      });
    });
#pragma unroll
    for (int j = 0; j < EPT; ++j) sG[(r0 + RS * j) * LDG + c] = a[j];
    if (UT == 1 && tid == 0) sh.diag[0] = __int_as_float(fail);
  }
  __syncthreads();
  if constexpr (UT == 1) fail = __float_as_int(sh.diag[0]);
  return fail;
}

// Gram matrix of one group into sh.G (users padded to UP with zero rows and
// columns; the sweep puts the identity block there), ending with a barrier.
// A is read from global memory (ld = U) or, when A_LDS, first staged into sA
// (rows of U+1). part: LDS scratch for split Gram partials (zf_dyn_lds).
template <int UT, int NTH, bool A_LDS>
__device__ __forceinline__ void zf_gram(const float2* A, uint32_t U, uint32_t B,
                                        ZfShared<UT>& sh, float2* sA, float2* part) {
  constexpr int LDG = ZfShared<UT>::LDG;
  constexpr int NW = NTH / 64;
  using Plan = ZfPlan<UT, NW>;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int i16 = lane & 15, q4 = lane >> 4;
  const uint32_t ks = (B + 3) / 4;
  const uint32_t nch = gram_chunks(B);
  const uint32_t lda = A_LDS ? U + 1 : U;
  const float2* Asrc = A_LDS ? sA : A;
  float2* sG = sh.G;
  if (A_LDS) {
    stage_a<NTH>(A, U, B, sA);
    __syncthreads();
  }
  auto aval = [&](uint32_t u, uint32_t b) -> float2 {
    return (u < U && b < B) ? Asrc[u + (uint64_t)lda * b] : make_float2(0.f, 0.f);
  };
  // From global memory (no staging): gram_item_fetch / gram_item_mfma.
  GramBuf gbuf[2];
  // Gram of tile (mi, mj) over antenna chunk ch: this lane's 4 entries.
  auto gram_chunk = [&](int mi, int mj, uint32_t ch, float2 (&out)[4]) {
    const uint32_t s0 = ch * kGramChunk;
    v4f gr1 = {0, 0, 0, 0}, gr2 = {0, 0, 0, 0}, gi1 = {0, 0, 0, 0}, gi2 = {0, 0, 0, 0};
    auto step = [&](float2 ai, float2 aj) {
      gr1 = mfma4(ai.x, aj.x, gr1);
      gr2 = mfma4(ai.y, aj.y, gr2);
      gi1 = mfma4(ai.y, aj.x, gi1);
      gi2 = mfma4(-ai.x, aj.y, gi2);
    };
    if constexpr (A_LDS) {
      const uint32_t ui = 16 * mi + i16, uj = 16 * mj + i16;
      const uint32_t s1 = min(s0 + kGramChunk, ks);
      for (uint32_t s = s0; s < s1; ++s) {
        const uint32_t b = 4 * s + q4;
        const float2 ai = aval(ui, b);
        step(ai, (mi == mj) ? ai : aval(uj, b));
      }
    } else {
      gram_item_fetch(A, U, B, mi, mj, ch, gbuf[0]);
      gram_item_mfma(gbuf[0], B, ch, gr1, gr2, gi1, gi2);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) out[r] = make_float2(gr1[r] + gr2[r], gi1[r] + gi2[r]);
  };
  auto put_tile = [&](int mi, int mj, int l, int r, float2 v) {
    gram_put_tile<LDG>(sG, U, mi, mj, l, r, v);
  };

  // ---- 1. Gram ---------------------------------------------------------------
  if constexpr (!Plan::SPLIT) {
    for (int tile = wave; tile < Plan::NT; tile += NW) {
      int mi, mj;
      tile_of(tile, mi, mj);
      float2 sum[4], c[4];
      gram_chunk(mi, mj, 0, sum);
      for (uint32_t ch = 1; ch < nch; ++ch) {
        gram_chunk(mi, mj, ch, c);
#pragma unroll
        for (int r = 0; r < 4; ++r) sum[r] = make_float2(sum[r].x + c[r].x, sum[r].y + c[r].y);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) put_tile(mi, mj, lane, r, sum[r]);
    }
  } else {
    auto item_tile = [&](uint32_t item, int& mi, int& mj) { tile_of(item / nch, mi, mj); };
    auto put_part = [&](uint32_t item, const v4f& r1, const v4f& r2, const v4f& i1,
                        const v4f& i2) {
#pragma unroll
      for (int r = 0; r < 4; ++r)
        part[item * 256 + r * 64 + lane] = make_float2(r1[r] + r2[r], i1[r] + i2[r]);
    };
    if constexpr (A_LDS) {
      for (uint32_t item = wave; item < Plan::NT * nch; item += NW) {
        int mi, mj;
        item_tile(item, mi, mj);
        float2 c[4];
        gram_chunk(mi, mj, item % nch, c);
#pragma unroll
        for (int r = 0; r < 4; ++r) part[item * 256 + r * 64 + lane] = c[r];
      }
    } else {
      // the next item's loads are in flight during this item's MFMAs
      const uint32_t nitems = Plan::NT * nch;
      int mi, mj;
      if ((uint32_t)wave < nitems) {
        item_tile(wave, mi, mj);
        gram_item_fetch(A, U, B, mi, mj, wave % nch, gbuf[0]);
      }
      for (uint32_t it0 = wave; it0 < nitems; it0 += 2 * NW) {
#pragma unroll
        for (int r = 0; r < 2; ++r) {
          const uint32_t item = it0 + r * NW;
          if (item >= nitems) break;
          if (item + NW < nitems) {
            item_tile(item + NW, mi, mj);
            gram_item_fetch(A, U, B, mi, mj, (item + NW) % nch, gbuf[r ^ 1]);
          }
          v4f gr1 = {0, 0, 0, 0}, gr2 = {0, 0, 0, 0}, gi1 = {0, 0, 0, 0}, gi2 = {0, 0, 0, 0};
          gram_item_mfma(gbuf[r], B, item % nch, gr1, gr2, gi1, gi2);
          put_part(item, gr1, gr2, gi1, gi2);
        }
      }
    }
    __syncthreads();
    for (int e = tid; e < Plan::NT * 256; e += NTH) {
      const int tile = e / 256, r = (e % 256) / 64, l = e % 64;
      int mi, mj;
      tile_of(tile, mi, mj);
      float2 v = part[(tile * nch) * 256 + r * 64 + l];
      for (uint32_t ch = 1; ch < nch; ++ch) {
        const float2 q = part[(tile * nch + ch) * 256 + r * 64 + l];
        v = make_float2(v.x + q.x, v.y + q.y);
      }
      put_tile(mi, mj, l, r, v);
    }
  }
  __syncthreads();

}

// W = G^-1 A from -G^-1 in sh.G (the sweep result): W[u + ld_w * b], w2
// (optional) a second copy with ld = U. A as for zf_gram (sA when A_LDS).
template <int UT, int NTH, bool A_LDS>
__device__ __forceinline__ void zf_apply(const float2* A, uint32_t U, uint32_t B,
                                         ZfShared<UT>& sh, const float2* sA, float2* w,
                                         uint32_t ld_w, float2* w2) {
  constexpr int UP = ZfShared<UT>::UP;
  constexpr int LDG = ZfShared<UT>::LDG;
  constexpr int NW = NTH / 64;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int i16 = lane & 15, q4 = lane >> 4;
  const uint32_t lda = A_LDS ? U + 1 : U;
  const float2* Asrc = A_LDS ? sA : A;
  const float2* res = sh.G;
  auto aval = [&](uint32_t u, uint32_t b) -> float2 {
    return (u < U && b < B) ? Asrc[u + (uint64_t)lda * b] : make_float2(0.f, 0.f);
  };
  // ---- 3. W = G^-1 A = -(res) A -------------------------------------------------
  constexpr int KSU = UP / 4;  // K steps over users
  const uint32_t nb_tiles = (B + 15) / 16;
  // A columns of the wave's antenna tiles, the next tile's in flight during
  // this tile's MFMAs (global A: clamped addresses, users past U zeroed)
  float2 abuf[2][KSU];
  auto fetch_a = [&](uint32_t nb, float2 (&d)[KSU]) {
    if constexpr (A_LDS) {
#pragma unroll
      for (int s = 0; s < KSU; ++s) d[s] = aval(4 * s + q4, 16 * nb + i16);
    } else {
      const uint32_t bc = min(16 * min(nb, nb_tiles - 1) + i16, B - 1);
#pragma unroll
      for (int s = 0; s < KSU; ++s) d[s] = A[min(4u * s + q4, U - 1) + (uint64_t)U * bc];
    }
  };
  auto apply_tile = [&](uint32_t nb, const float2 (&cur)[KSU]) {
    const uint32_t b = 16 * nb + i16;
    float2 av[KSU];
#pragma unroll
    for (int s = 0; s < KSU; ++s) {
      const bool ok = 4u * s + q4 < U;
      av[s] = ok ? cur[s] : make_float2(0.f, 0.f);
    }
    v4f res_r[UT], res_i[UT];
#pragma unroll
    for (int mi = 0; mi < UT; ++mi) {
      v4f wr1 = {0, 0, 0, 0}, wr2 = {0, 0, 0, 0}, wi1 = {0, 0, 0, 0}, wi2 = {0, 0, 0, 0};
#pragma unroll
      for (int s = 0; s < KSU; ++s) {
        const float2 m = res[(16 * mi + i16) * LDG + 4 * s + q4];
        const float gr = -m.x, gi = -m.y;  // G^-1 = -res
        wr1 = mfma4(gr, av[s].x, wr1);
        wr2 = mfma4(-gi, av[s].y, wr2);
        wi1 = mfma4(gr, av[s].y, wi1);
        wi2 = mfma4(gi, av[s].x, wi2);
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
  };
  // A wave owns the antenna tiles wave, wave+NW, ...; only its own columns
  // are read (ahead, from global memory) and overwritten, so W may replace A
  // in place.
  if constexpr (NTH >= 1024) {
    // 16-wave workgroups (128 registers a lane): one A buffer, the user
    // tiles two at a time; the same products and sums as apply_tile
    for (uint32_t nb = wave; nb < nb_tiles; nb += NW) {
      fetch_a(nb, abuf[0]);
      const uint32_t b = 16 * nb + i16;
#pragma unroll
      for (int s = 0; s < KSU; ++s)
        if (4u * s + q4 >= U) abuf[0][s] = make_float2(0.f, 0.f);
      // the -G^-1 reads are the same for every antenna tile: an offset the
      // compiler cannot see through (always 0) keeps it from holding all of
      // them in registers across tiles; the pointer itself stays sh.G's
      uint32_t zero = 0;
      __asm__ __volatile__("" : "+s"(zero));
      const float2* rs = res + zero;
#pragma unroll
      for (int m0 = 0; m0 < UT; m0 += 2) {
        v4f rr[2], ri[2];
#pragma unroll
        for (int mi = 0; mi < 2; ++mi) {
          v4f wr1 = {0, 0, 0, 0}, wr2 = {0, 0, 0, 0}, wi1 = {0, 0, 0, 0}, wi2 = {0, 0, 0, 0};
#pragma unroll
          for (int s = 0; s < KSU; ++s) {
            const float2 m = rs[(16 * (m0 + mi) + i16) * LDG + 4 * s + q4];
            const float gr = -m.x, gi = -m.y;
            wr1 = mfma4(gr, abuf[0][s].x, wr1);
            wr2 = mfma4(-gi, abuf[0][s].y, wr2);
            wi1 = mfma4(gr, abuf[0][s].y, wi1);
            wi2 = mfma4(gi, abuf[0][s].x, wi2);
          }
          rr[mi] = wr1 + wr2;
          ri[mi] = wi1 + wi2;
        }
        if (b < B) {
#pragma unroll
          for (int mi = 0; mi < 2; ++mi)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const uint32_t u = 16 * (m0 + mi) + 4 * q4 + r;
              if (u < U) {
                const float2 v = make_float2(rr[mi][r], ri[mi][r]);
                w[u + (uint64_t)ld_w * b] = v;
                if (w2) w2[u + (uint64_t)U * b] = v;
              }
            }
        }
      }
    }
  } else if constexpr (A_LDS) {
    for (uint32_t nb = wave; nb < nb_tiles; nb += NW) {
      fetch_a(nb, abuf[0]);
      apply_tile(nb, abuf[0]);
    }
  } else {
    fetch_a(wave, abuf[0]);
    for (uint32_t n0 = wave; n0 < nb_tiles; n0 += 2 * NW) {
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        const uint32_t nb = n0 + r * NW;
        if (nb >= nb_tiles) break;
        if (nb + NW < nb_tiles) fetch_a(nb + NW, abuf[r ^ 1]);
        apply_tile(nb, abuf[r]);
      }
    }
  }
}


// Returns the group status (0 = ok, k+1 = pivot k failed), identical in every
// thread: zf_gram, zf_sweep, zf_apply. W is written as W[u + ld_w * b] (ld_w =
// U for the CSI layout, or a padded LDS row); w2 (optional) receives a
// second copy with ld = U; on failure W = 0.
template <int UT, int NTH, bool A_LDS>
__device__ __forceinline__ int zf_group(const float2* A, uint32_t U, uint32_t B,
                                        ZfShared<UT>& sh, float2* sA, float2* part,
                                        float2* w, uint32_t ld_w, float2* w2) {
  const int tid = threadIdx.x;
  zf_gram<UT, NTH, A_LDS>(A, U, B, sh, sA, part);
  const int fail = zf_sweep<UT, NTH>(sh, U);
  if (fail) {
    for (uint32_t e = tid; e < U * B; e += NTH) {
      const uint32_t u = e % U, b = e / U;
      w[u + (uint64_t)ld_w * b] = make_float2(0.f, 0.f);
      if (w2) w2[e] = make_float2(0.f, 0.f);
    }
    return fail;
  }
  zf_apply<UT, NTH, A_LDS>(A, U, B, sh, sA, w, ld_w, w2);
  return 0;
}

}  // namespace detail
}  // namespace mega
