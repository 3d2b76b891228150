// Learned warm start for the ZF inverse (the learned form). Per subcarrier
// group and slot t, with G = A A^H:
//   features   [upper(G_t - G_{t-1}) / s, upper(X_{t-1}) * s],
//              s = tr(G_{t-1}) / U (upper = the U(U+1)/2 entries c >= i as
//              (re, im) pairs, row by row), as BF16 rows for the network;
//   warm start X0 = X_{t-1} - X_{t-1} D X_{t-1} + C / s, D = G_t - G_{t-1}
//              (first-order perturbation of the inverse plus the network's
//              correction C, Hermitian from its upper triangle);
//   refine     n Newton-Schulz steps X <- X (2I - G_t X);
//   check      r = |X G_t - I|_F; groups with r above the bound (or not
//              finite) are recomputed by the plain FP32 sweep (zf_core.h),
//              the fallback;
//   apply      W = X A on the matrix cores (zf_apply); X becomes X_{t-1}
//              of the next slot.
// The network itself runs on the kernels of mlp.hip.
#include "zf_core.h"

namespace mega {
namespace detail {
namespace {

template <int UT>
constexpr int lz_threads() { return UT == 1 ? 64 : 256; }

// G (row-major [U][U] per group) and X = G^-1 from the plain path.
template <int UT, bool A_LDS>
__global__ __launch_bounds__(lz_threads<UT>()) void gram_inv_kernel(
    const float2* __restrict__ in, float2* G_out, float2* X_out, int32_t* info, uint32_t U,
    uint32_t B) {
  constexpr int NTH = lz_threads<UT>();
  constexpr int LDG = ZfShared<UT>::LDG;
  __shared__ ZfShared<UT> sh;
  extern __shared__ float sDyn[];
  float2* sA = reinterpret_cast<float2*>(sDyn);
  float2* part = reinterpret_cast<float2*>(sDyn + (A_LDS ? 2 * (U + 1) * B : 0));
  const uint32_t g = blockIdx.x;
  zf_gram<UT, NTH, A_LDS>(in + (uint64_t)g * U * B, U, B, sh, sA, part);
  const uint64_t o = (uint64_t)g * U * U;
  if (G_out)
    for (uint32_t e = threadIdx.x; e < U * U; e += NTH) G_out[o + e] = sh.G[(e / U) * LDG + e % U];
  if (!X_out) return;  // Gram only (uniform)
  // zf_sweep's first barrier orders these reads before any of its writes
  const int fail = zf_sweep<UT, NTH>(sh, U);
  for (uint32_t e = threadIdx.x; e < U * U; e += NTH) {
    const float2 v = sh.G[(e / U) * LDG + e % U];
    X_out[o + e] = fail ? make_float2(0.f, 0.f) : make_float2(-v.x, -v.y);
  }
  if (threadIdx.x == 0 && info) info[g] = fail;
}

__device__ __forceinline__ void tri_index(uint32_t t, uint32_t U, uint32_t& i, uint32_t& c) {
  uint32_t row = 0, start = 0;
  while (start + (U - row) <= t) {
    start += U - row;
    ++row;
  }
  i = row;
  c = row + (t - start);
}

__global__ __launch_bounds__(256) void features_kernel(const float2* __restrict__ G,
                                                       const float2* __restrict__ Gp,
                                                       const float2* __restrict__ Xp,
                                                       __bf16* __restrict__ feat,
                                                       float* __restrict__ scale, uint32_t U) {
  __shared__ float s_sh;
  const uint32_t g = blockIdx.x;
  const uint64_t o = (uint64_t)g * U * U;
  if (threadIdx.x < 64) {  // trace of G_prev: one diagonal entry per lane (U <= 64)
    float tr = threadIdx.x < U ? Gp[o + threadIdx.x * (U + 1)].x : 0.f;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) tr += __shfl_xor(tr, off, 64);
    if (threadIdx.x == 0) {
      s_sh = tr > 0.f ? tr / (float)U : 1.f;
      scale[g] = s_sh;
    }
  }
  __syncthreads();
  const float s = s_sh, inv = 1.f / s;
  const uint32_t nt = U * (U + 1) / 2;
  __bf16* f = feat + (uint64_t)g * 4 * nt;
  for (uint32_t t = threadIdx.x; t < nt; t += blockDim.x) {
    uint32_t i, c;
    tri_index(t, U, i, c);
    const float2 a = G[o + i * U + c], b = Gp[o + i * U + c], x = Xp[o + i * U + c];
    f[2 * t] = (__bf16)((a.x - b.x) * inv);
    f[2 * t + 1] = (__bf16)((a.y - b.y) * inv);
    f[2 * nt + 2 * t] = (__bf16)(x.x * s);
    f[2 * nt + 2 * t + 1] = (__bf16)(x.y * s);
  }
}

// Training target: upper((X_true - X_base) * s) as (re, im) pairs, FP32.
__global__ __launch_bounds__(256) void target_kernel(const float2* __restrict__ Xt,
                                                     const float2* __restrict__ Xb,
                                                     const float* __restrict__ scale,
                                                     float* __restrict__ tgt, uint32_t U) {
  const uint32_t g = blockIdx.x;
  const uint64_t o = (uint64_t)g * U * U;
  const uint32_t nt = U * (U + 1) / 2;
  const float s = scale[g];
  float* t_out = tgt + (uint64_t)g * 2 * nt;
  for (uint32_t t = threadIdx.x; t < nt; t += blockDim.x) {
    uint32_t i, c;
    tri_index(t, U, i, c);
    const float2 a = Xt[o + i * U + c], b = Xb[o + i * U + c];
    t_out[2 * t] = (a.x - b.x) * s;
    t_out[2 * t + 1] = (a.y - b.y) * s;
  }
}

// Products of UP x UP complex matrices in LDS (rows of UP+1, zero past U) on
// the FP32 matrix cores: one 16 x 16 output tile per wave and step
// (v_mfma_f32_16x16x4_f32, four real products per complex step). Operand
// lane l: A[row l%16][k l/16], B[k l/16][col l%16]; result lane l, register
// r: row 4*(l/16)+r, column l%16. epi(row, col, re, im) consumes each entry.
template <int UP>
struct RefineLayout {
  static constexpr int NTH = 256, LD = UP + 1, NW = NTH / 64, TT = (UP / 16) * (UP / 16);
};

template <int UP, typename Epi>
__device__ __forceinline__ void cmm(const float2* A, const float2* B, Epi&& epi) {
  using L = RefineLayout<UP>;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i16 = lane & 15, q4 = lane >> 4;
  for (int tile = wave; tile < L::TT; tile += L::NW) {  // uniform per wave
    const int ti = tile / (UP / 16), tj = tile % (UP / 16);
    v4f cre = {0.f, 0.f, 0.f, 0.f}, cim = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < UP / 4; ++ks) {
      const float2 a = A[(16 * ti + i16) * L::LD + 4 * ks + q4];
      const float2 b = B[(4 * ks + q4) * L::LD + 16 * tj + i16];
      cre = mfma4(a.x, b.x, cre);
      cre = mfma4(-a.y, b.y, cre);
      cim = mfma4(a.x, b.y, cim);
      cim = mfma4(a.y, b.x, cim);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) epi(16 * ti + 4 * q4 + r, 16 * tj + i16, cre[r], cim[r]);
  }
}

// Warm start, Newton-Schulz refinement and the residual check (users padded
// to UP with zeros; the identity only on the first U diagonal entries).
template <int UP>
__global__ __launch_bounds__(256) void refine_kernel(
    const float2* __restrict__ G, const float2* __restrict__ Gp, const float2* __restrict__ Xp,
    const float* __restrict__ corr, uint32_t corr_ld, const float* __restrict__ scale,
    float2* __restrict__ X_out, float* __restrict__ res, int32_t* __restrict__ flag,
    float bound, uint32_t iters, uint32_t U) {
  using L = RefineLayout<UP>;
  constexpr int LD = L::LD;
  extern __shared__ float2 sm[];  // G, X, T, V: UP x LD each
  float2* sG = sm;
  float2* sX = sm + UP * LD;
  float2* sT = sm + 2 * UP * LD;
  float2* sV = sm + 3 * UP * LD;
  __shared__ float red[4];
  const uint32_t g = blockIdx.x;
  const uint64_t o = (uint64_t)g * U * U;
  for (int e = threadIdx.x; e < UP * UP; e += L::NTH) {
    const int i = e / UP, c = e % UP;
    const bool in = i < (int)U && c < (int)U;
    const float2 a = in ? G[o + i * U + c] : make_float2(0.f, 0.f);
    const float2 b = in ? Gp[o + i * U + c] : make_float2(0.f, 0.f);
    sG[i * LD + c] = a;
    sT[i * LD + c] = make_float2(a.x - b.x, a.y - b.y);  // D
    sX[i * LD + c] = in ? Xp[o + i * U + c] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  auto store = [](float2* M) {
    return [M](int i, int c, float re, float im) { M[i * LD + c] = make_float2(re, im); };
  };
  cmm<UP>(sX, sT, store(sV));  // Xp D
  __syncthreads();
  cmm<UP>(sV, sX, store(sT));  // Xp D Xp
  __syncthreads();
  // X0 = Xp - Xp D Xp + C / s (C Hermitian from its upper triangle)
  const float inv = corr ? 1.f / scale[g] : 0.f;
  const float* cg = corr ? corr + (uint64_t)g * corr_ld : nullptr;
  for (int e = threadIdx.x; e < UP * UP; e += L::NTH) {
    const int i = e / UP, c = e % UP;
    float cr = 0.f, ci = 0.f;
    if (cg && i < (int)U && c < (int)U) {
      const uint32_t a = min(i, c), b = max(i, c);
      const uint32_t t = a * U - a * (a - 1) / 2 + (b - a);  // upper-triangle index
      cr = cg[2 * t] * inv;
      ci = i > c ? -cg[2 * t + 1] * inv : i == c ? 0.f : cg[2 * t + 1] * inv;
    }
    const float2 x = sX[i * LD + c], d = sT[i * LD + c];
    sX[i * LD + c] = make_float2(x.x - d.x + cr, x.y - d.y + ci);
  }
  __syncthreads();
  for (uint32_t it = 0; it < iters; ++it) {
    cmm<UP>(sG, sX, [&](int i, int c, float re, float im) {  // 2I - G X
      sT[i * LD + c] = make_float2((i == c && c < (int)U ? 2.f : 0.f) - re, -im);
    });
    __syncthreads();
    cmm<UP>(sX, sT, store(sV));  // X (2I - G X)
    __syncthreads();
    float2* t = sX;
    sX = sV;
    sV = t;
  }
  float ss = 0.f;
  cmm<UP>(sX, sG, [&](int i, int c, float re, float im) {  // |X G - I|_F
    const float d = re - (i == c && c < (int)U ? 1.f : 0.f);
    ss += d * d + im * im;
  });
  for (int e = threadIdx.x; e < (int)(U * U); e += L::NTH) {
    const int i = e / U, c = e % U;
    X_out[o + e] = sX[i * LD + c];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) ss += __shfl_xor(ss, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ss;
  __syncthreads();
  if (threadIdx.x == 0) {
    const float r = sqrtf((red[0] + red[1]) + (red[2] + red[3]));
    res[g] = r;
    flag[g] = (r <= bound) ? 0 : 1;  // NaN fails too
  }
}

// Final W = X A: groups flagged by the residual check get X from the plain sweep of
// their Gram matrix (the fallback), the others use the refined X. X (row-
// major [U][U]) is updated in place for the next slot.
template <int UT>
__global__ __launch_bounds__(lz_threads<UT>()) void finish_kernel(
    const float2* __restrict__ in, const float2* __restrict__ Gm, float2* X,
    const int32_t* __restrict__ flag, float2* w_out, int32_t* info, uint32_t U, uint32_t B) {
  constexpr int NTH = lz_threads<UT>();
  constexpr int LDG = ZfShared<UT>::LDG;
  constexpr int UP = ZfShared<UT>::UP;
  __shared__ ZfShared<UT> sh;
  extern __shared__ float sDyn[];
  float2* part = reinterpret_cast<float2*>(sDyn);
  const uint32_t g = blockIdx.x;
  const float2* A = in + (uint64_t)g * U * B;
  float2* w = w_out + (uint64_t)g * U * B;
  const uint64_t o = (uint64_t)g * U * U;
  int fail = 0;
  if (flag[g]) {
    if (Gm) {  // the Gram matrix computed earlier in the slot (row-major [U][U])
      for (uint32_t e = threadIdx.x; e < (uint32_t)(UP * UP); e += NTH) {
        const uint32_t i = e / UP, c = e % UP;
        sh.G[i * LDG + c] = (i < U && c < U) ? Gm[o + i * U + c] : make_float2(0.f, 0.f);
      }
      __syncthreads();
    } else {
      zf_gram<UT, NTH, false>(A, U, B, sh, nullptr, part);
    }
    fail = zf_sweep<UT, NTH>(sh, U);
    for (uint32_t e = threadIdx.x; e < U * U; e += NTH) {
      const float2 v = sh.G[(e / U) * LDG + e % U];
      X[o + e] = fail ? make_float2(0.f, 0.f) : make_float2(-v.x, -v.y);
    }
  } else {
    for (uint32_t e = threadIdx.x; e < (uint32_t)(UP * UP); e += NTH) {
      const uint32_t i = e / UP, c = e % UP;
      const float2 v = (i < U && c < U) ? X[o + i * U + c] : make_float2(0.f, 0.f);
      sh.G[i * LDG + c] = make_float2(-v.x, -v.y);  // zf_apply reads -G^-1
    }
    __syncthreads();
  }
  if (fail) {
    for (uint32_t e = threadIdx.x; e < U * B; e += NTH) w[e] = make_float2(0.f, 0.f);
  } else {
    zf_apply<UT, NTH, false>(A, U, B, sh, nullptr, w, U, nullptr);
  }
  if (threadIdx.x == 0 && info) info[g] = fail;
}

template <int UT>
hipError_t launch_gram_inv_ut(uint32_t U, uint32_t B, uint32_t batch, const float2* in,
                              float2* G, float2* X, int32_t* info, hipStream_t st) {
  constexpr int NTH = lz_threads<UT>();
  const size_t with_a = zf_dyn_lds<UT, NTH>(U, B, true);
  const bool a_lds = UT == 1 && sizeof(ZfShared<UT>) + with_a <= 120 * 1024;
  const size_t lds = a_lds ? with_a : zf_dyn_lds<UT, NTH>(U, B, false);
  const void* fn = a_lds ? reinterpret_cast<const void*>(&gram_inv_kernel<UT, true>)
                         : reinterpret_cast<const void*>(&gram_inv_kernel<UT, false>);
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  if (a_lds)
    hipLaunchKernelGGL((gram_inv_kernel<UT, true>), dim3(batch), dim3(NTH), lds, st, in, G, X,
                       info, U, B);
  else
    hipLaunchKernelGGL((gram_inv_kernel<UT, false>), dim3(batch), dim3(NTH), lds, st, in, G, X,
                       info, U, B);
  return hipGetLastError();
}

template <int UT>
hipError_t launch_finish_ut(uint32_t U, uint32_t B, uint32_t batch, const float2* in,
                            const float2* Gm, float2* X, const int32_t* flag, float2* w,
                            int32_t* info, hipStream_t st) {
  constexpr int NTH = lz_threads<UT>();
  const size_t lds = zf_dyn_lds<UT, NTH>(U, B, false);
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&finish_kernel<UT>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((finish_kernel<UT>), dim3(batch), dim3(NTH), lds, st, in, Gm, X, flag, w,
                     info, U, B);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_gram_inv(uint32_t U, uint32_t B, uint32_t batch, const float2* in, float2* G,
                           float2* X, int32_t* info, hipStream_t st) {
  if (batch == 0) return hipSuccess;
  if (U <= 16) return launch_gram_inv_ut<1>(U, B, batch, in, G, X, info, st);
  if (U <= 32) return launch_gram_inv_ut<2>(U, B, batch, in, G, X, info, st);
  return launch_gram_inv_ut<4>(U, B, batch, in, G, X, info, st);
}

hipError_t launch_learned_features(uint32_t U, uint32_t batch, const float2* G, const float2* Gp,
                                   const float2* Xp, __bf16* feat, float* scale, hipStream_t st) {
  if (batch == 0) return hipSuccess;
  hipLaunchKernelGGL(features_kernel, dim3(batch), dim3(256), 0, st, G, Gp, Xp, feat, scale, U);
  return hipGetLastError();
}

hipError_t launch_learned_target(uint32_t U, uint32_t batch, const float2* Xt, const float2* Xb,
                                 const float* scale, float* tgt, hipStream_t st) {
  if (batch == 0) return hipSuccess;
  hipLaunchKernelGGL(target_kernel, dim3(batch), dim3(256), 0, st, Xt, Xb, scale, tgt, U);
  return hipGetLastError();
}

template <int UP>
hipError_t launch_refine_up(uint32_t U, uint32_t batch, const float2* G, const float2* Gp,
                            const float2* Xp, const float* corr, uint32_t corr_ld,
                            const float* scale, float2* X_out, float* res, int32_t* flag,
                            float bound, uint32_t iters, hipStream_t st) {
  const size_t lds = (size_t)4 * UP * (UP + 1) * sizeof(float2);
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&refine_kernel<UP>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(refine_kernel<UP>, dim3(batch), dim3(256), lds, st, G, Gp, Xp, corr,
                     corr_ld, scale, X_out, res, flag, bound, iters, U);
  return hipGetLastError();
}

hipError_t launch_learned_refine(uint32_t U, uint32_t batch, const float2* G, const float2* Gp,
                                 const float2* Xp, const float* corr, uint32_t corr_ld,
                                 const float* scale, float2* X_out, float* res, int32_t* flag,
                                 float bound, uint32_t iters, hipStream_t st) {
  if (batch == 0) return hipSuccess;
  if (U <= 16)
    return launch_refine_up<16>(U, batch, G, Gp, Xp, corr, corr_ld, scale, X_out, res, flag,
                                bound, iters, st);
  if (U <= 32)
    return launch_refine_up<32>(U, batch, G, Gp, Xp, corr, corr_ld, scale, X_out, res, flag,
                                bound, iters, st);
  return launch_refine_up<64>(U, batch, G, Gp, Xp, corr, corr_ld, scale, X_out, res, flag,
                              bound, iters, st);
}

hipError_t launch_learned_finish(uint32_t U, uint32_t B, uint32_t batch, const float2* in,
                                 const float2* Gm, float2* X, const int32_t* flag, float2* w,
                                 int32_t* info, hipStream_t st) {
  if (batch == 0) return hipSuccess;
  if (U <= 16) return launch_finish_ut<1>(U, B, batch, in, Gm, X, flag, w, info, st);
  if (U <= 32) return launch_finish_ut<2>(U, B, batch, in, Gm, X, flag, w, info, st);
  return launch_finish_ut<4>(U, B, batch, in, Gm, X, flag, w, info, st);
}

}  // namespace detail
}  // namespace mega
