// Zero-forcing weights per subcarrier group, W_g = (A_g A_g^H)^-1 A_g with
// A_g = H_g^H (U x B, column-major), and the downlink precoder normalisation.
//
// Reference: src/operators/beamform/beamform.cu:74-122 runs a batched Cgemm
// (Gram), a batched potrf (lower Cholesky) and G*B single-right-hand-side
// potrs solves through cuBLAS/cuSOLVER. Here one workgroup owns one group and
// fuses all three steps:
//   1. Gram: A is streamed through LDS in chunks of kChunk antennas (one
//      contiguous 8*U*kChunk-byte read per chunk); each thread accumulates
//      its Gram entries in registers.
//   2. Cholesky: right-looking, in LDS, with the pivot status recorded per
//      group (0 = ok, k+1 = pivot k not positive; W_g is then zero).
//   3. Solves: a thread owns an antenna column b, holds a = A[:, b] in
//      registers and applies L^-1 then L^-H with L read by LDS broadcast;
//      W[:, b] overwrites A[:, b] (in place, as the reference).
// U is a template bound (8/16/32/64): the loops are unrolled to the bound and
// rows >= U are skipped by a uniform branch.
#include "internal.h"

namespace mega {
namespace detail {
namespace {

constexpr int kThreads = 256;
constexpr int kChunk = 64;

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

template <int UMAX>
__global__ __launch_bounds__(kThreads) void zf_kernel(
    const float2* in, float2* out, int32_t* info, uint32_t U,
    uint32_t B) {
  constexpr int kEpt = (UMAX * UMAX + kThreads - 1) / kThreads;
  __shared__ float2 sL[UMAX * UMAX];  // Gram, then L (row-major, ld UMAX)
  __shared__ float sInv[UMAX];        // 1 / L[k][k]
  __shared__ float2 sA[kChunk * UMAX];
  __shared__ int sFail;

  const uint32_t g = blockIdx.x;
  const float2* A = in + (uint64_t)g * U * B;
  float2* W = out + (uint64_t)g * U * B;
  const int t = threadIdx.x;

  // ---- 1. Gram G[i][j] = sum_b A[i,b] conj(A[j,b]) -------------------------
  float2 acc[kEpt];
#pragma unroll
  for (int r = 0; r < kEpt; ++r) acc[r] = make_float2(0.f, 0.f);
  for (uint32_t c0 = 0; c0 < B; c0 += kChunk) {
    const uint32_t nb = min((uint32_t)kChunk, B - c0);
    for (uint32_t e = t; e < nb * U; e += kThreads) sA[e] = A[(uint64_t)c0 * U + e];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kEpt; ++r) {
      const uint32_t e = t + r * kThreads;
      if (e < U * U) {
        const uint32_t i = e % U, j = e / U;
        float2 s = acc[r];
        for (uint32_t b = 0; b < nb; ++b) {
          const float2 x = sA[b * U + i], y = sA[b * U + j];
          s.x += x.x * y.x + x.y * y.y;
          s.y += x.y * y.x - x.x * y.y;
        }
        acc[r] = s;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < kEpt; ++r) {
    const uint32_t e = t + r * kThreads;
    if (e < U * U) sL[(e % U) * UMAX + e / U] = acc[r];
  }
  if (t == 0) sFail = 0;
  __syncthreads();

  // ---- 2. Cholesky, lower, right-looking ------------------------------------
  for (uint32_t k = 0; k < U; ++k) {
    if (t == 0) {
      const float d = sL[k * UMAX + k].x;
      if (!(d > 0.f) || !isfinite(d)) {
        sFail = (int)k + 1;
      } else {
        const float r = sqrtf(d);
        sL[k * UMAX + k] = make_float2(r, 0.f);
        sInv[k] = 1.0f / r;
      }
    }
    __syncthreads();
    if (sFail) break;
    const float inv = sInv[k];
    for (uint32_t i = k + 1 + t; i < U; i += kThreads) {
      float2 v = sL[i * UMAX + k];
      sL[i * UMAX + k] = make_float2(v.x * inv, v.y * inv);
    }
    __syncthreads();
    const uint32_t m = U - k - 1;
    for (uint32_t e = t; e < m * m; e += kThreads) {
      const uint32_t i = k + 1 + e % m, j = k + 1 + e / m;
      if (j <= i) {
        const float2 l = cmulc(sL[i * UMAX + k], sL[j * UMAX + k]);
        float2 v = sL[i * UMAX + j];
        sL[i * UMAX + j] = make_float2(v.x - l.x, v.y - l.y);
      }
    }
    __syncthreads();
  }
  const int fail = sFail;
  if (t == 0 && info) info[g] = fail;
  if (fail) {
    for (uint32_t e = t; e < U * B; e += kThreads) W[e] = make_float2(0.f, 0.f);
    return;
  }

  // ---- 3. per column: z = L^-1 a, w = L^-H z ---------------------------------
  for (uint32_t b = t; b < B; b += kThreads) {
    float2 v[UMAX];
    const float2* ac = A + (uint64_t)b * U;
#pragma unroll
    for (int i = 0; i < UMAX; ++i)
      v[i] = (i < (int)U) ? ac[i] : make_float2(0.f, 0.f);
#pragma unroll
    for (int i = 0; i < UMAX; ++i) {
      if (i < (int)U) {
        float2 z = v[i];
#pragma unroll
        for (int j = 0; j < i; ++j) {
          const float2 p = cmul(sL[i * UMAX + j], v[j]);
          z.x -= p.x;
          z.y -= p.y;
        }
        const float s = sInv[i];
        v[i] = make_float2(z.x * s, z.y * s);
      }
    }
#pragma unroll
    for (int i = UMAX - 1; i >= 0; --i) {
      if (i < (int)U) {
        float2 z = v[i];
#pragma unroll
        for (int j = i + 1; j < UMAX; ++j) {
          if (j < (int)U) {
            // conj(L[j][i]) * w_j
            const float2 l = sL[j * UMAX + i];
            const float2 p = cmul(make_float2(l.x, -l.y), v[j]);
            z.x -= p.x;
            z.y -= p.y;
          }
        }
        const float s = sInv[i];
        v[i] = make_float2(z.x * s, z.y * s);
      }
    }
    float2* wc = W + (uint64_t)b * U;
#pragma unroll
    for (int i = 0; i < UMAX; ++i)
      if (i < (int)U) wc[i] = v[i];
  }
}

// P_g = W_g^T * (1 / max|W_g|) (beam_utils.cuh:40-80): U x B column-major in,
// B x U column-major out.
__global__ __launch_bounds__(256) void scale_absmax_kernel(
    const float2* __restrict__ in, float2* __restrict__ out, uint32_t U,
    uint32_t B, uint32_t batch) {
  __shared__ float sMax[256 / 64];
  const uint32_t n = U * B;
  for (uint32_t g = blockIdx.x; g < batch; g += gridDim.x) {
    const float2* x = in + (uint64_t)g * n;
    float2* o = out + (uint64_t)g * n;
    float m = 0.f;
    for (uint32_t e = threadIdx.x; e < n; e += blockDim.x) {
      const float2 v = x[e];
      m = fmaxf(m, sqrtf(v.x * v.x + v.y * v.y));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) sMax[threadIdx.x >> 6] = m;
    __syncthreads();
    float mx = 0.f;
    for (uint32_t w = 0; w < (blockDim.x + 63) / 64; ++w) mx = fmaxf(mx, sMax[w]);
    const float s = 1.0f / mx;
    for (uint32_t e = threadIdx.x; e < n; e += blockDim.x) {
      const float2 v = x[e];
      o[(e % U) * B + e / U] = make_float2(v.x * s - v.y * 0.f, v.x * 0.f + v.y * s);
    }
    __syncthreads();
  }
}

}  // namespace

hipError_t launch_zf(uint32_t U, uint32_t B, uint32_t batch, const float2* in,
                     float2* out, int32_t* info, hipStream_t st,
                     const char** why) {
  if (batch == 0) return hipSuccess;
  dim3 grid(batch), block(kThreads);
  if (U <= 8)
    hipLaunchKernelGGL(zf_kernel<8>, grid, block, 0, st, in, out, info, U, B);
  else if (U <= 16)
    hipLaunchKernelGGL(zf_kernel<16>, grid, block, 0, st, in, out, info, U, B);
  else if (U <= 32)
    hipLaunchKernelGGL(zf_kernel<32>, grid, block, 0, st, in, out, info, U, B);
  else if (U <= 64)
    hipLaunchKernelGGL(zf_kernel<64>, grid, block, 0, st, in, out, info, U, B);
  else {
    *why = "more than 64 users";
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_scale_absmax(uint32_t U, uint32_t B, uint32_t batch,
                               const float2* in, float2* out,
                               hipStream_t st) {
  if (batch == 0) return hipSuccess;
  const uint32_t n = U * B;
  const uint32_t threads = n >= 256 ? 256 : (n + 63) / 64 * 64;
  const uint32_t blocks = batch < 65536 ? batch : 65535;
  hipLaunchKernelGGL(scale_absmax_kernel, dim3(blocks), dim3(threads), 0, st, in,
                     out, U, B, batch);
  return hipGetLastError();
}

}  // namespace detail
}  // namespace mega
