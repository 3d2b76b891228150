// Kernels of the learned warm start's network (a BF16 MLP): a strided GEMM
// on the BF16 matrix cores with FP32 accumulation, the fused bias +
// activation epilogues of the forward pass, their backward, the MSE loss
// gradient and the Adam step. Training and inference run on these kernels
// only; the host side (learned.py) sequences them.
//
// GEMM: C[M][N] = sum_k A(i,k) B(k,j), A(i,k) = A[i*sai + k*sak],
// B(k,j) = B[k*sbk + j*sbj] (any transpose through the strides), C row-major
// FP32 (ldc) and/or BF16 with an epilogue: + bias[j], then GELU (tanh form)
// or identity. 64 x 64 tiles per 256-thread workgroup, K in steps of 32
// staged through LDS as BF16 (the next step's vectors loaded during this
// step's products); each wave owns a 32 x 32 quarter (2 x 2
// v_mfma_f32_16x16x32_bf16 tiles). Small M x N grids split K over more
// workgroups and add the partial sums in a second kernel.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "internal.h"
#include "mlp.h"

namespace mega {
namespace detail {
namespace {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float gelu(float z) {
  const float k0 = 0.7978845608f, k1 = 0.044715f;  // sqrt(2/pi)
  return 0.5f * z * (1.f + tanhf(k0 * (z + k1 * z * z * z)));
}

__device__ __forceinline__ float gelu_grad(float z) {
  const float k0 = 0.7978845608f, k1 = 0.044715f;
  const float u = k0 * (z + k1 * z * z * z);
  const float t = tanhf(u);
  return 0.5f * (1.f + t) + 0.5f * z * (1.f - t * t) * k0 * (1.f + 3.f * k1 * z * z);
}

constexpr int TM = 64, TN = 64, TK = 32;

// Bias, activation and the FP32 / BF16 stores of one output element.
__device__ __forceinline__ void gemm_out(float acc, uint32_t i, uint32_t j,
                                         const float* __restrict__ bias, int act, float* Cf,
                                         __bf16* Cb, uint32_t ldc, float beta) {
  const float z = acc + (bias ? bias[j] : 0.f);
  if (Cf) {
    const uint64_t o = (uint64_t)i * ldc + j;
    Cf[o] = beta != 0.f ? z + beta * Cf[o] : z;  // FP32 output (pre-activation)
  }
  if (Cb) Cb[(uint64_t)i * ldc + j] = (__bf16)(act == 1 ? gelu(z) : z);
}

// blockIdx.z = K split: K range [z * kchunk, (z+1) * kchunk). With part, the
// split's raw sums go to part[z][M][N] (gemm_reduce adds the splits in order
// and applies the epilogue); else the epilogue is applied here.
__global__ __launch_bounds__(256) void gemm_bf16(const __bf16* __restrict__ A, int64_t sai,
                                                 int64_t sak, const __bf16* __restrict__ B,
                                                 int64_t sbk, int64_t sbj, uint32_t M,
                                                 uint32_t N, uint32_t K, uint32_t kchunk,
                                                 const float* __restrict__ bias, int act,
                                                 float* __restrict__ Cf, __bf16* __restrict__ Cb,
                                                 uint32_t ldc, float beta, float* part) {
  __shared__ __bf16 sA[TM][TK + 8];
  __shared__ __bf16 sB[TN][TK + 8];  // transposed: sB[j][k]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int i16 = lane & 15, q4 = lane >> 4;
  const uint32_t m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  const uint32_t kb0 = blockIdx.z * kchunk, kb1 = min(K, kb0 + kchunk);
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  v4f acc[2][2] = {};
  const bool a_kfast = sak == 1, b_jfast = sbj == 1;
  // 16-byte staging needs the rows 16-byte aligned
  const bool a_vec = a_kfast && (sai % 8) == 0 && (reinterpret_cast<uintptr_t>(A) % 16) == 0;
  const bool b_vec = b_jfast && (sbk % 8) == 0 && (reinterpret_cast<uintptr_t>(B) % 16) == 0;
  // Whole 16-byte vectors when the unit-stride dimension is the tile's
  // contiguous one and the tile lies inside the matrix (loaded one K step
  // ahead into registers); else element by element with consecutive threads
  // on the unit-stride dimension.
  const int ai = tid / 4, ak = (tid % 4) * 8;   // A vector slot
  const int bk = tid / 8, bj = (tid % 8) * 8;   // B vector slot
  auto a_fast = [&](uint32_t k0) { return a_vec && m0 + TM <= M && k0 + TK <= kb1; };
  auto b_fast = [&](uint32_t k0) { return b_vec && n0 + TN <= N && k0 + TK <= kb1; };
  uint4 ra = {}, rb = {};
  auto fetch = [&](uint32_t k0) {
    if (a_fast(k0)) ra = *reinterpret_cast<const uint4*>(A + (uint64_t)(m0 + ai) * sai + k0 + ak);
    if (b_fast(k0)) rb = *reinterpret_cast<const uint4*>(B + (uint64_t)(k0 + bk) * sbk + n0 + bj);
  };
  fetch(kb0);
  for (uint32_t k0 = kb0; k0 < kb1; k0 += TK) {
    if (a_fast(k0)) {
      *reinterpret_cast<uint4*>(&sA[ai][ak]) = ra;
    } else {
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const int e = tid + 256 * r;  // 0..2047
        const int ii = a_kfast ? e / TK : e % TM, kk = a_kfast ? e % TK : e / TM;
        const uint32_t gi = m0 + ii, gk = k0 + kk;
        sA[ii][kk] = (gi < M && gk < kb1) ? A[gi * sai + gk * sak] : (__bf16)0.f;
      }
    }
    if (b_fast(k0)) {
      const __bf16* e8 = reinterpret_cast<const __bf16*>(&rb);
#pragma unroll
      for (int q = 0; q < 8; ++q) sB[bj + q][bk] = e8[q];
    } else {
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const int e = tid + 256 * r;
        const int jj = b_jfast ? e % TN : e / TK, kb = b_jfast ? e / TN : e % TK;
        const uint32_t gj = n0 + jj, gkb = k0 + kb;
        sB[jj][kb] = (gj < N && gkb < kb1) ? B[gkb * sbk + gj * sbj] : (__bf16)0.f;
      }
    }
    __syncthreads();
    if (k0 + TK < kb1) fetch(k0 + TK);  // in flight during this step's products
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      v8bf av;
#pragma unroll
      for (int v = 0; v < 8; ++v) av[v] = sA[wm + 16 * a + i16][8 * q4 + v];
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        v8bf bv;
#pragma unroll
        for (int v = 0; v < 8; ++v) bv[v] = sB[wn + 16 * b + i16][8 * q4 + v];
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[a][b], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // D lane l, reg r = row 4*(l/16)+r, column l%16 of each 16 x 16 tile
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const uint32_t i = m0 + wm + 16 * a + 4 * q4 + r;
        const uint32_t j = n0 + wn + 16 * b + i16;
        if (i >= M || j >= N) continue;
        if (part)
          part[((uint64_t)blockIdx.z * M + i) * N + j] = acc[a][b][r];
        else
          gemm_out(acc[a][b][r], i, j, bias, act, Cf, Cb, ldc, beta);
      }
}

// Sum of the K splits (in split order) and the epilogue.
__global__ void gemm_reduce(const float* __restrict__ part, uint32_t splits, uint32_t M,
                            uint32_t N, const float* __restrict__ bias, int act,
                            float* __restrict__ Cf, __bf16* __restrict__ Cb, uint32_t ldc,
                            float beta) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t mn = (uint64_t)M * N;
  if (e >= mn) return;
  float s = part[e];
  for (uint32_t z = 1; z < splits; ++z) s += part[z * mn + e];
  gemm_out(s, (uint32_t)(e / N), (uint32_t)(e % N), bias, act, Cf, Cb, ldc, beta);
}

// dZ = dY * act'(Z) (BF16 out), Z the saved FP32 pre-activation.
__global__ void act_backward(const float* __restrict__ dY, const float* __restrict__ Z,
                             __bf16* __restrict__ dZ, float* __restrict__ dZf, uint64_t n, int act) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = act == 1 ? dY[i] * gelu_grad(Z[i]) : dY[i];
  dZ[i] = (__bf16)g;
  if (dZf) dZf[i] = g;
}

// db[j] = sum_i dZ[i][j] (FP32): 64 columns per workgroup, the rows split
// over 4 waves (coalesced row segments), partial sums added through LDS.
__global__ __launch_bounds__(256) void col_sum(const float* __restrict__ dZ,
                                               float* __restrict__ db, uint32_t M, uint32_t N) {
  __shared__ float part[4][64];
  const uint32_t j = blockIdx.x * 64 + (threadIdx.x & 63);
  const uint32_t w = threadIdx.x >> 6;
  float s = 0.f;
  if (j < N)
    for (uint32_t i = w; i < M; i += 4) s += dZ[(uint64_t)i * N + j];
  part[w][threadIdx.x & 63] = s;
  __syncthreads();
  if (w == 0 && j < N) db[j] = (part[0][threadIdx.x] + part[1][threadIdx.x]) +
                                (part[2][threadIdx.x] + part[3][threadIdx.x]);
}

// Gradient (y - t)/M of the MSE loss 0.5/M * sum (y - t)^2.
__global__ void mse_grad(const float* __restrict__ Y, const float* __restrict__ T,
                         float* __restrict__ dY, uint64_t n, float inv_m) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  dY[i] = (Y[i] - T[i]) * inv_m;
}

// Adam on FP32 master weights, BF16 copy for the GEMMs.
__global__ void adam_step(float* __restrict__ w, __bf16* __restrict__ wb, const float* __restrict__ g,
                          float* __restrict__ m, float* __restrict__ v, uint64_t n, float lr,
                          float b1, float b2, float eps, float c1, float c2) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wb) wb[i] = (__bf16)wi;
}

__global__ void to_bf16(const float* __restrict__ x, __bf16* __restrict__ y, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (__bf16)x[i];
}

uint32_t blocks_for(uint64_t n, uint32_t t) { return (uint32_t)((n + t - 1) / t); }

}  // namespace

// K splits: enough workgroups to cover the CUs (>= 256) while every split
// keeps at least 8 K steps. The partial sums live in the library's workspace
// of the (device, stream) (stream_workspace): work queued on one stream runs
// in order and shares it safely, streams never share one. It grows on demand
// with hipFree/hipMalloc, so a split GEMM is refused while its stream is
// being captured into a graph.
hipError_t mlp_gemm(const MlpGemm& g, hipStream_t st) {
  if (g.M == 0 || g.N == 0) return hipSuccess;
  const uint32_t tiles = ((g.N + TN - 1) / TN) * ((g.M + TM - 1) / TM);
  const uint32_t ksteps = (g.K + TK - 1) / TK;
  uint32_t splits = 1;
  while (tiles * splits < 256 && ksteps / (splits + 1) >= 8) ++splits;
  const uint32_t kchunk = ((ksteps + splits - 1) / splits) * TK;
  splits = (g.K + kchunk - 1) / kchunk;
  float* part = nullptr;
  if (splits > 1) {
    void* w = nullptr;
    const hipError_t e =
        stream_workspace(st, (size_t)splits * g.M * g.N * sizeof(float), &w);
    if (e != hipSuccess) return e;
    part = static_cast<float*>(w);
  }
  const dim3 grid((g.N + TN - 1) / TN, (g.M + TM - 1) / TM, splits);
  hipLaunchKernelGGL(gemm_bf16, grid, dim3(256), 0, st, g.A, g.sai, g.sak, g.B, g.sbk, g.sbj,
                     g.M, g.N, g.K, splits > 1 ? kchunk : g.K, g.bias, g.act, g.Cf, g.Cb, g.ldc,
                     g.beta, part);
  if (splits > 1) {
    const uint64_t mn = (uint64_t)g.M * g.N;
    hipLaunchKernelGGL(gemm_reduce, dim3(blocks_for(mn, 256)), dim3(256), 0, st, part, splits,
                       g.M, g.N, g.bias, g.act, g.Cf, g.Cb, g.ldc, g.beta);
  }
  return hipGetLastError();
}

hipError_t mlp_act_backward(const float* dY, const float* Z, __bf16* dZ, float* dZf, uint64_t n,
                            int act, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(act_backward, dim3(blocks_for(n, 256)), dim3(256), 0, st, dY, Z, dZ, dZf, n,
                     act);
  return hipGetLastError();
}

hipError_t mlp_col_sum(const float* dZ, float* db, uint32_t M, uint32_t N, hipStream_t st) {
  if (N == 0) return hipSuccess;
  hipLaunchKernelGGL(col_sum, dim3(blocks_for(N, 64)), dim3(256), 0, st, dZ, db, M, N);
  return hipGetLastError();
}

hipError_t mlp_mse_grad(const float* Y, const float* T, float* dY, uint64_t n, uint32_t M,
                        hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(mse_grad, dim3(blocks_for(n, 256)), dim3(256), 0, st, Y, T, dY, n,
                     1.f / (float)M);
  return hipGetLastError();
}

hipError_t mlp_adam(float* w, __bf16* wb, const float* g, float* m, float* v, uint64_t n,
                    float lr, float b1, float b2, float eps, uint32_t step, hipStream_t st) {
  if (n == 0) return hipSuccess;
  float c1 = 1.f, c2 = 1.f;
  for (uint32_t i = 0; i < step; ++i) {
    c1 *= b1;
    c2 *= b2;
  }
  c1 = 1.f - c1;
  c2 = 1.f - c2;
  hipLaunchKernelGGL(adam_step, dim3(blocks_for(n, 256)), dim3(256), 0, st, w, wb, g, m, v, n,
                     lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

hipError_t mlp_to_bf16(const float* x, __bf16* y, uint64_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(to_bf16, dim3(blocks_for(n, 256)), dim3(256), 0, st, x, y, n);
  return hipGetLastError();
}

}  // namespace detail
}  // namespace mega
