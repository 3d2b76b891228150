// Launchers of the learned warm start's network kernels (mlp.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mega {
namespace detail {

// C[M][N] (+)= A B (+ bias, activation); see mlp.hip.
struct MlpGemm {
  const __bf16* A;
  int64_t sai, sak;
  const __bf16* B;
  int64_t sbk, sbj;
  uint32_t M, N, K;
  const float* bias;  // [N] or null
  int act;            // 0 identity, 1 GELU (tanh form)
  float* Cf;          // FP32 pre-activation out (or null)
  __bf16* Cb;         // BF16 activation out (or null)
  uint32_t ldc;
  float beta;         // Cf = z + beta * Cf
};

hipError_t mlp_gemm(const MlpGemm& g, hipStream_t st);
hipError_t mlp_act_backward(const float* dY, const float* Z, __bf16* dZ, float* dZf, uint64_t n,
                            int act, hipStream_t st);
hipError_t mlp_col_sum(const float* dZ, float* db, uint32_t M, uint32_t N, hipStream_t st);
hipError_t mlp_mse_grad(const float* Y, const float* T, float* dY, uint64_t n, uint32_t M,
                        hipStream_t st);
hipError_t mlp_adam(float* w, __bf16* wb, const float* g, float* m, float* v, uint64_t n,
                    float lr, float b1, float b2, float eps, uint32_t step, hipStream_t st);
hipError_t mlp_to_bf16(const float* x, __bf16* y, uint64_t n, hipStream_t st);

}  // namespace detail
}  // namespace mega
