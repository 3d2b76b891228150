// Zero-forcing weights per subcarrier group, W_g = (A_g A_g^H)^-1 A_g with
// A_g = H_g^H (U x B, column-major), and the downlink precoder normalisation.
//
// Reference: src/operators/beamform/beamform.cu:74-122 runs a batched Cgemm
// (Gram), a batched potrf (lower Cholesky) and G*B single-right-hand-side
// potrs solves through cuBLAS/cuSOLVER. Here one workgroup owns one group and
// fuses the three steps, with both products on the FP32 matrix cores
// (v_mfma_f32_16x16x4_f32):
//   1. Gram G = A A^H, users padded to 16-row tiles, K = 4 antennas per step:
//        Re G += Re A Re A^T + Im A Im A^T,  Im G += Im A Re A^T - Re A Im A^T.
//      For the 16x16x4 instruction the A operand of lane l is X[l%16][l/16]
//      and the B operand is Y[l/16][l%16]^T-wise the same pattern, so one
//      load of A[u = l%16][b = 4s + l/16] (64 contiguous complex for U = 16)
//      feeds both operands of a diagonal tile.
//   2. G^-1 by in-place Gauss-Jordan elimination in LDS (no pivoting: G is
//      Hermitian positive definite, and its pivots are exactly the squared
//      Cholesky diagonals). A pivot not above U * 2^-23 * G_kk marks the group
//      as numerically singular: status k+1 and W_g = 0 (the reference never
//      reads cuSOLVER's status and would emit Inf/NaN).
//   3. W = G^-1 A, M = users, N = 16 antennas, K = 4 users per step:
//        Re W = Re Gi Re A - Im Gi Im A,  Im W = Re Gi Im A + Im Gi Re A.
//      A wave owns whole antenna columns (all user tiles), loads them once
//      and only then stores, so the in-place case (out == in) is safe.
#include "internal.h"

namespace mega {
namespace detail {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4f mfma4(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

constexpr float kPivotTol = 1.1920929e-7f;  // 2^-23, times U

template <int UT>
constexpr int zf_threads() { return UT == 1 ? 64 : 256; }

template <int UT>  // U <= 16 * UT
__global__ __launch_bounds__(zf_threads<UT>()) void zf_mfma(
    const float2* in, float2* out, int32_t* info, uint32_t U, uint32_t B) {
  constexpr int UP = 16 * UT;
  constexpr int LDG = UP + 1;  // padded row (complex elements)
  constexpr int NTH = zf_threads<UT>();
  constexpr int NW = NTH / 64;
  constexpr int EPT = (UP * UP + NTH - 1) / NTH;
  __shared__ float2 sG[UP * LDG];
  __shared__ float sDiag[UP];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int i16 = lane & 15, q4 = lane >> 4;
  const uint32_t g = blockIdx.x;
  const float2* A = in + (uint64_t)g * U * B;
  float2* W = out + (uint64_t)g * U * B;
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
    sDiag[i] = sG[i * LDG + i].x;
  }
  __syncthreads();

  // ---- 2. Gauss-Jordan inverse in place ----------------------------------------
  int fail = 0;
  const float tol = kPivotTol * (float)U;
  for (uint32_t k = 0; k < U; ++k) {
    const float piv = sG[k * LDG + k].x;
    if (!(piv > tol * sDiag[k]) || !isfinite(piv)) {
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
  if (tid == 0 && info) info[g] = fail;
  if (fail) {
    for (uint32_t e = tid; e < U * B; e += NTH) W[e] = make_float2(0.f, 0.f);
    return;
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
          if (u < U) W[u + (uint64_t)U * b] = make_float2(res_r[mi][r], res_i[mi][r]);
        }
    }
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
  if (batch > 0x7fffffffu) {
    *why = "batch too large";
    return hipErrorInvalidValue;
  }
  dim3 grid(batch);
  if (U <= 16)
    hipLaunchKernelGGL(zf_mfma<1>, grid, dim3(zf_threads<1>()), 0, st, in, out, info, U, B);
  else if (U <= 32)
    hipLaunchKernelGGL(zf_mfma<2>, grid, dim3(zf_threads<2>()), 0, st, in, out, info, U, B);
  else if (U <= 64)
    hipLaunchKernelGGL(zf_mfma<4>, grid, dim3(zf_threads<4>()), 0, st, in, out, info, U, B);
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
