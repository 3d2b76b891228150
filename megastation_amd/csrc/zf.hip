// Zero-forcing weights per subcarrier group (one workgroup per group, the
// fused Gram / Gauss-Jordan / MFMA apply of zf_core.h) and the downlink
// precoder normalisation (src/operators/beamform/beam_utils.cuh:40-80).
#include "zf_core.h"

namespace mega {
namespace detail {
namespace {

// Threads per group: one wave for up to 16 users, four waves up to 32, 16
// waves for 33 to 128 (at 64 users the sweep keeps 4 entries a thread, and
// the Gram's 10 tiles and the apply's antenna tiles each go to a wave of
// their own; W is bit-identical to four waves' and the launch at 256 x 64,
// 273 groups, takes 125 us against 166, at 48 users 113 against 142).
// MEGA_ZF_NTH overrides for tuning sweeps: 512 for 17 to 32 users, 256 for
// 33 to 64.
template <int UT>
constexpr int zf_threads() { return UT == 1 ? 64 : UT == 2 ? 256 : 1024; }

// Four-wave groups of 17-32 users are held to 168 registers a lane (no
// spills), so three groups fit a CU instead of two: a ZF launch beside an
// equalize launch, or over two slots' 546 groups, keeps more of them resident
// (256x32 slot pipeline 0.066 against 0.077 ms a slot; one 273-group launch
// alone unchanged).
template <int UT, bool A_LDS, int NTH>  // U <= 16 * UT
__global__ __launch_bounds__(NTH, UT == 2 && NTH == 256 ? 3 : 1) void zf_mfma(
    const float2* in, float2* out, int32_t* info, uint32_t U, uint32_t B) {
  __shared__ ZfShared<UT> sh;
  extern __shared__ float sDyn[];
  float2* sA = reinterpret_cast<float2*>(sDyn);
  float2* part = reinterpret_cast<float2*>(sDyn + (A_LDS ? 2 * (U + 1) * B : 0));
  const uint32_t g = blockIdx.x;
  const int fail = zf_group<UT, NTH, A_LDS>(
      in + (uint64_t)g * U * B, U, B, sh, sA, part, out + (uint64_t)g * U * B, U, nullptr);
  if (threadIdx.x == 0 && info) info[g] = fail;
}

// Static LDS of ZfShared<UT> plus the dynamic part must stay within the
// 160 KB a workgroup may use; A is staged (up to 16 users) only when that
// holds. Wider groups read A from global memory so that several groups fit
// a CU at once.
template <int UT, int NTH>
hipError_t launch_zf_nth(uint32_t U, uint32_t B, uint32_t batch, const float2* in,
                         float2* out, int32_t* info, hipStream_t st) {
  if constexpr (UT == 1) {
    const size_t with_a = zf_dyn_lds<UT, NTH>(U, B, true);
    if (sizeof(ZfShared<UT>) + with_a <= 120 * 1024) {
      if (with_a > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&zf_mfma<UT, true, NTH>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)with_a);
        if (e != hipSuccess) return e;
      }
      hipLaunchKernelGGL((zf_mfma<UT, true, NTH>), dim3(batch), dim3(NTH), with_a, st, in, out,
                         info, U, B);
      return hipGetLastError();
    }
  }
  const size_t lds = zf_dyn_lds<UT, NTH>(U, B, false);
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&zf_mfma<UT, false, NTH>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((zf_mfma<UT, false, NTH>), dim3(batch), dim3(NTH), lds, st, in, out, info,
                     U, B);
  return hipGetLastError();
}

template <int UT>
hipError_t launch_zf_ut(uint32_t U, uint32_t B, uint32_t batch, const float2* in,
                        float2* out, int32_t* info, hipStream_t st) {
  if constexpr (UT == 2) {
    const int nth = tune(kZfNth);
    if (nth == 512) return launch_zf_nth<2, 512>(U, B, batch, in, out, info, st);
  }
  if constexpr (UT == 4) {
    const int nth = tune(kZfNth);
    if (nth == 256) return launch_zf_nth<4, 256>(U, B, batch, in, out, info, st);
  }
  return launch_zf_nth<UT, zf_threads<UT>()>(U, B, batch, in, out, info, st);
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
                     const char** why, void* work, uint64_t work_bytes, bool may_alloc) {
  if (batch == 0) return hipSuccess;
  if (batch > 0x7fffffffu) {
    *why = "batch too large";
    return hipErrorInvalidValue;
  }
  // few wide groups: the three-kernel form when a workspace is at hand
  // (MEGA_ZF_WIDE = 1 never, 2 wherever the shape allows)
  const int wide = tune(kZfWide);
  const uint64_t need = zf_wide_work_bytes(U, B, batch);
  if (need > 0 && wide != 1 && (wide == 2 || zf_wide_pays(U, B, batch))) {
    void* ws = work && work_bytes >= need ? work : nullptr;
    if (!ws && !work && may_alloc && stream_workspace(st, need, &ws) != hipSuccess) {
      ws = nullptr;
      (void)hipGetLastError();  // a refused workspace leaves the one-workgroup form
    }
    if (ws) return launch_zf_wide(U, B, batch, in, out, info, ws, st, why);
  }
  if (U <= 16) return launch_zf_ut<1>(U, B, batch, in, out, info, st);
  if (U <= 32) return launch_zf_ut<2>(U, B, batch, in, out, info, st);
  if (U <= 64) return launch_zf_ut<4>(U, B, batch, in, out, info, st);
  if (U <= 128) return launch_zf_ut<8>(U, B, batch, in, out, info, st);
  *why = "more than 128 users";
  return hipErrorInvalidValue;
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
