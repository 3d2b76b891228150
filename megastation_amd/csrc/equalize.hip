// Fused equalize + max-log demod: x_k = W_{k/sg} y_k, then binary16 LLRs.
//
// Specification: the naive kernel of src/operators/sgemv/equalize_kernel.cuh
// :415-480 (one dot product per (user, subcarrier), demod, store bit j while
// kQ + j < valid_bits), plus the tiled kernel's row bound (offset < ue_stride,
// :286,319). Stores are exact per (user, subcarrier), so a subcarrier group
// that is not a multiple of any tile width never touches its neighbour's
// outputs (the reference's tiled kernel stores zeros there when sg is not a
// multiple of its tile width).
//
// Work decomposition (one workgroup per subcarrier group x slot x symbol
// chunk): W_g (U x B complex, 8UB bytes) is staged in LDS once and applied to
// every (symbol, subcarrier) column of the chunk, so W is read from HBM once
// per slot instead of once per symbol. A thread owns one column: it streams
// y[b, k] down the antennas (lanes = consecutive subcarriers, coalesced) and
// keeps all U accumulators in registers; W[u, b] reads are LDS broadcasts.
#include "demod.h"
#include "internal.h"

namespace mega {
namespace detail {
namespace {

struct EqParams {
  const float2* csi;
  const float2* y;
  _Float16* llr;
  float2* xhat;
  uint32_t U, B, N, sg, S, spb;
  uint64_t ue_stride, cb_size, cb_stride, zeros, valid_bits;
  uint64_t csi_slot, y_sym, y_slot, llr_sym, llr_slot;
  uint32_t store32;  // every LLR offset even and the base 4-byte aligned
};

template <int Q>
__device__ __forceinline__ void store_llrs(const EqParams& p, _Float16* row,
                                           uint64_t k, float re, float im) {
  _Float16 v[Q];
  demod::demod<Q>(re, im, v);
  const uint64_t bit = k * Q;
  const uint64_t off = (bit / p.cb_size) * p.cb_stride + p.zeros + bit % p.cb_size;
  _Float16* dst = row + off;
  if (p.store32 && bit + Q <= p.valid_bits && off + Q <= p.ue_stride) {
#pragma unroll
    for (int j = 0; j < Q / 2; ++j) {
      uint32_t lo = __builtin_bit_cast(uint16_t, v[2 * j]);
      uint32_t hi = __builtin_bit_cast(uint16_t, v[2 * j + 1]);
      reinterpret_cast<uint32_t*>(dst)[j] = lo | (hi << 16);
    }
  } else {
#pragma unroll
    for (int j = 0; j < Q; ++j)
      if (bit + j < p.valid_bits && off + j < p.ue_stride) dst[j] = v[j];
  }
}

template <int UMAX, int Q>
__global__ __launch_bounds__(256) void equalize_valu(const EqParams p) {
  extern __shared__ float2 sW[];  // [B][U] = W_g, column-major as in HBM
  const uint32_t g = blockIdx.x;
  const uint32_t slot = blockIdx.z;
  const uint32_t s0 = blockIdx.y * p.spb;
  const uint32_t ns = min(p.spb, p.S - s0);
  const uint32_t k0 = g * p.sg;
  const uint32_t nk = min(p.sg, p.N - k0);
  const uint32_t U = p.U;

  const float2* w = p.csi + slot * p.csi_slot + (uint64_t)g * U * p.B;
  for (uint32_t e = threadIdx.x; e < U * p.B; e += blockDim.x) sW[e] = w[e];
  __syncthreads();

  const uint32_t ncol = nk * ns;
  for (uint32_t c = threadIdx.x; c < ncol; c += blockDim.x) {
    const uint32_t s = s0 + c / nk;
    const uint64_t k = k0 + c % nk;
    const float2* yp = p.y + slot * p.y_slot + s * p.y_sym + k;
    float ar[UMAX], ai[UMAX];
#pragma unroll
    for (int u = 0; u < UMAX; ++u) ar[u] = ai[u] = 0.f;
#pragma unroll 2
    for (uint32_t b = 0; b < p.B; ++b) {
      const float2 yv = yp[(uint64_t)b * p.N];
      const float2* wc = sW + b * U;
#pragma unroll
      for (int u = 0; u < UMAX; ++u) {
        if (u < (int)U) {
          const float2 a = wc[u];
          ar[u] += a.x * yv.x - a.y * yv.y;
          ai[u] += a.x * yv.y + a.y * yv.x;
        }
      }
    }
    _Float16* row = p.llr + slot * p.llr_slot + s * p.llr_sym;
#pragma unroll
    for (int u = 0; u < UMAX; ++u) {
      if (u < (int)U) {
        store_llrs<Q>(p, row + u * p.ue_stride, k, ar[u], ai[u]);
        if (p.xhat)
          p.xhat[(((uint64_t)slot * p.S + s) * U + u) * p.N + k] =
              make_float2(ar[u], ai[u]);
      }
    }
  }
}

template <int UMAX, int Q>
hipError_t launch_one(const EqParams& p, dim3 grid, dim3 block, size_t lds,
                      hipStream_t st) {
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute(
        reinterpret_cast<const void*>(&equalize_valu<UMAX, Q>),
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((equalize_valu<UMAX, Q>), grid, block, lds, st, p);
  return hipGetLastError();
}

template <int UMAX>
hipError_t launch_q(const EqParams& p, dim3 grid, dim3 block, size_t lds,
                    uint32_t Q, hipStream_t st) {
  switch (Q) {
    case 2: return launch_one<UMAX, 2>(p, grid, block, lds, st);
    case 4: return launch_one<UMAX, 4>(p, grid, block, lds, st);
    case 6: return launch_one<UMAX, 6>(p, grid, block, lds, st);
    case 8: return launch_one<UMAX, 8>(p, grid, block, lds, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

hipError_t launch_equalize(const EqualizeArgs& a, hipStream_t st,
                           const char** why) {
  if (a.U == 0 || a.B == 0 || a.N == 0 || a.sg == 0 || a.S == 0 || a.slots == 0)
    return hipSuccess;
  const size_t lds = (size_t)a.U * a.B * sizeof(float2);
  int dev = 0, max_lds = 65536;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess)
    e = hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  if (e != hipSuccess) return e;
  if (lds > (size_t)max_lds) {
    *why = "U*B*8 bytes of W exceed the workgroup LDS limit";
    return hipErrorInvalidValue;
  }
  EqParams p;
  p.csi = a.csi;
  p.y = a.y;
  p.llr = a.llr;
  p.xhat = a.xhat;
  p.U = a.U;
  p.B = a.B;
  p.N = a.N;
  p.sg = a.sg;
  p.S = a.S;
  p.ue_stride = a.ue_stride;
  p.cb_size = a.cb_size;
  p.cb_stride = a.cb_stride;
  p.zeros = a.zeros;
  p.valid_bits = a.valid_bits;
  p.csi_slot = a.csi_slot;
  p.y_sym = a.y_sym;
  p.y_slot = a.y_slot;
  p.llr_sym = a.llr_sym;
  p.llr_slot = a.llr_slot;
  p.store32 = ((reinterpret_cast<uintptr_t>(a.llr) & 3) == 0) &&
              (a.ue_stride % 2 == 0) && (a.cb_stride % 2 == 0) &&
              (a.zeros % 2 == 0) && (a.cb_size % 2 == 0) &&
              (a.llr_sym % 2 == 0) && (a.llr_slot % 2 == 0);
  // Split the symbols of a slot over workgroups until the grid fills the GPU.
  uint32_t spb = a.S;
  while (spb > 1 && (uint64_t)a.G * a.slots * ((a.S + spb - 1) / spb) < 2048)
    spb = (spb + 1) / 2;
  p.spb = spb;
  const uint32_t ncol = a.sg * spb;
  const uint32_t threads = ncol >= 256 ? 256 : (ncol + 63) / 64 * 64;
  dim3 grid(a.G, (a.S + spb - 1) / spb, a.slots);
  dim3 block(threads);
  if (a.U <= 8) return launch_q<8>(p, grid, block, lds, a.Q, st);
  if (a.U <= 16) return launch_q<16>(p, grid, block, lds, a.Q, st);
  if (a.U <= 32) return launch_q<32>(p, grid, block, lds, a.Q, st);
  if (a.U <= 64) return launch_q<64>(p, grid, block, lds, a.Q, st);
  *why = "more than 64 users";
  return hipErrorInvalidValue;
}

}  // namespace detail
}  // namespace mega
