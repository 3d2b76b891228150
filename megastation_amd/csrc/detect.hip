// Fused detection: ZF weights and equalize + demod of every symbol of a slot
// in one workgroup per (slot, subcarrier group).
//
// The drop-in path runs beamform_uplink (W written over the CSI buffer,
// beamform.cu:74-122) and then one equalize per symbol
// (equalize.cu:20-51), so W makes a round trip through HBM between the two.
// Here the workgroup computes W_g = (A_g A_g^H)^-1 A_g into LDS (zf_core.h),
// loads its equalize operands from there and streams the slot's y through the
// MFMA GEMM (eq_core.h): HBM traffic is A + y + LLR only. W is also written
// to w_out when the caller wants it (downlink precoding, or to reuse it for
// later symbols of the frame). Results equal ZF followed by equalize of the
// same kernels' arithmetic.
//
// Shapes: up to 16 users (one 16-row Gram tile) and 256 antennas (four waves
// of 64 antennas); other shapes run the two kernels back to back through
// w_out.
#include <cstdlib>
#include <mutex>

#include "zf_core.h"

namespace mega {
namespace detail {
namespace {

// Dynamic LDS: A staged as rows of U+1 complex (W is written over it and
// read from there by the equalize), then one region used first for the Gram
// partials and afterwards for the equalize reduction.
template <int NW>
size_t detect_lds(uint32_t U, uint32_t B, uint32_t MT) {
  const size_t a = (size_t)(U + 1) * B * sizeof(float2);
  const size_t part = zf_dyn_lds<1, 64 * NW>(U, B, false);
  const size_t red = NW > 1 ? (size_t)2 * NW * MT * 256 * sizeof(float) : 0;
  return a + (part > red ? part : red);
}

template <int MT, int KS, int NW, int Q, bool PF>
__global__ __launch_bounds__(64 * NW) void detect_fused(const EqParams p) {
  __shared__ ZfShared<1> sh;
  extern __shared__ float sDyn[];
  float2* sA = reinterpret_cast<float2*>(sDyn);
  float* region = sDyn + 2 * (p.U + 1) * p.B;
  float2* part = reinterpret_cast<float2*>(region);
  float* sRed = region;  // reused after the ZF phase
  const uint32_t wk = threadIdx.x >> 6;
  const uint32_t ldw = p.U + 1;

  const uint32_t logical = xcd_linear(blockIdx.x, gridDim.x);
  const uint32_t g = logical % p.G;
  const uint32_t slot = logical / p.G;
  const uint64_t goff = slot * p.csi_slot + (uint64_t)g * p.U * p.B;
  float2* wout = p.w_out ? p.w_out + goff : nullptr;

  // The first y tile does not depend on W: issue it before the ZF phase.
  float2 ynext[KS];
  eq_load_tile<KS>(p, slot, g, wk, 0, ynext);

  // W overwrites A in LDS (each wave owns whole antenna columns of both).
  const int fail =
      zf_group<1, 64 * NW, true>(p.csi + goff, p.U, p.B, sh, sA, part, sA, ldw, wout);
  if (threadIdx.x == 0 && p.info) p.info[(uint64_t)slot * p.G + g] = fail;
  __syncthreads();

  const WLds<MT, KS> ws{sA, ldw, p.U, p.B, 0, wk};
  const uint32_t nk = min(p.sg, p.N - g * p.sg);
  const uint32_t ntiles = (nk * p.S + 15) / 16;
  eq_tiles<MT, KS, Q, WLds<MT, KS>, PF>(p, ws, slot, g, 0, wk, 0, ntiles, sRed, ynext);
}

template <int MT, int NW, int Q, bool PF>
hipError_t launch_one(const EqParams& p, dim3 grid, size_t lds, hipStream_t st) {
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute(
        reinterpret_cast<const void*>(&detect_fused<MT, 16, NW, Q, PF>),
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((detect_fused<MT, 16, NW, Q, PF>), grid, dim3(64 * NW), lds, st, p);
  return hipGetLastError();
}

// MEGA_DET_PF=0 turns the next-tile y prefetch off (tuning sweeps).
bool detect_prefetch() { return tune(kDetPf) != 0; }

template <int MT, int NW>
hipError_t launch_q(const EqParams& p, dim3 grid, size_t lds, uint32_t Q,
                    hipStream_t st) {
  const bool pf = detect_prefetch();
  switch (Q * 2 + (pf ? 1 : 0)) {
    case 5: return launch_one<MT, NW, 2, true>(p, grid, lds, st);
    case 4: return launch_one<MT, NW, 2, false>(p, grid, lds, st);
    case 9: return launch_one<MT, NW, 4, true>(p, grid, lds, st);
    case 8: return launch_one<MT, NW, 4, false>(p, grid, lds, st);
    case 13: return launch_one<MT, NW, 6, true>(p, grid, lds, st);
    case 12: return launch_one<MT, NW, 6, false>(p, grid, lds, st);
    case 17: return launch_one<MT, NW, 8, true>(p, grid, lds, st);
    case 16: return launch_one<MT, NW, 8, false>(p, grid, lds, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

bool detect_fused_supported(uint32_t U, uint32_t B) { return U <= 16 && B <= 256; }

hipError_t launch_detect(const EqualizeArgs& a, float2* w_out, int32_t* info,
                         hipStream_t st, const char** why, bool may_alloc) {
  if (a.U == 0 || a.B == 0 || a.N == 0 || a.sg == 0 || a.S == 0 || a.slots == 0)
    return hipSuccess;
  if (!detect_fused_supported(a.U, a.B)) {
    if (!w_out) {
      *why = "shape outside the fused kernel (U > 16 or B > 256) needs w_out";
      return hipErrorInvalidValue;
    }
    hipError_t e = hipSuccess;
    if (a.csi_slot == (uint64_t)a.G * a.U * a.B || a.slots == 1) {
      e = launch_zf(a.U, a.B, a.slots * a.G, a.csi, w_out, info, st, why, nullptr, 0,
                    may_alloc);
    } else {  // padded slot stride (w_out has the same layout): one launch per slot
      for (uint32_t s = 0; s < a.slots && e == hipSuccess; ++s)
        e = launch_zf(a.U, a.B, a.G, a.csi + s * a.csi_slot, w_out + s * a.csi_slot,
                      info ? info + (uint64_t)s * a.G : nullptr, st, why, nullptr, 0,
                      may_alloc);
    }
    if (e != hipSuccess) return e;
    EqualizeArgs b = a;
    b.csi = w_out;
    return launch_equalize(b, st, why);
  }
  EqParams p;
  hipError_t e = fill_eq_params(a, p, why);
  if (e != hipSuccess) return e;
  p.w_out = w_out;
  p.info = info;
  const uint32_t MT = a.U <= 8 ? 1 : 2;
  // KS = 16 antenna steps (64 antennas) per wave; 3 waves' worth run as 4
  // (the last wave's antennas are all beyond B and contribute zeros).
  const uint32_t need = (a.B + 63) / 64;
  const uint32_t nw = need == 1 ? 1 : need == 2 ? 2 : 4;
  p.nmg = 1;
  p.nwk = nw;
  p.tiles_per_wg = p.col_tiles;
  p.chunks = 1;
  const uint64_t nblocks = (uint64_t)a.slots * a.G;
  if (nblocks > 0x7fffffffull) {
    *why = "grid too large";
    return hipErrorInvalidValue;
  }
  const size_t lds = nw == 1 ? detect_lds<1>(a.U, a.B, MT)
                    : nw == 2 ? detect_lds<2>(a.U, a.B, MT) : detect_lds<4>(a.U, a.B, MT);
  dim3 grid((uint32_t)nblocks);
  switch (MT * 10 + nw) {
    case 11: return launch_q<1, 1>(p, grid, lds, a.Q, st);
    case 12: return launch_q<1, 2>(p, grid, lds, a.Q, st);
    case 14: return launch_q<1, 4>(p, grid, lds, a.Q, st);
    case 21: return launch_q<2, 1>(p, grid, lds, a.Q, st);
    case 22: return launch_q<2, 2>(p, grid, lds, a.Q, st);
    case 24: return launch_q<2, 4>(p, grid, lds, a.Q, st);
    default:
      *why = "no fused configuration";
      return hipErrorInvalidValue;
  }
}

hipError_t launch_detect_mixed(const EqualizeArgs& a, float2* w_out, void* work,
                               int32_t* info, int32_t* route, hipStream_t st,
                               const char** why) {
  if (a.U == 0 || a.B == 0 || a.N == 0 || a.sg == 0 || a.S == 0 || a.slots == 0)
    return hipSuccess;
  if (!work) {
    *why = "the mixed form needs a work buffer";
    return hipErrorInvalidValue;
  }
  // split planes per slot, in 8-byte units (dense work buffer)
  const uint64_t per_slot = (uint64_t)a.G * a.U * ((a.B + 15) / 16 * 16);
  hipError_t e = hipSuccess;
  if (a.csi_slot == (uint64_t)a.G * a.U * a.B || a.slots == 1) {
    e = launch_zf_mixed(a.U, a.B, a.slots * a.G, a.csi, w_out, work, info, route, st, why);
  } else {
    for (uint32_t s = 0; s < a.slots && e == hipSuccess; ++s)
      e = launch_zf_mixed(a.U, a.B, a.G, a.csi + s * a.csi_slot,
                          w_out ? w_out + s * a.csi_slot : nullptr,
                          static_cast<uint2*>(work) + s * per_slot,  // 2 words per unit
                          info ? info + (uint64_t)s * a.G : nullptr,
                          route ? route + (uint64_t)s * a.G : nullptr, st, why);
  }
  if (e != hipSuccess) return e;
  EqualizeArgs b = a;
  b.csi_slot = per_slot;
  return launch_equalize_mixed(b, work, st, why);
}

// Scratch of the no-allocation entry point: the mixed form's split weight
// planes. Every GPU the library runs on holds its own copy for the life of
// the process. The array's name is not a device address in host code, so
// the address is looked up once per device, with that device current, and
// only that address is handed to kernels. Calls that overlap on one GPU
// share the array and race; the kernels write every word they later read
// (pads included), so nothing needs clearing between calls.
static __device__ __align__(256) uint8_t scratch[kDetectScratchBytes];

hipError_t detect_scratch(void** out, const char** why) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static void* addr[kMaxDevices] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) {
    *why = "device index above 63";
    return hipErrorInvalidDevice;
  }
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      *why = "hipGetSymbolAddress of the scratch array";
      return e != hipSuccess ? e : hipErrorInvalidSymbol;
    }
    addr[dev] = p;
  }
  *out = addr[dev];
  return hipSuccess;
}

}  // namespace detail
}  // namespace mega
