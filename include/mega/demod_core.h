// Max-log QAM demodulation to binary16 LLRs, for device and host code.
//
// Restates src/operators/modulation/demodulation.h:19-67 operation by
// operation so the FP16 results are bit-identical:
//   QPSK      llr = half((c * -20) * sqrtf(2))                      (:30-33)
//   M-QAM     l0 = half(100 re), l1 = half(100 im),
//             l(2i+2) = T_i - |l(2i)|, l(2i+3) = T_i - |l(2i+1)|    (:35-67)
// where T_i = half(int / sqrtf(M-dependent)) and "-" is a binary16
// subtraction rounded to nearest even. LLR > 0 means bit 0.
//
// sqrtf(2), sqrtf(10), sqrtf(42) and sqrtf(170) are spelled as their correctly
// rounded float values so no run-time square root is ever evaluated; the
// thresholds are then the same IEEE float divisions the reference writes,
// folded at compile time.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mega {
namespace demod {

using half_t = _Float16;

constexpr float kSqrt2 = 0x1.6a09e6p+0f;
constexpr float kSqrt10 = 0x1.94c584p+1f;
constexpr float kSqrt42 = 0x1.9ec474p+2f;
constexpr float kSqrt170 = 0x1.a13a9cp+3f;

__host__ __device__ __forceinline__ half_t habs(half_t x) {
  return x < (half_t)0.0f ? (half_t)(-x) : x;
}

// Rounds a float product to binary16 in two steps (product to float, then
// float to half), as the reference does. On gfx950 the compiler otherwise
// folds the multiply and the conversion into one mixed-precision FMA
// (v_fma_mixlo_f16) that rounds the exact product once, which differs from
// the reference on rare double-rounding cases; the empty register constraint
// pins the float product in a VGPR so the plain v_cvt_f16_f32 is used.
__host__ __device__ __forceinline__ half_t to_half(float p) {
#if defined(__HIP_DEVICE_COMPILE__)
  __asm__("" : "+v"(p));
#endif
  return (half_t)p;
}

// Writes Q LLRs (Q in {2,4,6,8}); other Q write nothing.
template <int Q>
__host__ __device__ __forceinline__ void demod(float re, float im,
                                               half_t* llr) {
  if constexpr (Q == 2) {
    llr[0] = to_half(re * -20.f * kSqrt2);
    llr[1] = to_half(im * -20.f * kSqrt2);
  } else if constexpr (Q == 4) {
    const half_t t = (half_t)(200.f / kSqrt10);
    llr[0] = to_half(100.f * re);
    llr[1] = to_half(100.f * im);
    llr[2] = t - habs(llr[0]);
    llr[3] = t - habs(llr[1]);
  } else if constexpr (Q == 6) {
    const half_t t1 = (half_t)(400.f / kSqrt42);
    const half_t t2 = (half_t)(200.f / kSqrt42);
    llr[0] = to_half(100.f * re);
    llr[1] = to_half(100.f * im);
    llr[2] = t1 - habs(llr[0]);
    llr[3] = t1 - habs(llr[1]);
    llr[4] = t2 - habs(llr[2]);
    llr[5] = t2 - habs(llr[3]);
  } else if constexpr (Q == 8) {
    const half_t t1 = (half_t)(800.f / kSqrt170);
    const half_t t2 = (half_t)(400.f / kSqrt170);
    const half_t t3 = (half_t)(200.f / kSqrt170);
    llr[0] = to_half(100.f * re);
    llr[1] = to_half(100.f * im);
    llr[2] = t1 - habs(llr[0]);
    llr[3] = t1 - habs(llr[1]);
    llr[4] = t2 - habs(llr[2]);
    llr[5] = t2 - habs(llr[3]);
    llr[6] = t3 - habs(llr[4]);
    llr[7] = t3 - habs(llr[5]);
  }
}

// Runtime-Q form (host helpers and tests). Returns false for bad Q.
__host__ __device__ inline bool demod_any(float re, float im, int Q,
                                          half_t* llr) {
  switch (Q) {
    case 2: demod<2>(re, im, llr); return true;
    case 4: demod<4>(re, im, llr); return true;
    case 6: demod<6>(re, im, llr); return true;
    case 8: demod<8>(re, im, llr); return true;
    default: return false;
  }
}

}  // namespace demod
}  // namespace mega
