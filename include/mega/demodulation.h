// Max-log demodulation functions with the reference signatures
// (src/operators/modulation/demodulation.h:19-75), host and device:
//   demodQPSK / demod16QAM / demod64QAM / demod256QAM(const Complex&, half[])
// half is _Float16 (binary16). Results are bit-identical to the fused kernel's
// (same arithmetic, mega/demod_core.h).
#pragma once

#include "mega/demod_core.h"
#include "mega/mega_complex.h"

#define SCALE_BYTE_QPSK 20
#define SCALE_BYTE_QAM16 100
#define SCALE_BYTE_QAM64 100
#define SCALE_BYTE_QAM256 100

namespace mega {

using half = _Float16;

__host__ __device__ inline void demodQPSK(const Complex& c, half llr[]) {
  demod::demod<2>(c.re, c.im, llr);
}
__host__ __device__ inline void demod16QAM(const Complex& c, half llr[]) {
  demod::demod<4>(c.re, c.im, llr);
}
__host__ __device__ inline void demod64QAM(const Complex& c, half llr[]) {
  demod::demod<6>(c.re, c.im, llr);
}
__host__ __device__ inline void demod256QAM(const Complex& c, half llr[]) {
  demod::demod<8>(c.re, c.im, llr);
}

typedef void (*demodPtr)(const Complex&, half[]);

}  // namespace mega
