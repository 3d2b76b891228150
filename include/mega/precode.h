// mega::Precode — downlink precode, drop-in for the reference class
// (src/operators/sgemv/precode.h:23-62) over libmegadetect's C ABI.
//
//   Precode::precode(precoded, ueVec, bsVec, precode_skip, bs_start, stream)
// with B = precoded.dim(0), U = precoded.dim(1), N = ueVec.dim(0) and
// bs_stride = bsVec.stride(1) (precode.cu:59-69): per subcarrier k of group
// k / precode_skip, bsVec[b][(k + bs_start + bs_stride/2) % bs_stride] =
// sum_u precoded_g[b][u] ueVec[u][k], for the N / precode_skip full groups.
// Failures throw std::runtime_error with the library's message.
#pragma once

#include <hip/hip_runtime_api.h>

#include <stdexcept>

#include "megadetect.h"
#include "mega/matrix.h"
#include "mega/mega_complex.h"

namespace mega {

class Precode {
 public:
  static void precode(uint64_t problem_size_row, uint64_t problem_size_col,
                      uint64_t batch_count, const Complex* bPrecoded, const Complex* bUeVec,
                      Complex* bBsVec, uint64_t bs_stride, uint64_t precode_skip,
                      uint64_t bs_start, hipStream_t stream = nullptr) {
    const int rc = mega_precode(problem_size_row, problem_size_col, batch_count, bPrecoded,
                                bUeVec, bBsVec, bs_stride, precode_skip, bs_start, stream);
    if (rc < 0) throw std::runtime_error(mega_last_error());
  }

  static void precode(const Matrix& precoded, const Matrix& ueVec, const Matrix& bsVec,
                      uint64_t precode_skip, uint64_t bs_start, hipStream_t stream = nullptr) {
    precode(precoded.dim(0), precoded.dim(1), ueVec.dim(0), precoded.ptr<Complex>(),
            ueVec.ptr<Complex>(), bsVec.ptr<Complex>(), bsVec.stride(1), precode_skip,
            bs_start, stream);
  }
};

}  // namespace mega
