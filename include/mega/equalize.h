// mega::Equalize — fused equalize + max-log demod, drop-in for the reference
// class (src/operators/sgemv/equalize.h:25-74) over libmegadetect's C ABI.
//
//   Equalize::equalize(csi, bsVec, ueVec, csi_skip, mod_order, cb_size, zeros,
//                      cb_blocks, stream)
// with U = csi.dim(0), B = csi.dim(1), N = bsVec.dim(0),
// cb_stride = ueVec.stride(1), ue_stride = cb_stride * cb_blocks,
// valid_bits = cb_size * cb_blocks (equalize.cu:53-67). An unsupported
// mod_order writes nothing and does not throw, as in the reference. A CSI
// buffer with fewer than ceil(N / csi_skip) groups throws (the reference would
// read past it); other failures throw std::runtime_error with the library's
// message.
#pragma once

#include <hip/hip_runtime_api.h>

#include <stdexcept>

#include "megadetect.h"
#include "mega/matrix.h"

namespace mega {

class Equalize {
 public:
  static void equalize(const Matrix& csi, const Matrix& bsVec, const Matrix& ueVec,
                       uint64_t csi_skip, uint8_t mod_order, uint64_t cb_size,
                       uint64_t zeros, uint16_t cb_blocks, hipStream_t stream = nullptr) {
    const uint64_t U = csi.dim(0), B = csi.dim(1), N = bsVec.dim(0);
    const uint64_t groups = csi.nDim() == 3 ? csi.dim(2) : 1;
    if (csi_skip && groups < (N + csi_skip - 1) / csi_skip)
      throw std::runtime_error("CSI holds fewer groups than ceil(N / csi_skip)");
    const uint64_t cb_stride = ueVec.stride(1);
    const uint64_t ue_stride = cb_stride * cb_blocks;
    const uint64_t valid_bits = cb_size * cb_blocks;
    const int rc = mega_equalize(U, B, N, csi.ptr(), bsVec.ptr(), ueVec.ptr(), ue_stride,
                                 csi_skip, mod_order, cb_size, cb_stride, zeros,
                                 valid_bits, stream);
    if (rc == MEGA_ERR_UNSUPPORTED && !(mod_order == 2 || mod_order == 4 ||
                                        mod_order == 6 || mod_order == 8))
      return;
    if (rc < 0) throw std::runtime_error(mega_last_error());
  }
};

}  // namespace mega
