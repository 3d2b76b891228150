// mega::Beamform — zero-forcing beamformer, drop-in for the reference class
// (src/operators/beamform/beamform.h:27-84) over libmegadetect's C ABI.
//
// Same constructor, methods, argument meaning and exceptions:
//   Beamform(num_bs, num_ue, batch_count)
//   beamformer(csi, precode, stream)      W in place + normalised precoder
//   beamform_uplink(csi, stream)          W in place
//   static beamform_downlink(csi, precode, stream)
// with std::runtime_error("Csi & Precode mismatch"), ("Batch size ... exceeds
// the maximum batch size ...") and ("Precode dimension mismatch") as in
// beamform.cu:124-168. Copies share the native handle; the last copy frees
// it (beamform.cu:33-48). hipStream_t replaces cudaStream_t.
// Additionally, info() returns the per-group Cholesky status, which the
// reference computes but never reads.
#pragma once

#include <hip/hip_runtime_api.h>

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "megadetect.h"
#include "mega/matrix.h"

namespace mega {

class Beamform {
  std::shared_ptr<struct mega_beamform> handle_;
  uint32_t num_bs_ = 0, num_ue_ = 0, batch_count_ = 0;

  static void check(int rc) {
    if (rc < 0) throw std::runtime_error(mega_last_error());
  }

 public:
  Beamform() = default;
  Beamform(uint32_t num_bs, uint32_t num_ue, uint32_t batch_count)
      : num_bs_(num_bs), num_ue_(num_ue), batch_count_(batch_count) {
    mega_beamform_t h = nullptr;
    check(mega_beamform_create(num_bs, num_ue, batch_count, &h));
    handle_ = std::shared_ptr<struct mega_beamform>(
        h, [](mega_beamform_t p) { (void)mega_beamform_destroy(p); });
  }
  ~Beamform() = default;

  void beamformer(const Matrix& csi, const Matrix& precode, hipStream_t stream = nullptr) {
    const uint32_t batch = csi.dim(csi.nDim() - 1);
    if (precode.nDim() != csi.nDim() || batch != precode.dim(precode.nDim() - 1))
      throw std::runtime_error("Csi & Precode mismatch");
    if (batch > batch_count_)
      throw std::runtime_error("Batch size exceeds the maximum batch size");
    check(mega_beamformer(handle_.get(), csi.ptr(), precode.ptr(), batch, stream));
  }

  void beamform_uplink(const Matrix& csi, hipStream_t stream = nullptr) {
    const uint32_t batch = csi.dim(csi.nDim() - 1);
    if (batch > batch_count_)
      throw std::runtime_error("Batch size " + std::to_string(batch) +
                               " exceeds the maximum batch size " +
                               std::to_string(batch_count_));
    check(mega_beamform_uplink(handle_.get(), csi.ptr(), batch, stream));
  }

  static void beamform_downlink(const Matrix& csi, const Matrix& precode,
                                hipStream_t stream = nullptr) {
    const uint32_t batch = csi.dim(csi.nDim() - 1);
    const uint32_t num_ue = csi.dim(0), num_bs = csi.dim(1);
    if (precode.nDim() != csi.nDim() || batch != precode.dim(precode.nDim() - 1))
      throw std::runtime_error("Csi & Precode mismatch");
    if (num_bs != precode.dim(0) || num_ue != precode.dim(1))
      throw std::runtime_error("Precode dimension mismatch");
    check(mega_beamform_downlink(csi.ptr(), precode.ptr(), num_ue, num_bs, batch, stream));
  }

  // Per-group status of the last call (0 ok, k+1 pivot k failed); synchronises.
  std::vector<int32_t> info(uint32_t batch, hipStream_t stream = nullptr) const {
    std::vector<int32_t> out(batch);
    check(mega_beamform_info(handle_.get(), out.data(), batch, stream));
    return out;
  }
};

}  // namespace mega
