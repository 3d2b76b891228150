// mega::UplinkFFT / mega::PilotFFT / mega::PilotSign / mega::DownlinkIFFT —
// the FFT stages, drop-in for the reference classes
// (src/operators/fft/fft_op.h:27-113) over
// libmegadetect's C ABI: the constructor keeps the reference's arguments,
// operator() runs the fused load + FFT + store kernel (mega_uplink_fft /
// mega_pilot_fft / mega_downlink_ifft) on `stream`. QuanT is short (16-bit samples) or char
// (8-bit). PilotSign::init builds the pilot table on the host
// (mega_pilot_sign) and copies it to each device, as the reference does;
// PilotFFT reads the current device's table. Failures throw
// std::runtime_error with the library's message.
#pragma once

#include <hip/hip_runtime_api.h>

#include <stdexcept>
#include <type_traits>
#include <vector>

#include "megadetect.h"
#include "mega/matrix.h"
#include "mega/mega_complex.h"

namespace mega {

struct PilotSign {
  static inline std::vector<Complex*> pilot_sign_ptrs;  // per device

  static void init(uint32_t ofdm_num, uint8_t num_device = 1) {
    std::vector<float> table(2 * (size_t)ofdm_num);
    if (mega_pilot_sign(ofdm_num, table.data()) < 0) throw std::runtime_error(mega_last_error());
    destroy();
    pilot_sign_ptrs.assign(num_device, nullptr);
    int cur = 0;
    (void)hipGetDevice(&cur);
    for (uint8_t i = 0; i < num_device; ++i) {
      if (hipSetDevice(i) != hipSuccess ||
          hipMalloc(&pilot_sign_ptrs[i], table.size() * sizeof(float)) != hipSuccess ||
          hipMemcpy(pilot_sign_ptrs[i], table.data(), table.size() * sizeof(float),
                    hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipSetDevice(cur);
        throw std::runtime_error("PilotSign::init: device allocation or copy failed");
      }
    }
    (void)hipSetDevice(cur);
  }

  static void destroy() {
    for (auto& p : pilot_sign_ptrs)
      if (p) {
        (void)hipFree(p);
        p = nullptr;
      }
  }
};

namespace fft_detail {
template <typename QuanT>
constexpr int32_t quant_bits() {
  static_assert(std::is_same_v<QuanT, short> || std::is_same_v<QuanT, char>,
                "samples are short or char pairs");
  return std::is_same_v<QuanT, short> ? 16 : 8;
}
}  // namespace fft_detail

template <typename QuanT = short>
class UplinkFFT {
 public:
  UplinkFFT() = default;
  UplinkFFT(uint32_t ofdm_start, uint32_t ofdm_ca, uint32_t ofdm_num, uint32_t bs_num)
      : start_(ofdm_start), ca_(ofdm_ca), num_(ofdm_num), bs_(bs_num) {}

  // in: [bs][ofdm_ca] QuanT (re, im) pairs; out: [bs][ofdm_num] Complex
  void operator()(const Matrix& in, const Matrix& out, hipStream_t stream = nullptr) const {
    mega_fft_desc d{start_, ca_, num_, bs_, 1, fft_detail::quant_bits<QuanT>(), in.ptr(), 0,
                    out.ptr(), 0};
    if (mega_uplink_fft(&d, stream) < 0) throw std::runtime_error(mega_last_error());
  }

 private:
  uint32_t start_ = 0, ca_ = 0, num_ = 0, bs_ = 0;
};

template <typename QuanT = short>
class PilotFFT {
 public:
  PilotFFT() = default;
  PilotFFT(uint32_t ofdm_start, uint32_t ofdm_ca, uint32_t ofdm_num, uint32_t bs_num,
           uint32_t ue_num, uint32_t sc_group)
      : start_(ofdm_start), ca_(ofdm_ca), num_(ofdm_num), bs_(bs_num), ue_(ue_num),
        sg_(sc_group) {}

  // in: [bs][ofdm_ca] QuanT pairs; out: CSI [ceil(ofdm_num / sc_group)][bs][ue]
  void operator()(const Matrix& in, const Matrix& out, hipStream_t stream = nullptr) const {
    (*this)(in, out, 0, stream);
  }
  void operator()(const Matrix& in, const Matrix& out, uint32_t ue_start,
                  hipStream_t stream = nullptr) const {
    int dev = 0;
    (void)hipGetDevice(&dev);
    if ((size_t)dev >= PilotSign::pilot_sign_ptrs.size() || !PilotSign::pilot_sign_ptrs[dev])
      throw std::runtime_error("PilotFFT: PilotSign::init has not run for this device");
    mega_fft_desc d{start_, ca_, num_, bs_, 1, fft_detail::quant_bits<QuanT>(), in.ptr(), 0,
                    out.ptr(), 0};
    if (mega_pilot_fft(&d, ue_, sg_, ue_start, PilotSign::pilot_sign_ptrs[dev], stream) < 0)
      throw std::runtime_error(mega_last_error());
  }

 private:
  uint32_t start_ = 0, ca_ = 0, num_ = 0, bs_ = 0, ue_ = 0, sg_ = 0;
};

template <typename QuanT = short>
class DownlinkIFFT {
 public:
  DownlinkIFFT() = default;
  DownlinkIFFT(uint32_t ofdm_start, uint32_t ofdm_ca, uint32_t ofdm_num, uint32_t bs_num)
      : start_(ofdm_start), ca_(ofdm_ca), num_(ofdm_num), bs_(bs_num) {}

  // in: [bs][ofdm_ca] Complex; out: [bs][ofdm_ca] QuanT (re, im) pairs
  void operator()(const Matrix& in, const Matrix& out, hipStream_t stream = nullptr) const {
    mega_fft_desc d{start_, ca_, num_, bs_, 1, fft_detail::quant_bits<QuanT>(), in.ptr(), 0,
                    out.ptr(), 0};
    if (mega_downlink_ifft(&d, stream) < 0) throw std::runtime_error(mega_last_error());
  }

 private:
  uint32_t start_ = 0, ca_ = 0, num_ = 0, bs_ = 0;
};

}  // namespace mega
