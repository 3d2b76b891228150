// mega::LDPCConfig and mega::LDPCDecoder -- the LDPC consumer of the
// detection path's LLRs, drop-in for the reference classes
// (src/operators/ldpc/config.h:27-73, src/operators/ldpc/decoder.h:28-87)
// over libmegadetect's C ABI (mega_ldpc_decode) instead of cuPHY.
//
//   LDPCConfig(mb, bg, Zc, max_iter): info_nodes = 22 (BG1) or 10 (BG2),
//     encoded_bits = (info_nodes - 2 + mb) * Zc, decoded_bits = info_nodes *
//     Zc, punctured_bits = 2 * Zc (config.h:52-72);
//   LDPCDecoder(llr_type, config)(tLLRs, tOut, stream): tLLRs kHalf or kFloat,
//     one column of Zc * (Kb + mb) LLRs (the punctured prefix zero) per code
//     block, code blocks stride(1) apart; tOut kBit, Zc * Kb bits per code
//     block. Decoding is layered normalized min-sum (megadetect.h).
// Argument errors throw std::invalid_argument / std::runtime_error with the
// library's message.
#pragma once

#include <hip/hip_runtime_api.h>

#include <stdexcept>

#include "megadetect.h"
#include "mega/padded_tensor.h"

namespace mega {

struct LDPCConfig {
  uint16_t parity_nodes;     // nRows / mb
  uint16_t base_graph;       // 1 or 2
  uint16_t lifting_factor;   // Zc
  uint16_t max_decode_iter;  // decoding iterations
  uint16_t info_nodes;       // Kb

  uint64_t encoded_bits;
  uint64_t decoded_bits;
  uint64_t punctured_bits;

  static inline uint16_t num_input_cols(uint16_t base_graph) {
    return base_graph == 1 ? 22 : 10;
  }
  static inline uint64_t num_decoded_bits(uint16_t base_graph, uint16_t lifting_factor) {
    return num_input_cols(base_graph) * uint64_t(lifting_factor);
  }
  static inline constexpr uint16_t num_punc_cols() { return 2; }

  LDPCConfig() = default;
  LDPCConfig(uint16_t mb_, uint16_t bg_, uint16_t zc_, uint16_t md_iter_)
      : parity_nodes(mb_), base_graph(bg_), lifting_factor(zc_), max_decode_iter(md_iter_) {
    info_nodes = num_input_cols(base_graph);
    encoded_bits = (info_nodes - num_punc_cols() + parity_nodes) * uint64_t(lifting_factor);
    decoded_bits = info_nodes * uint64_t(lifting_factor);
    punctured_bits = num_punc_cols() * uint64_t(lifting_factor);
  }
};

class LDPCDecoder {
  mega_ldpc_desc base_{};

 public:
  LDPCDecoder() = default;
  LDPCDecoder(CuphyTensor::cuphy_dtype_t llr_type, const LDPCConfig& config) {
    if (llr_type != CuphyTensor::kHalf && llr_type != CuphyTensor::kFloat)
      throw std::invalid_argument("LLRs must be half or float");
    base_.base_graph = config.base_graph;
    base_.lifting_size = config.lifting_factor;
    base_.parity_nodes = config.parity_nodes;
    base_.max_iterations = config.max_decode_iter;
    base_.normalization = 0.f;
    base_.llr_bits = llr_type == CuphyTensor::kHalf ? 16 : 32;
  }

  void operator()(const CuphyTensor& tLLRs, const CuphyTensor& tOut,
                  hipStream_t stream = nullptr) const {
    const uint32_t bits = tLLRs.dType() == CuphyTensor::kHalf ? 16 : 32;
    if (tLLRs.dType() == CuphyTensor::kBit || bits != (uint32_t)base_.llr_bits)
      throw std::invalid_argument("LLR tensor type differs from the decoder's");
    if (tOut.dType() != CuphyTensor::kBit) throw std::invalid_argument("output must be kBit");
    mega_ldpc_desc d = base_;
    d.num_cb = tLLRs.nDim() > 1 ? tLLRs.dim(1) : 1;
    if ((tOut.nDim() > 1 ? tOut.dim(1) : 1) != d.num_cb)
      throw std::invalid_argument("LLR and output tensors hold different code block counts");
    d.llr = tLLRs.ptr();
    d.llr_stride = tLLRs.nDim() > 1 ? tLLRs.stride(1) : tLLRs.dim(0);
    d.out = tOut.ptr<uint32_t>();
    d.out_stride = (tOut.nDim() > 1 ? tOut.stride(1) : (tOut.dim(0) + 31) / 32 * 32) / 32;
    const int rc = mega_ldpc_decode(&d, stream);
    if (rc < 0) throw std::runtime_error(mega_last_error());
  }
};

}  // namespace mega
