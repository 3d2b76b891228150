// mega::PaddedTensor (alias mega::CuphyTensor) — a Matrix whose first
// dimension is padded the way the LDPC decoder's tensors are: the LLR buffer
// the equalizer writes is such a tensor, and its row stride (stride(1)) is the
// cb_stride of the equalize call.
//
// Padding rules, as the reference's cuPHY-backed descriptor produces them
// (src/operators/matrix/cuphy_tensor.h:29-280, pinned by
// test/matrix/test_cuphy_tensor.cc:18-104):
//   kBit tensors pack 32 elements per word: the first dimension is rounded up
//   to a multiple of 32 bits, also for a 1-D vector;
//   kCoalesce rounds the first dimension up to a multiple of 128 bytes;
//   kTight and kDefault add nothing beyond the bit-word rule.
// Byte sizes per dimension follow the reference's derivation
// (sz_bytes[d] = stride[d] * element_bits / 8 with element_bits = total bits /
// (stride[n-1] * dim[n-1])), so a sub-byte stride reports 0 bytes.
// No cuPHY: this is plain descriptor arithmetic.
#pragma once

#include "mega/matrix.h"

namespace mega {

class PaddedTensor : public Matrix {
 public:
  enum cuphy_dtype_t : uint8_t { kBit, kHalf, kFloat };
  enum cuphy_align_t : uint8_t { kDefault, kTight, kCoalesce };

 private:
  cuphy_dtype_t dtype_ = kHalf;
  cuphy_align_t align_ = kDefault;

  static uint32_t elem_bits(cuphy_dtype_t t) { return t == kBit ? 1 : t == kHalf ? 16 : 32; }

  static uint64_t padded_first(cuphy_dtype_t t, cuphy_align_t a, uint32_t d0) {
    const uint32_t bits = elem_bits(t);
    uint64_t n = d0;
    if (t == kBit) n = (n + 31) / 32 * 32;
    if (a == kCoalesce) {
      const uint64_t per_line = 1024 / bits;  // elements per 128 bytes
      n = (n + per_line - 1) / per_line * per_line;
    }
    return n;
  }

  void init(uint8_t n, const uint32_t* d, const uint32_t* s) {
    check_dims(n);
    if (n == 0) throw std::invalid_argument("Number of dimensions exceeds maximum or below minimum");
    nDims = n;
    for (uint8_t i = 0; i < kDimMax; ++i) dims[i] = i < n ? d[i] : 0;
    if (s) {
      for (uint8_t i = 0; i < kDimMax; ++i) strides[i] = i < n ? s[i] : 0;
    } else {
      uint64_t acc = 1;
      for (uint8_t i = 0; i < kDimMax; ++i) {
        if (i >= n) {
          strides[i] = 0;
          continue;
        }
        strides[i] = (uint32_t)acc;
        acc *= (i == 0) ? padded_first(dtype_, align_, d[0]) : d[i];
      }
    }
    const uint64_t total_elems = (n == 1) ? padded_first(dtype_, align_, d[0])
                                          : (uint64_t)strides[n - 1] * dims[n - 1];
    const uint64_t total_bits = total_elems * elem_bits(dtype_);
    for (uint8_t i = 0; i <= kDimMax; ++i) sz_bytes[i] = 0;
    sz_bytes[n] = (total_bits + 7) / 8;
    const uint64_t denom = (uint64_t)strides[n - 1] * dims[n - 1];
    const uint64_t bits_per = denom ? (sz_bytes[n] * 8) / denom : 0;
    for (uint8_t i = 0; i < n; ++i) sz_bytes[i] = (strides[i] * bits_per) / 8;
  }

  PaddedTensor(cuphy_dtype_t t, uint8_t n, const uint32_t* d, const uint32_t* s,
               const uint64_t* sz, alloc_type_t a, cuphy_align_t al,
               const std::shared_ptr<void>& p)
      : Matrix(n, d, s, sz, a, p), dtype_(t), align_(al) {}

 public:
  PaddedTensor() = default;

  PaddedTensor(cuphy_dtype_t t, uint8_t n, const uint32_t* d,
               const uint32_t* s = nullptr, alloc_type_t a = kDesc,
               cuphy_align_t al = kDefault, const std::shared_ptr<void>& p = nullptr)
      : dtype_(t), align_(al) {
    init(n, d, s);
    allocate(a, p, sz_bytes[nDims]);
  }
  PaddedTensor(cuphy_dtype_t t, uint32_t nVec, alloc_type_t a = kDesc,
               cuphy_align_t al = kDefault, const std::shared_ptr<void>& p = nullptr)
      : PaddedTensor(t, 1, &nVec, nullptr, a, al, p) {}
  PaddedTensor(cuphy_dtype_t t, uint32_t nCols, uint32_t nRows, alloc_type_t a = kDesc,
               cuphy_align_t al = kDefault, const std::shared_ptr<void>& p = nullptr)
      : PaddedTensor(t, 2, std::array<uint32_t, 2>{nCols, nRows}.data(), nullptr, a, al, p) {}
  PaddedTensor(cuphy_dtype_t t, uint32_t nCols, uint32_t nRows, uint32_t nDepths,
               alloc_type_t a = kDesc, cuphy_align_t al = kDefault,
               const std::shared_ptr<void>& p = nullptr)
      : PaddedTensor(t, 3, std::array<uint32_t, 3>{nCols, nRows, nDepths}.data(), nullptr,
                     a, al, p) {}

  inline cuphy_dtype_t dType() const { return dtype_; }
  inline cuphy_align_t alignment() const { return align_; }

  inline PaddedTensor operator[](uint32_t idx) const {
    if (nDims == 0 || idx >= dims[nDims - 1]) throw std::out_of_range("Index out of range");
    if (sz_bytes[nDims - 1] == 0)
      throw std::underflow_error("Size of element is less than a byte");
    uint8_t* p = ptr<uint8_t>() + idx * sz_bytes[nDims - 1];
    return PaddedTensor(dtype_, nDims - 1, dims.data(), strides.data(), sz_bytes.data(),
                        kRef, align_, std::shared_ptr<void>(data, p));
  }
  inline PaddedTensor wrap(void* p) const {
    return PaddedTensor(dtype_, nDims, dims.data(), strides.data(), sz_bytes.data(), kRef,
                        align_, std::shared_ptr<void>(p, [](void*) {}));
  }
};

// The reference's name for this descriptor.
using CuphyTensor = PaddedTensor;

}  // namespace mega
