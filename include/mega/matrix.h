// mega::Matrix — up to 3-D strided buffer descriptor with shared ownership,
// the argument type of the Beamform / Equalize drop-in classes.
//
// Semantics follow the reference descriptor (src/operators/matrix/matrix.h:
// 30-426): dims[0] is the fastest dimension, strides in elements, szBytes(d)
// = element size * stride(d) and szBytes(nDim()) = the whole array;
// operator[] drops the last dimension, sub(start, end) slices it, wrap(ptr)
// reuses the descriptor for foreign memory. Storage is allocated on the host
// (malloc), in pinned host memory (hipHostMalloc) or on the current HIP device
// (hipMalloc), and freed when the last Matrix sharing it goes away.
#pragma once

#include <hip/hip_runtime_api.h>

#include <array>
#include <cstdint>
#include <cstdlib>
#include <memory>
#include <new>
#include <stdexcept>

namespace mega {

class Matrix {
 public:
  enum alloc_type_t : uint8_t {
    kNull,   // empty matrix (default constructor)
    kDesc,   // descriptor only, no storage
    kRef,    // refers to storage owned elsewhere
    kHost,   // malloc
    kHPin,   // pinned host memory
    kDevice  // device memory of the current HIP device
  };

 protected:
  static constexpr uint8_t kDimMax = 3;

  uint8_t nDims = 0;
  std::array<uint32_t, kDimMax> dims{};
  std::array<uint32_t, kDimMax> strides{};
  std::array<uint64_t, kDimMax + 1> sz_bytes{};
  alloc_type_t alloc = kNull;
  std::shared_ptr<void> data;

  static void check_dims(uint8_t n) {
    if (n > kDimMax)
      throw std::invalid_argument("Number of dimensions exceeds maximum or below minimum");
  }

  void set_shape(uint8_t sz_elem, uint8_t n, const uint32_t* dims_,
                 const uint32_t* strides_) {
    check_dims(n);
    nDims = n;
    for (uint8_t i = 0; i < kDimMax; ++i) dims[i] = i < n ? dims_[i] : 0;
    uint32_t s = 1;
    for (uint8_t i = 0; i < kDimMax; ++i) {
      if (i < n) {
        strides[i] = strides_ ? strides_[i] : s;
        s *= dims_[i];
      } else {
        strides[i] = 0;
      }
    }
    for (uint8_t i = 0; i <= kDimMax; ++i) sz_bytes[i] = 0;
    for (uint8_t i = 0; i < n; ++i) sz_bytes[i] = (uint64_t)sz_elem * strides[i];
    if (n > 0) sz_bytes[n] = (uint64_t)sz_elem * strides[n - 1] * dims[n - 1];
  }

  void allocate(alloc_type_t a, const std::shared_ptr<void>& data_, uint64_t size) {
    alloc = a;
    if (data_) {
      data = data_;
      return;
    }
    void* p = nullptr;
    switch (a) {
      case kHost:
        p = std::malloc(size ? size : 1);
        if (!p) throw std::bad_alloc();
        data = std::shared_ptr<void>(p, std::free);
        break;
      case kHPin:
        if (hipHostMalloc(&p, size ? size : 1, hipHostMallocDefault) != hipSuccess || !p)
          throw std::bad_alloc();
        data = std::shared_ptr<void>(p, [](void* q) { (void)hipHostFree(q); });
        break;
      case kDevice:
        if (hipMalloc(&p, size ? size : 1) != hipSuccess || !p) throw std::bad_alloc();
        data = std::shared_ptr<void>(p, [](void* q) { (void)hipFree(q); });
        break;
      default:
        data = nullptr;
    }
  }

  // Descriptor copy with explicit byte sizes (views).
  Matrix(uint8_t n, const uint32_t* dims_, const uint32_t* strides_,
         const uint64_t* sz_bytes_, alloc_type_t a, const std::shared_ptr<void>& data_)
      : nDims(n) {
    check_dims(n);
    for (uint8_t i = 0; i < kDimMax; ++i) {
      dims[i] = i < n ? dims_[i] : 0;
      strides[i] = i < n ? strides_[i] : 0;
    }
    for (uint8_t i = 0; i <= kDimMax; ++i) sz_bytes[i] = i <= n ? sz_bytes_[i] : 0;
    alloc = a;
    data = data_;
  }

 public:
  Matrix() = default;
  ~Matrix() = default;

  Matrix(uint8_t sz_elem, uint8_t n, const uint32_t* dims_,
         const uint32_t* strides_ = nullptr, alloc_type_t a = kDesc,
         const std::shared_ptr<void>& data_ = nullptr) {
    set_shape(sz_elem, n, dims_, strides_);
    allocate(a, data_, sz_bytes[nDims]);
  }
  Matrix(uint8_t sz_elem, uint32_t nVec, alloc_type_t a = kDesc,
         const std::shared_ptr<void>& data_ = nullptr)
      : Matrix(sz_elem, 1, &nVec, nullptr, a, data_) {}
  Matrix(uint8_t sz_elem, uint32_t nCols, uint32_t nRows, alloc_type_t a = kDesc,
         const std::shared_ptr<void>& data_ = nullptr)
      : Matrix(sz_elem, 2, std::array<uint32_t, 2>{nCols, nRows}.data(), nullptr, a, data_) {}
  Matrix(uint8_t sz_elem, uint32_t nCols, uint32_t nRows, uint32_t nDepths,
         alloc_type_t a = kDesc, const std::shared_ptr<void>& data_ = nullptr)
      : Matrix(sz_elem, 3, std::array<uint32_t, 3>{nCols, nRows, nDepths}.data(), nullptr,
               a, data_) {}

  template <typename T = void>
  inline T* ptr() const { return reinterpret_cast<T*>(data.get()); }
  template <typename T>
  inline T& deref(uint64_t idx) const { return ptr<T>()[idx]; }

  inline uint8_t nDim() const { return nDims; }
  inline uint32_t dim(uint8_t i) const {
    if (i >= kDimMax) throw std::out_of_range("Index out of range");
    return dims[i];
  }
  inline uint32_t stride(uint8_t i) const {
    if (i >= kDimMax) throw std::out_of_range("Index out of range");
    return strides[i];
  }
  inline uint64_t szBytes(uint8_t i) const {
    if (i > kDimMax) throw std::out_of_range("Index out of range");
    return sz_bytes[i];
  }
  inline uint64_t szBytes() const { return sz_bytes[nDims]; }
  inline alloc_type_t allocType() const { return alloc; }
  inline std::shared_ptr<void> raw() const { return data; }

  inline Matrix operator[](uint32_t idx) const {
    if (nDims == 0 || idx >= dims[nDims - 1]) throw std::out_of_range("Index out of range");
    uint8_t* p = ptr<uint8_t>() + idx * sz_bytes[nDims - 1];
    return Matrix(nDims - 1, dims.data(), strides.data(), sz_bytes.data(), kRef,
                  std::shared_ptr<void>(data, p));
  }
  inline Matrix wrap(void* p) const {
    return Matrix(nDims, dims.data(), strides.data(), sz_bytes.data(), kRef,
                  std::shared_ptr<void>(p, [](void*) {}));
  }
  inline Matrix sub(uint32_t start, uint32_t end) const {
    if (nDims == 0 || start >= end || end > dims[nDims - 1])
      throw std::out_of_range("Index out of range");
    uint8_t* p = ptr<uint8_t>() + start * sz_bytes[nDims - 1];
    auto d = dims;
    d[nDims - 1] = end - start;
    return Matrix((uint8_t)sz_bytes[0], nDims, d.data(), nullptr, kRef,
                  std::shared_ptr<void>(data, p));
  }
};

}  // namespace mega
