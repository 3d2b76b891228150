// mega::Complex — interleaved single-precision complex value, the element type
// of every CSI, weight and received-sample buffer of the detection path.
//
// Same layout (float re, float im) and operator set as the reference type
// (src/operators/mega_complex.h:25-83); usable in host and device code.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace mega {

struct Complex {
  float re;
  float im;

  __host__ __device__ Complex(float re_, float im_) : re(re_), im(im_) {}
  __host__ __device__ Complex() : re(0.f), im(0.f) {}
  __host__ __device__ Complex(const float2& v) : re(v.x), im(v.y) {}

  __host__ __device__ inline Complex operator+(const Complex& o) const {
    return Complex(re + o.re, im + o.im);
  }
  __host__ __device__ inline Complex operator-(const Complex& o) const {
    return Complex(re - o.re, im - o.im);
  }
  __host__ __device__ inline Complex operator*(const Complex& o) const {
    return Complex(re * o.re - im * o.im, re * o.im + im * o.re);
  }
  __host__ __device__ inline Complex operator-() const { return Complex(-re, -im); }
  __host__ __device__ inline Complex operator+=(const Complex& o) {
    re += o.re;
    im += o.im;
    return *this;
  }
  __host__ __device__ inline Complex operator-=(const Complex& o) {
    re -= o.re;
    im -= o.im;
    return *this;
  }
  __host__ __device__ inline Complex operator*=(const Complex& o) {
    const float r = re * o.re - im * o.im;
    im = re * o.im + im * o.re;
    re = r;
    return *this;
  }
  __host__ __device__ inline bool operator==(const Complex& o) const {
    return re == o.re && im == o.im;
  }
  __host__ __device__ inline bool operator!=(const Complex& o) const {
    return !(*this == o);
  }
  __host__ __device__ inline Complex conj() const { return Complex(re, -im); }
  __host__ __device__ inline float abs() const { return sqrtf(re * re + im * im); }
  __host__ __device__ inline float pow2() const { return re * re + im * im; }
  __host__ __device__ inline operator float2() const { return make_float2(re, im); }
  __host__ __device__ inline Complex operator/(const float& d) const {
    return Complex(re / d, im / d);
  }
  __host__ __device__ inline Complex operator/=(const float& d) {
    re /= d;
    im /= d;
    return *this;
  }
};

static_assert(sizeof(Complex) == 8, "Complex must be two packed floats");

}  // namespace mega
