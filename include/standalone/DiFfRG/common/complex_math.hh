// DiFfRG/common/complex_math.hh -- a small host/device complex number for complex-valued
// integrands (NT = complex<float|double> in the integrator templates).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

namespace DiFfRG
{
  template <typename T> struct complex {
    T re, im;
    constexpr __host__ __device__ complex(T r = T(0), T i = T(0)) : re(r), im(i) {}
    template <typename U> constexpr __host__ __device__ complex(const complex<U> &o) : re(T(o.re)), im(T(o.im)) {}
    constexpr __host__ __device__ T real() const { return re; }
    constexpr __host__ __device__ T imag() const { return im; }
    __host__ __device__ complex &operator+=(const complex &o) { re += o.re; im += o.im; return *this; }
    __host__ __device__ complex &operator-=(const complex &o) { re -= o.re; im -= o.im; return *this; }
    __host__ __device__ complex &operator*=(const complex &o) { *this = *this * o; return *this; }
    friend constexpr __host__ __device__ complex operator+(complex a, complex b) { return {a.re + b.re, a.im + b.im}; }
    friend constexpr __host__ __device__ complex operator-(complex a, complex b) { return {a.re - b.re, a.im - b.im}; }
    friend constexpr __host__ __device__ complex operator-(complex a) { return {-a.re, -a.im}; }
    friend constexpr __host__ __device__ complex operator*(complex a, complex b)
    {
      return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    }
    friend constexpr __host__ __device__ complex operator*(complex a, T s) { return {a.re * s, a.im * s}; }
    friend constexpr __host__ __device__ complex operator*(T s, complex a) { return {a.re * s, a.im * s}; }
    friend constexpr __host__ __device__ complex operator/(complex a, complex b)
    {
      const T d = b.re * b.re + b.im * b.im;
      return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
    }
    friend constexpr __host__ __device__ complex operator/(complex a, T s) { return {a.re / s, a.im / s}; }
    friend constexpr __host__ __device__ bool operator==(complex a, complex b) { return a.re == b.re && a.im == b.im; }
  };

  template <typename T> struct is_complex : std::false_type {};
  template <typename T> struct is_complex<complex<T>> : std::true_type {};

  template <typename T> constexpr __host__ __device__ T real(const complex<T> &z) { return z.re; }
  template <typename T> constexpr __host__ __device__ T imag(const complex<T> &z) { return z.im; }
  template <typename T>
    requires std::is_arithmetic_v<T>
  constexpr __host__ __device__ T real(const T &x)
  {
    return x;
  }
  template <typename T>
    requires std::is_arithmetic_v<T>
  constexpr __host__ __device__ T imag(const T &)
  {
    return T(0);
  }
} // namespace DiFfRG
