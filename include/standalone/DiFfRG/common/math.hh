// DiFfRG/common/math.hh -- numeric helpers used inside integrands and integrators.
// Interface of the reference's DiFfRG/include/DiFfRG/common/math.hh:45-195 (powr, V_d, S_d,
// S_d_prec, heaviside_theta, sign, is_close), written for hipcc host + device code.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <limits>
#include <type_traits>

#include <DiFfRG/common/complex_math.hh>

namespace DiFfRG
{
  using uint = unsigned int;

  // x^n for integer n, evaluated as x * x^(n-1) (math.hh:45-55)
  template <int n, typename NumberType> constexpr __forceinline__ __host__ __device__ NumberType powr(const NumberType x)
  {
    if constexpr (n == 0)
      return NumberType(1.);
    else if constexpr (n < 0)
      return NumberType(1.) / powr<-n, NumberType>(x);
    else if constexpr (n > 1)
      return x * powr<n - 1, NumberType>(x);
    else
      return x;
  }

  // volume and surface of the d-ball / d-sphere (math.hh:63-96)
  template <typename NT> __forceinline__ __host__ __device__ double V_d(NT d)
  {
    return pow(M_PI, d / 2.) / tgamma(d / 2. + 1.);
  }
  template <typename NT1, typename NT2> __forceinline__ __host__ __device__ double V_d(NT1 d, NT2 extent)
  {
    return pow(M_PI, d / 2.) / tgamma(d / 2. + 1.) * pow(extent, d);
  }
  template <typename NT> __forceinline__ __host__ __device__ double S_d(NT d) { return 2. * pow(M_PI, d / 2.) / tgamma(d / 2.); }

  // S_d as a compile-time constant (math.hh:104-121)
  template <typename NT> constexpr NT S_d_prec(uint d)
  {
    if (d == 1) return 2;
    if (d == 2) return static_cast<NT>(2) * static_cast<NT>(M_PI);
    if (d == 3) return static_cast<NT>(4) * static_cast<NT>(M_PI);
    if (d == 4) return static_cast<NT>(2) * powr<2>(static_cast<NT>(M_PI));
    if (d == 5) return static_cast<NT>(8) * powr<2>(static_cast<NT>(M_PI)) / static_cast<NT>(3);
    if (d == 6) return powr<3>(static_cast<NT>(M_PI));
    if (d == 7) return static_cast<NT>(16) * powr<3>(static_cast<NT>(M_PI)) / static_cast<NT>(15);
    return std::numeric_limits<NT>::quiet_NaN();
  }

  template <typename NumberType> constexpr __forceinline__ __host__ __device__ auto heaviside_theta(const NumberType x)
  {
    return x >= static_cast<NumberType>(0) ? static_cast<NumberType>(1) : static_cast<NumberType>(0);
  }
  template <typename NumberType> constexpr __forceinline__ __host__ __device__ auto sign(const NumberType x)
  {
    return x >= static_cast<NumberType>(0) ? static_cast<NumberType>(1) : static_cast<NumberType>(-1);
  }

  // absolute OR relative agreement within eps (math.hh:156-195)
  template <typename T1, typename T2, typename T3>
    requires std::is_floating_point_v<T3>
  __forceinline__ __host__ __device__ bool is_close(T1 a, T2 b, T3 eps_)
  {
    if constexpr (is_complex<T1>::value || is_complex<T2>::value) {
      return is_close(real(a), real(b), eps_) && is_close(imag(a), imag(b), eps_);
    } else {
      const auto diff = fabs(a - b);
      if (diff <= eps_) return true;
      if (diff <= fmax(fabs(a), fabs(b)) * eps_) return true;
      return false;
    }
  }
  template <typename T1, typename T2> __forceinline__ __host__ __device__ bool is_close(T1 a, T2 b)
  {
    if constexpr (is_complex<T1>::value || is_complex<T2>::value) {
      return is_close(real(a), real(b)) && is_close(imag(a), imag(b));
    } else {
      constexpr auto eps_ = std::numeric_limits<T1>::epsilon() > std::numeric_limits<T2>::epsilon()
                                ? std::numeric_limits<T1>::epsilon()
                                : std::numeric_limits<T2>::epsilon();
      return is_close(a, b, eps_);
    }
  }
} // namespace DiFfRG
