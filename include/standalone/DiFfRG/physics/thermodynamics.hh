// DiFfRG/physics/thermodynamics.hh -- finite-temperature helpers used inside finite-T kernels
// (reference DiFfRG/include/DiFfRG/physics/thermodynamics.hh:16-135).
#pragma once

#include <cmath>
#include <type_traits>

#include <DiFfRG/common/math.hh>

namespace DiFfRG
{
  template <typename T> __forceinline__ __host__ __device__ auto Cosh(const T x) { return cosh(x); }
  template <typename T> __forceinline__ __host__ __device__ auto Sinh(const T x) { return sinh(x); }
  template <typename T> __forceinline__ __host__ __device__ auto Tanh(const T x) { return tanh(x); }
  template <typename T> __forceinline__ __host__ __device__ auto Coth(const T x) { return 1. / tanh(x); }
  template <typename T> __forceinline__ __host__ __device__ auto Sech(const T x) { return 1. / cosh(x); }
  template <typename T> __forceinline__ __host__ __device__ auto Csch(const T x) { return 1. / sinh(x); }

  // coth(x/2T), with the T/x -> 0 limit +-1 (thermodynamics.hh:28-33)
  template <typename T1, typename T2>
    requires std::is_arithmetic_v<T2>
  __forceinline__ __host__ __device__ auto CothFiniteT(const T1 x, const T2 T)
  {
    using R = decltype(Coth(x / (2 * T)));
    return is_close(T / x, T2(0)) ? (R)(real(x) < 0 ? -1 : 1) : Coth(x / (2 * T));
  }
  template <typename T1, typename T2>
    requires std::is_arithmetic_v<T2>
  __forceinline__ __host__ __device__ auto TanhFiniteT(const T1 x, const T2 T)
  {
    using R = decltype(Tanh(x / (2 * T)));
    return is_close(T / x, T2(0)) ? (R)(real(x) < 0 ? -1 : 1) : Tanh(x / (2 * T));
  }
  template <typename T1, typename T2>
    requires std::is_arithmetic_v<T2>
  __forceinline__ __host__ __device__ auto SechFiniteT(const T1 x, const T2 T)
  {
    using R = decltype(Cosh(x / (2 * T)));
    const auto res = Cosh(x / (2 * T));
    return !std::isfinite(res) ? (R)0 : (R)(1) / res;
  }
  template <typename T1, typename T2>
    requires std::is_arithmetic_v<T2>
  __forceinline__ __host__ __device__ auto CschFiniteT(const T1 x, const T2 T)
  {
    using R = decltype(Sinh(x / (2 * T)));
    const auto res = Sinh(x / (2 * T));
    return !std::isfinite(res) ? (R)0 : (R)(1) / res;
  }

  template <typename T1, typename T2> __forceinline__ __host__ __device__ auto cothS(const T1 e, const T2 T)
  {
    return 1. / tanh(e / (T * 2.));
  }
  template <typename T1, typename T2> __forceinline__ __host__ __device__ auto dcothS(const T1 e, const T2 T)
  {
    return -1. / powr<2>(sinh(e / (T * 2.))) / (2. * T);
  }
  // second and third e-derivatives of coth(u), u = e/2T: in u they are 2 csch^2(u) coth(u) and
  // -2 csch^2(u) (2 coth^2(u) + csch^2(u)); each e-derivative adds a factor 1/2T
  template <typename T1, typename T2> __forceinline__ __host__ __device__ auto ddcothS(const T1 e, const T2 T)
  {
    const auto c = 1. / tanh(e / (T * 2.)), s = 1. / sinh(e / (T * 2.));
    return s * s * c / (2. * T * T);
  }
  template <typename T1, typename T2> __forceinline__ __host__ __device__ auto dddcothS(const T1 e, const T2 T)
  {
    const auto c = 1. / tanh(e / (T * 2.)), s = 1. / sinh(e / (T * 2.));
    return -s * s * (2. * c * c + s * s) / (4. * T * T * T);
  }
  template <typename T1, typename T2> __forceinline__ __host__ __device__ auto tanhS(const T1 e, const T2 T)
  {
    return tanh(e / (T * 2.));
  }
  template <typename T1, typename T2> __forceinline__ __host__ __device__ auto dtanhS(const T1 e, const T2 T)
  {
    return 1. / powr<2>(cosh(e / (T * 2.))) / (2. * T);
  }
  // second e-derivative of tanh(u): -2 sech^2(u) tanh(u) in u, times (1/2T)^2
  template <typename T1, typename T2> __forceinline__ __host__ __device__ auto ddtanhS(const T1 e, const T2 T)
  {
    const auto h = 1. / cosh(e / (T * 2.));
    return -h * h * tanh(e / (T * 2.)) / (2. * T * T);
  }
  template <typename T1, typename T2> __forceinline__ __host__ __device__ auto sechS(const T1 e, const T2 T)
  {
    return 1. / cosh(e / (T * 2.));
  }
  template <typename T1, typename T2> __forceinline__ __host__ __device__ auto cschS(const T1 e, const T2 T)
  {
    return 1. / sinh(e / (T * 2.));
  }

  // Bose-Einstein / Fermi-Dirac distributions (thermodynamics.hh:106-135)
  template <typename T1, typename T2> __forceinline__ __host__ __device__ auto nB(const T1 e, const T2 T)
  {
    return 1. / expm1(e / T);
  }
  template <typename T1, typename T2> __forceinline__ __host__ __device__ auto dnB(const T1 e, const T2 T)
  {
    return -exp(e / T) / powr<2>(expm1(e / T)) / T;
  }
  // second e-derivative of nB: with y = e/T, d^2/dy^2 (e^y - 1)^-1 = e^y (e^y + 1)/(e^y - 1)^3
  template <typename T1, typename T2> __forceinline__ __host__ __device__ auto ddnB(const T1 e, const T2 T)
  {
    const auto ey = exp(e / T), em1 = expm1(e / T);
    return ey * (ey + 1.) / (em1 * em1 * em1) / (T * T);
  }
  template <typename T1, typename T2> __forceinline__ __host__ __device__ auto nF(const T1 e, const T2 T)
  {
    return 1. / (exp(e / T) + 1.);
  }
  template <typename T1, typename T2> __forceinline__ __host__ __device__ auto dnF(const T1 e, const T2 T)
  {
    return -exp(e / T) / powr<2>(exp(e / T) + 1.) / T;
  }
} // namespace DiFfRG
