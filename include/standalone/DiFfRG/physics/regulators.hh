// DiFfRG/physics/regulators.hh -- regulator shape functions with the reference interface
// (DiFfRG/include/DiFfRG/physics/regulators.hh): static __host__ __device__ RB, RBdot, dq2RB,
// RF, RFdot, dq2RF of (k^2, q^2).
#pragma once

#include <cmath>
#include <utility>

#include <DiFfRG/common/math.hh>

namespace DiFfRG
{
  namespace internal
  {
    // x^b for a compile-time exponent given as a double
    template <double b, typename T> __forceinline__ __host__ __device__ T xpow(const T x)
    {
      if constexpr (b == (double)(int)b)
        return powr<(int)b>(x);
      else
        return pow(x, T(b));
    }
  } // namespace internal

  // R = (k^2 - q^2) theta(k^2 - q^2) (regulators.hh:24-51)
  template <class Dummy = void> struct LitimRegulator {
    template <typename NT1, typename NT2> static __forceinline__ __host__ __device__ auto RB(const NT1 k2, const NT2 q2)
    {
      return (k2 - q2) * (k2 > q2);
    }
    template <typename NT1, typename NT2> static __forceinline__ __host__ __device__ auto RBdot(const NT1 k2, const NT2 q2)
    {
      return 2. * k2 * (k2 > q2);
    }
    template <typename NT1, typename NT2> static __forceinline__ __host__ __device__ auto RF(const NT1 k2, const NT2 q2)
    {
      return sqrt(RB(k2, q2) + q2) - sqrt(q2);
    }
    template <typename NT1, typename NT2> static __forceinline__ __host__ __device__ auto RFdot(const NT1 k2, const NT2 q2)
    {
      return 0.5 * RBdot(k2, q2) / sqrt(RB(k2, q2) + q2);
    }
    template <typename NT1, typename NT2> static void dq2RB(const NT1, const NT2) = delete;
  };

  struct BosonicRegulatorOpts {
    static constexpr double b = 2.;
  };
  // R = q^2 x^(b-1)/(exp(x^b) - 1), x = q^2/k^2 (regulators.hh:72-123)
  template <class OPTS = BosonicRegulatorOpts> struct BosonicRegulator {
    static constexpr double b = OPTS::b;
    template <typename NT1, typename NT2> static __forceinline__ __host__ __device__ auto RB(const NT1 k2, const NT2 q2)
    {
      return q2 * internal::xpow<b - 1>(q2 / k2) / expm1(internal::xpow<b>(q2 / k2));
    }
    template <typename NT1, typename NT2> static __forceinline__ __host__ __device__ auto dq2RB(const NT1 k2, const NT2 q2)
    {
      const auto xb = internal::xpow<b>(q2 / k2);
      const auto mexp = exp(xb);
      const auto mexpm1 = expm1(xb);
      return b * mexp * internal::xpow<b - 1>(q2 / k2) * (xb - mexpm1) / powr<2>(mexpm1);
    }
    template <typename NT1, typename NT2> static __forceinline__ __host__ __device__ auto RBdot(const NT1 k2, const NT2 q2)
    {
      const auto xb = internal::xpow<b>(q2 / k2);
      const auto mexp = exp(xb);
      const auto mexpm1 = expm1(xb);
      return 2. * k2 * xb * (mexpm1 * (1. - b) + b * mexp * xb) / powr<2>(mexpm1);
    }
    template <typename NT1, typename NT2> static __forceinline__ __host__ __device__ auto RF(const NT1 k2, const NT2 q2)
    {
      return sqrt(RB(k2, q2) + q2) - sqrt(q2);
    }
    template <typename NT1, typename NT2> static __forceinline__ __host__ __device__ auto dq2RF(const NT1 k2, const NT2 q2)
    {
      const auto q = sqrt(q2);
      return (-1. + q * (1. + dq2RB(k2, q2)) / (q + RF(k2, q2))) / (2. * q);
    }
    template <typename NT1, typename NT2> static __forceinline__ __host__ __device__ auto RFdot(const NT1 k2, const NT2 q2)
    {
      return 0.5 * RBdot(k2, q2) / sqrt(RB(k2, q2) + q2);
    }
  };

  struct ExponentialRegulatorOpts {
    static constexpr double b = 2.;
  };
  // R = k^2 exp(-x^b), x = q^2/k^2 (regulators.hh:144-191)
  template <class OPTS = ExponentialRegulatorOpts> struct ExponentialRegulator {
    static constexpr double b = OPTS::b;
    template <typename NT1, typename NT2> static __forceinline__ __host__ __device__ auto RB(const NT1 k2, const NT2 q2)
    {
      const auto xb = internal::xpow<b>(q2 / k2);
      return k2 * exp(-xb);
    }
    template <typename NT1, typename NT2> static __forceinline__ __host__ __device__ auto dq2RB(const NT1 k2, const NT2 q2)
    {
      const auto xbm1 = internal::xpow<b - 1>(q2 / k2);
      const auto xb = xbm1 * (q2 / k2);
      return -b * xbm1 * exp(-xb);
    }
    template <typename NT1, typename NT2> static __forceinline__ __host__ __device__ auto RBdot(const NT1 k2, const NT2 q2)
    {
      const auto xb = internal::xpow<b>(q2 / k2);
      return 2. * exp(-xb) * k2 * (1. + b * xb);
    }
    template <typename NT1, typename NT2> static __forceinline__ __host__ __device__ auto RF(const NT1 k2, const NT2 q2)
    {
      return sqrt(RB(k2, q2) + q2) - sqrt(q2);
    }
    template <typename NT1, typename NT2> static __forceinline__ __host__ __device__ auto dq2RF(const NT1 k2, const NT2 q2)
    {
      const auto q = sqrt(q2);
      return (-1. + q * (1. + dq2RB(k2, q2)) / (q + RF(k2, q2))) / (2. * q);
    }
    template <typename NT1, typename NT2> static __forceinline__ __host__ __device__ auto RFdot(const NT1 k2, const NT2 q2)
    {
      return 0.5 * RBdot(k2, q2) / sqrt(RB(k2, q2) + q2);
    }
  };

  struct SmoothedLitimRegulatorOpts {
    static constexpr double alpha = 2e-3;
  };
  // R = (k^2 - q^2)/(1 + exp((x - 1)/alpha)) (regulators.hh:212-242)
  template <class OPTS = SmoothedLitimRegulatorOpts> struct SmoothedLitimRegulator {
    static constexpr double alpha = OPTS::alpha;
    template <typename NT1, typename NT2> static __forceinline__ __host__ __device__ auto RB(const NT1 k2, const NT2 q2)
    {
      return (k2 - q2) / (1. + exp((q2 / k2 - 1.) / alpha));
    }
    template <typename NT1, typename NT2> static __forceinline__ __host__ __device__ auto RBdot(const NT1 k2, const NT2 q2)
    {
      const auto x = q2 / k2;
      const auto mexp = exp((x - 1.) / alpha);
      return mexp > 1e50 ? 0. : 2. * k2 * (1. + mexp * (x - powr<2>(x) + alpha) / alpha) / powr<2>(1. + mexp);
    }
    template <typename NT1, typename NT2> static __forceinline__ __host__ __device__ auto RF(const NT1 k2, const NT2 q2)
    {
      return sqrt(RB(k2, q2) + q2) - sqrt(q2);
    }
    template <typename NT1, typename NT2> static __forceinline__ __host__ __device__ auto RFdot(const NT1 k2, const NT2 q2)
    {
      return 0.5 * RBdot(k2, q2) / sqrt(RB(k2, q2) + q2);
    }
  };

  struct PolynomialExpRegulatorOpts {
    static constexpr int order = 8;
  };
  // R = k^2 exp(-f(x)), f(x) = sum_{c=1..order} x^c / c (regulators.hh:625-695)
  template <class OPTS = PolynomialExpRegulatorOpts> struct PolynomialExpRegulator {
    static constexpr int order = OPTS::order;
    static_assert(order > 0, "PolynomialExpRegulator: order must be > 0 !");
    template <typename T> static __forceinline__ __host__ __device__ T get_f(const T x)
    {
      T f = T(0), xp = T(1);
      for (int c = 1; c <= order; ++c) {
        xp = xp * x;
        f = f + xp / T(c);
      }
      return f;
    }
    template <typename T> static __forceinline__ __host__ __device__ T get_df(const T x)
    {
      T df = T(0), xp = T(1);
      for (int c = 0; c < order; ++c) {
        df = df + xp;
        xp = xp * x;
      }
      return df;
    }
    template <typename NT1, typename NT2> static __forceinline__ __host__ __device__ auto RB(const NT1 k2, const NT2 q2)
    {
      const auto x = q2 / k2;
      return k2 * exp(-get_f(x));
    }
    template <typename NT1, typename NT2> static __forceinline__ __host__ __device__ auto RBdot(const NT1 k2, const NT2 q2)
    {
      const auto x = q2 / k2;
      return 2 * k2 * exp(-get_f(x)) * (1 + get_df(x) * x);
    }
    template <typename NT1, typename NT2> static __forceinline__ __host__ __device__ auto dq2RB(const NT1 k2, const NT2 q2)
    {
      const auto x = q2 / k2;
      return -exp(-get_f(x)) * get_df(x);
    }
    template <typename NT1, typename NT2> static __forceinline__ __host__ __device__ auto RF(const NT1 k2, const NT2 q2)
    {
      return sqrt(RB(k2, q2) + q2) - sqrt(q2);
    }
    template <typename NT1, typename NT2> static __forceinline__ __host__ __device__ auto RFdot(const NT1 k2, const NT2 q2)
    {
      using T = decltype(k2 * q2);
      return T(0.5) * RBdot(k2, q2) / sqrt(RB(k2, q2) + q2);
    }
    template <typename NT1, typename NT2> static __forceinline__ __host__ __device__ auto dq2RF(const NT1 k2, const NT2 q2)
    {
      const auto q = sqrt(q2);
      return (-1 + q * (1 + dq2RB(k2, q2)) / (q + RF(k2, q2))) / (2 * q);
    }
  };
} // namespace DiFfRG
