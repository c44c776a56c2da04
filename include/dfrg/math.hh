// dfrg/math.hh -- the numeric helpers the drop-in integrators and interpolators use themselves, in
// namespace dfrg so that they coexist with the reference's DiFfRG/common/math.hh and types.hh
// (whose DiFfRG::powr, S_d_prec, is_close and get_type::ctype user KERNELs keep calling):
//   powr<n>(x)     x^n for integer n as repeated products (reference math.hh:45-55)
//   S_d_prec<T>(d) surface of the unit sphere in d dimensions (math.hh:104-121)
//   is_close       absolute or relative agreement (math.hh:156-195), real scalars
//   ctype_t<NT>    the real type of a number type (types.hh:130-167)
//   JsonLike       what the integrator and provider constructors read from a parameter object
//                  (the getters of the reference's JSONValue, common/json.hh:62-139)
//   gld(p)         device load of *p as global memory (for pointers the compiler cannot place:
//                  members of argument tuples and node tables point into HBM)
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <limits>
#include <string>
#include <type_traits>
#include <utility>

#include "dfrg/dual.hh"

namespace dfrg
{
  // *p through a global-address-space pointer: a global load with a scalar base and 32-bit offsets
  // instead of a flat load with a 64-bit address per lane.  p must point to device-visible global
  // memory (HBM or mapped host memory), as every table of the integrators and interpolators does.
  template <class T> __device__ __forceinline__ T gld(const T *p)
  {
    if constexpr (std::is_scalar_v<T>) {
      return *(const __attribute__((address_space(1))) T *)p;
    } else {
      // class types (complex, dual numbers, node records): their words, then the bytes
      static_assert(std::is_trivially_copyable_v<T>, "gld: the type must be trivially copyable");
      using W = std::conditional_t<sizeof(T) % 8 == 0 && alignof(T) >= 8, unsigned long long, unsigned>;
      static_assert(sizeof(T) % sizeof(W) == 0, "gld: the size must be a multiple of 4 bytes");
      constexpr int n = sizeof(T) / sizeof(W);
      const __attribute__((address_space(1))) W *q = (const __attribute__((address_space(1))) W *)p;
      W w[n];
#pragma unroll
      for (int i = 0; i < n; ++i) w[i] = q[i];
      T r;
      __builtin_memcpy(&r, w, sizeof(T));
      return r;
    }
  }

  // *p: a global load on the device (gld), a plain read on the host; for tables that live in HBM
  // on the device and in host memory on the host
  template <class T> __host__ __device__ __forceinline__ T tld(const T *p)
  {
#ifdef __HIP_DEVICE_COMPILE__
    return gld(p);
#else
    return *p;
#endif
  }

  template <int n, typename NT> constexpr __forceinline__ __host__ __device__ NT powr(const NT x)
  {
    if constexpr (n == 0)
      return NT(1.);
    else if constexpr (n < 0)
      return NT(1.) / powr<-n, NT>(x);
    else if constexpr (n > 1)
      return x * powr<n - 1, NT>(x);
    else
      return x;
  }

  template <typename NT> constexpr NT S_d_prec(unsigned d)
  {
    constexpr NT pi = NT(3.14159265358979323846);
    if (d == 1) return NT(2);
    if (d == 2) return NT(2) * pi;
    if (d == 3) return NT(4) * pi;
    if (d == 4) return NT(2) * powr<2>(pi);
    if (d == 5) return NT(8) * powr<2>(pi) / NT(3);
    if (d == 6) return powr<3>(pi);
    if (d == 7) return NT(16) * powr<3>(pi) / NT(15);
    return std::numeric_limits<NT>::quiet_NaN();
  }

  template <typename T1, typename T2, typename T3>
  __forceinline__ __host__ __device__ bool is_close(T1 a, T2 b, T3 eps)
  {
    const auto diff = fabs(a - b);
    return diff <= eps || diff <= fmax(fabs(a), fabs(b)) * eps;
  }
  template <typename T1, typename T2> __forceinline__ __host__ __device__ bool is_close(T1 a, T2 b)
  {
    constexpr auto eps = std::numeric_limits<T1>::epsilon() > std::numeric_limits<T2>::epsilon()
                             ? std::numeric_limits<T1>::epsilon()
                             : std::numeric_limits<T2>::epsilon();
    return is_close(a, b, eps);
  }

  namespace detail
  {
    template <typename NT, typename = void> struct CType {
      static_assert(DualTraits<NT>::is_dual, "ctype: unsupported number type");
      using type = std::remove_cv_t<std::remove_reference_t<decltype(DualTraits<NT>::value(std::declval<NT>()))>>;
    };
    template <> struct CType<float> {
      using type = float;
    };
    template <> struct CType<double> {
      using type = double;
    };
    // complex numbers (anything with real() and imag(), e.g. std::complex or cuda::std::complex)
    template <typename NT>
    struct CType<NT, std::void_t<decltype(std::declval<const NT &>().real()), decltype(std::declval<const NT &>().imag())>> {
      using type = std::remove_cv_t<std::remove_reference_t<decltype(std::declval<const NT &>().real())>>;
    };
  } // namespace detail
  template <typename NT> using ctype_t = typename detail::CType<NT>::type;

  template <typename J>
  concept JsonLike = requires(const J &j, const std::string &key) {
    { j.get_double(key) };
    { j.get_uint(key, 0u) };
  };
} // namespace dfrg
