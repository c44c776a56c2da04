// dfrg/dual.hh -- forward-mode dual numbers (value, first derivative) as number types of the
// integrators and interpolators, the role autodiff::real (autodiff::Real<1, double>) plays in the
// reference (types.hh:130-167).
//
// A type opts in by specialising dfrg::DualTraits with is_dual = true and host/device value(x),
// derivative(x) and make(value, derivative).  The FP32 table interpolators then keep a second table
// for the derivative (reference tex_linear_interpolation_1D.hh:58-72, 86-106), and the GPU
// integrators reduce both channels.  autodiff::real opts in below where the autodiff library is
// installed; the accessors are the ones the reference calls (val, derivative).
#pragma once

#include <hip/hip_runtime.h>

#include <array>

namespace dfrg
{
  template <typename NT> struct DualTraits {
    static constexpr bool is_dual = false;
  };
} // namespace dfrg

#if __has_include(<autodiff/forward/real.hpp>)
#include <autodiff/forward/real.hpp>
namespace dfrg
{
  // autodiff's val() and derivative() by argument-dependent lookup, as the reference calls them
  __host__ __device__ inline double autodiff_value(const autodiff::real &x) { return val(x); }
  __host__ __device__ inline double autodiff_derivative(const autodiff::real &x) { return derivative(x); }

  template <> struct DualTraits<autodiff::real> {
    static constexpr bool is_dual = true;
    __host__ __device__ static double value(const autodiff::real &x) { return autodiff_value(x); }
    __host__ __device__ static double derivative(const autodiff::real &x) { return autodiff_derivative(x); }
    __host__ __device__ static autodiff::real make(double v, double d) { return std::array<double, 2>{{v, d}}; }
  };
} // namespace dfrg
#endif
