// DiFfRG/common/types.hh -- get_type::ctype<NT>: the real type of a number type
// (reference DiFfRG/include/DiFfRG/common/types.hh:130-167).
#pragma once

#include <type_traits>
#include <utility>

#include <DiFfRG/common/complex_math.hh>
#include "dfrg/dual.hh"

namespace DiFfRG
{
  // anything indexable, e.g. a grid of external momenta (types.hh:16 in the reference)
  template <typename T>
  concept IsContainer = requires(T x) { x[0]; };

  namespace get_type
  {
    namespace internal
    {
      // a dual number (dfrg/dual.hh): the type of its value, as for
      // autodiff::Real<1, T> in the reference (types.hh:89-96)
      template <typename CT> struct _ctype {
        static_assert(dfrg::DualTraits<CT>::is_dual, "ctype: unsupported number type");
        using value = std::remove_cv_t<
            std::remove_reference_t<decltype(dfrg::DualTraits<CT>::value(std::declval<CT>()))>>;
      };
      template <> struct _ctype<float> { using value = float; };
      template <> struct _ctype<double> { using value = double; };
      template <> struct _ctype<complex<float>> { using value = float; };
      template <> struct _ctype<complex<double>> { using value = double; };
    } // namespace internal
    template <typename CT> using ctype = typename internal::_ctype<CT>::value;
  } // namespace get_type
} // namespace DiFfRG
