// DiFfRG/physics/integration/integrator_constant.hh -- IntegratorConstant<d, NT, KERNEL> with the
// reference interface (integrator_constant.hh:12-55): no node sum, the KERNEL's value times the
// solid angle,
//   get(k, t...) = KERNEL::constant(k, t...) + KERNEL::kernel(k, t...) S_d / (2 pi)^d,
// evaluated on the host (as in the reference); the quadrature arguments are accepted and unused.
#pragma once

#include <array>
#include <cmath>
#include <future>

#include <DiFfRG/common/quadrature/quadrature_provider.hh>

#include "dfrg/math.hh"

namespace DiFfRG
{
  template <int d, typename NT, typename KERNEL> class IntegratorConstant
  {
  public:
    using ctype = typename dfrg::ctype_t<NT>;

    IntegratorConstant(QuadratureProvider &quadrature_provider, const std::array<unsigned, 1> grid_size,
                       const ctype x_extent, const unsigned max_block_size = 0)
        : IntegratorConstant(quadrature_provider, grid_size[0], x_extent, max_block_size)
    {
    }
    template <dfrg::JsonLike JSON>
    IntegratorConstant(QuadratureProvider &quadrature_provider, const std::array<unsigned, 1> grid_size,
                       const ctype x_extent, const JSON &)
        : IntegratorConstant(quadrature_provider, grid_size[0], x_extent)
    {
    }
    IntegratorConstant(QuadratureProvider &, const unsigned, const ctype, const unsigned = 0) {}
    IntegratorConstant() {}

    template <typename... T> NT get(const ctype k, const T &...t) const
    {
      return KERNEL::constant(k, t...) +
             KERNEL::kernel(k, t...) * dfrg::S_d_prec<ctype>(d) / dfrg::powr<d>(2 * (ctype)M_PI);
    }
    template <typename... T> std::future<NT> request(const ctype k, const T &...t) const
    {
      return std::async(std::launch::deferred, [=, this]() { return get(k, t...); });
    }
  };
} // namespace DiFfRG
