// DiFfRG/physics/integration/integrator_2Dpx0_cpu.hh -- Integrator2Dpx0TBB<NT, KERNEL>, the host-side
// 2+1 dimensional (q, cos, q0) integrator with the reference interface (integrator_2Dpx0_cpu.hh:24-137); the
// measure is that of Integrator2Dpx0GPU.
#pragma once

#include <array>

#include "dfrg/frequency_host.hh"

namespace DiFfRG
{
  template <typename NT, typename KERNEL>
  class Integrator2Dpx0TBB : public dfrg::tmpl::FrequencyHost<2, true, true, NT, KERNEL>
  {
    using Base = dfrg::tmpl::FrequencyHost<2, true, true, NT, KERNEL>;

  public:
    using ctype = typename Base::ctype;

    template <dfrg::JsonLike JSON>
    Integrator2Dpx0TBB(QuadratureProvider &quadrature_provider, const std::array<unsigned, 3> grid_sizes,
                       const ctype x_extent, const ctype x0_extent, const JSON &)
        : Integrator2Dpx0TBB(quadrature_provider, grid_sizes, x_extent, x0_extent)
    {
    }
    Integrator2Dpx0TBB(QuadratureProvider &quadrature_provider, std::array<unsigned, 3> grid_sizes, const ctype x_extent,
                       const ctype x0_extent, const unsigned max_block_size = 0)
        : Base(quadrature_provider, grid_sizes[0], grid_sizes[1], grid_sizes[2], x_extent, x0_extent)
    {
      (void)max_block_size;
    }
  };
} // namespace DiFfRG
