// DiFfRG/physics/integration/integrator_2Dpq0_cpu.hh -- Integrator2Dpq0TBB<NT, KERNEL>, the host-side
// 2+1 dimensional (q, cos, q0) integrator with the reference interface (integrator_2Dpq0_cpu.hh:22-135); the
// measure is that of Integrator2Dpq0GPU.
#pragma once

#include <array>

#include "dfrg/frequency_host.hh"

namespace DiFfRG
{
  template <typename NT, typename KERNEL>
  class Integrator2Dpq0TBB : public dfrg::tmpl::FrequencyHost<2, true, false, NT, KERNEL>
  {
    using Base = dfrg::tmpl::FrequencyHost<2, true, false, NT, KERNEL>;

  public:
    using ctype = typename Base::ctype;

    template <dfrg::JsonLike JSON>
    Integrator2Dpq0TBB(QuadratureProvider &quadrature_provider, const std::array<unsigned, 3> grid_sizes,
                       const ctype x_extent, const ctype q0_extent, const JSON &)
        : Integrator2Dpq0TBB(quadrature_provider, grid_sizes, x_extent, q0_extent)
    {
    }
    Integrator2Dpq0TBB(QuadratureProvider &quadrature_provider, std::array<unsigned, 3> grid_sizes, const ctype x_extent,
                       const ctype q0_extent, const unsigned max_block_size = 0)
        : Base(quadrature_provider, grid_sizes[0], grid_sizes[1], grid_sizes[2], x_extent, q0_extent)
    {
      (void)max_block_size;
    }
  };
} // namespace DiFfRG
