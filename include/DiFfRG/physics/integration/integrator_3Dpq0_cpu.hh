// DiFfRG/physics/integration/integrator_3Dpq0_cpu.hh -- Integrator3Dpq0TBB<NT, KERNEL>, the host-side
// 3+1 dimensional (q, cos, phi, q0) integrator with the reference interface (integrator_3Dpq0_cpu.hh:14-104);
// the measure and node sets are those of Integrator3Dpq0GPU (the q0 axis of order grid_sizes[3]).
#pragma once

#include <array>

#include "dfrg/frequency_host.hh"

namespace DiFfRG
{
  template <typename NT, typename KERNEL>
  class Integrator3Dpq0TBB : public dfrg::tmpl::FrequencyHost<3, true, false, NT, KERNEL>
  {
    using Base = dfrg::tmpl::FrequencyHost<3, true, false, NT, KERNEL>;

  public:
    using ctype = typename Base::ctype;

    template <dfrg::JsonLike JSON>
    Integrator3Dpq0TBB(QuadratureProvider &quadrature_provider, const std::array<unsigned, 4> grid_sizes,
                       const ctype x_extent, const ctype q0_extent, const JSON &)
        : Integrator3Dpq0TBB(quadrature_provider, grid_sizes, x_extent, q0_extent)
    {
    }
    Integrator3Dpq0TBB(QuadratureProvider &quadrature_provider, std::array<unsigned, 4> grid_sizes, const ctype x_extent,
                       const ctype q0_extent, const unsigned max_block_size = 0)
        : Base(quadrature_provider, grid_sizes[0], grid_sizes[1], grid_sizes[3], x_extent, q0_extent)
    {
      (void)max_block_size;
      if (grid_sizes[2] != grid_sizes[1])
        throw std::runtime_error("Grid sizes must be currently equal for all angular dimensions!");
    }
  };
} // namespace DiFfRG
