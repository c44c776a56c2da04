// DiFfRG/physics/integration/integrator_3D_cpu.hh -- Integrator3DTBB<NT, KERNEL>, the host-side
// (q, cos, phi) integrator with the reference interface (integrator_3D_cpu.hh:14-93):
//   get(k, t...) = KERNEL::constant(k, t...)
//                  + (2pi)^-3 int dq q^2 int dcos int_0^{2pi} dphi KERNEL::kernel(q, cos, phi, k, t...)
#pragma once

#include <array>

#include "dfrg/frequency_host.hh"

namespace DiFfRG
{
  template <typename NT, typename KERNEL>
  class Integrator3DTBB : public dfrg::tmpl::FrequencyHost<3, false, false, NT, KERNEL>
  {
    using Base = dfrg::tmpl::FrequencyHost<3, false, false, NT, KERNEL>;

  public:
    using ctype = typename Base::ctype;

    template <dfrg::JsonLike JSON>
    Integrator3DTBB(QuadratureProvider &quadrature_provider, const std::array<unsigned, 3> grid_sizes,
                    const ctype x_extent, const JSON &)
        : Integrator3DTBB(quadrature_provider, grid_sizes, x_extent)
    {
    }
    Integrator3DTBB(QuadratureProvider &quadrature_provider, std::array<unsigned, 3> grid_sizes, const ctype x_extent,
                    const unsigned max_block_size = 0)
        : Base(quadrature_provider, grid_sizes[0], grid_sizes[1], 0, x_extent, ctype(0))
    {
      (void)max_block_size;
      if (grid_sizes[2] != grid_sizes[1])
        throw std::runtime_error("Grid sizes must be currently equal for all angular dimensions!");
    }
  };
} // namespace DiFfRG
