// DiFfRG/physics/integration/integrator_1D_cartesian_cpu.hh -- Integrator1DCartesianTBB<NT, KERNEL>,
// the host-side 1D Cartesian integrator with the reference interface
// (integrator_1D_cartesian_cpu.hh:20-160):
//   get(k, t...) = KERNEL::constant(k, t...) + (2 pi)^-1 int_{qx_min}^{qx_max} dqx KERNEL::kernel(qx, k, t...)
#pragma once

#include <array>
#include <cmath>

#include "dfrg/cartesian_host.hh"

namespace DiFfRG
{
  template <typename NT, typename KERNEL> class Integrator1DCartesianTBB : public dfrg::tmpl::CartesianHost<1, NT, KERNEL>
  {
    using Base = dfrg::tmpl::CartesianHost<1, NT, KERNEL>;

  public:
    using ctype = typename Base::ctype;

    Integrator1DCartesianTBB(QuadratureProvider &quadrature_provider, const std::array<unsigned, 1> grid_size,
                             const ctype x_extent = 0., const unsigned max_block_size = 0,
                             const ctype qx_min = -M_PI, const ctype qx_max = M_PI)
        : Integrator1DCartesianTBB(quadrature_provider, grid_size[0], x_extent, qx_min, qx_max)
    {
      (void)max_block_size;
    }
    template <dfrg::JsonLike JSON>
    Integrator1DCartesianTBB(QuadratureProvider &quadrature_provider, const std::array<unsigned, 1> grid_size,
                             const JSON &json)
        : Integrator1DCartesianTBB(quadrature_provider, grid_size[0], 0.,
                                   json.get_double("/discretization/integration/qx_min", -M_PI),
                                   json.get_double("/discretization/integration/qx_max", M_PI))
    {
    }
    template <dfrg::JsonLike JSON>
    Integrator1DCartesianTBB(QuadratureProvider &quadrature_provider, const std::array<unsigned, 1> grid_size,
                             const ctype x_extent, const JSON &json)
        : Integrator1DCartesianTBB(quadrature_provider, grid_size[0], x_extent,
                                   json.get_double("/discretization/integration/qx_min", -M_PI),
                                   json.get_double("/discretization/integration/qx_max", M_PI))
    {
    }
    Integrator1DCartesianTBB(QuadratureProvider &quadrature_provider, const unsigned grid_size,
                             const ctype x_extent = 0., const ctype qx_min = -M_PI, const ctype qx_max = M_PI)
        : Base(quadrature_provider, {{grid_size}}, {{qx_min}}, {{qx_max}})
    {
      (void)x_extent;
    }

    void set_qx_min(const ctype qx_min) { this->set_min(0, qx_min); }
    void set_qx_max(const ctype qx_max) { this->set_max(0, qx_max); }
  };
} // namespace DiFfRG
