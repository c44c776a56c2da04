// DiFfRG/physics/integration/integrator_2D_cartesian_cpu.hh -- Integrator2DCartesianTBB<NT, KERNEL>,
// the host-side 2D Cartesian integrator with the reference interface
// (integrator_2D_cartesian_cpu.hh:20-159):
//   get(k, t...) = KERNEL::constant(k, t...) + (2 pi)^-2 int dqx dqy KERNEL::kernel(qx, qy, k, t...)
#pragma once

#include <array>
#include <cmath>

#include "dfrg/cartesian_host.hh"

namespace DiFfRG
{
  template <typename NT, typename KERNEL> class Integrator2DCartesianTBB : public dfrg::tmpl::CartesianHost<2, NT, KERNEL>
  {
    using Base = dfrg::tmpl::CartesianHost<2, NT, KERNEL>;

  public:
    using ctype = typename Base::ctype;

    template <dfrg::JsonLike JSON>
    Integrator2DCartesianTBB(QuadratureProvider &quadrature_provider, const std::array<unsigned, 2> grid_sizes,
                             const ctype x_extent, const JSON &json)
        : Integrator2DCartesianTBB(quadrature_provider, grid_sizes, x_extent, 0,
                                   json.get_double("/discretization/integration/qx_min", -M_PI),
                                   json.get_double("/discretization/integration/qy_min", -M_PI),
                                   json.get_double("/discretization/integration/qx_max", M_PI),
                                   json.get_double("/discretization/integration/qy_max", M_PI))
    {
    }
    Integrator2DCartesianTBB(QuadratureProvider &quadrature_provider, const std::array<unsigned, 2> grid_sizes,
                             const ctype x_extent = 0., const unsigned max_block_size = 0,
                             const ctype qx_min = -M_PI, const ctype qy_min = -M_PI, const ctype qx_max = M_PI,
                             const ctype qy_max = M_PI)
        : Base(quadrature_provider, grid_sizes, {{qx_min, qy_min}}, {{qx_max, qy_max}})
    {
      (void)x_extent;
      (void)max_block_size;
    }

    void set_qx_min(const ctype qx_min) { this->set_min(0, qx_min); }
    void set_qy_min(const ctype qy_min) { this->set_min(1, qy_min); }
    void set_qx_max(const ctype qx_max) { this->set_max(0, qx_max); }
    void set_qy_max(const ctype qy_max) { this->set_max(1, qy_max); }
  };
} // namespace DiFfRG
