// DiFfRG/physics/integration/integrator_3D_cartesian_cpu.hh -- Integrator3DCartesianTBB<NT, KERNEL>,
// the host-side 3D Cartesian integrator with the reference interface
// (integrator_3D_cartesian_cpu.hh:14-156):
//   get(k, t...) = KERNEL::constant(k, t...) + (2 pi)^-3 int dqx dqy dqz KERNEL::kernel(qx, qy, qz, k, t...)
#pragma once

#include <array>
#include <cmath>

#include "dfrg/cartesian_host.hh"

namespace DiFfRG
{
  template <typename NT, typename KERNEL> class Integrator3DCartesianTBB : public dfrg::tmpl::CartesianHost<3, NT, KERNEL>
  {
    using Base = dfrg::tmpl::CartesianHost<3, NT, KERNEL>;

  public:
    using ctype = typename Base::ctype;

    template <dfrg::JsonLike JSON>
    Integrator3DCartesianTBB(QuadratureProvider &quadrature_provider, const std::array<unsigned, 3> grid_sizes,
                             const ctype x_extent, const JSON &json)
        : Integrator3DCartesianTBB(quadrature_provider, grid_sizes, x_extent, 0,
                                   json.get_double("/discretization/integration/qx_min", -M_PI),
                                   json.get_double("/discretization/integration/qy_min", -M_PI),
                                   json.get_double("/discretization/integration/qz_min", -M_PI),
                                   json.get_double("/discretization/integration/qx_max", M_PI),
                                   json.get_double("/discretization/integration/qy_max", M_PI),
                                   json.get_double("/discretization/integration/qz_max", M_PI))
    {
    }
    Integrator3DCartesianTBB(QuadratureProvider &quadrature_provider, const std::array<unsigned, 3> grid_sizes,
                             const ctype x_extent = 0., const unsigned max_block_size = 0,
                             const ctype qx_min = -M_PI, const ctype qy_min = -M_PI, const ctype qz_min = -M_PI,
                             const ctype qx_max = M_PI, const ctype qy_max = M_PI, const ctype qz_max = M_PI)
        : Base(quadrature_provider, grid_sizes, {{qx_min, qy_min, qz_min}}, {{qx_max, qy_max, qz_max}})
    {
      (void)x_extent;
      (void)max_block_size;
    }

    void set_qx_min(const ctype qx_min) { this->set_min(0, qx_min); }
    void set_qy_min(const ctype qy_min) { this->set_min(1, qy_min); }
    void set_qz_min(const ctype qz_min) { this->set_min(2, qz_min); }
    void set_qx_max(const ctype qx_max) { this->set_max(0, qx_max); }
    void set_qy_max(const ctype qy_max) { this->set_max(1, qy_max); }
    void set_qz_max(const ctype qz_max) { this->set_max(2, qz_max); }
  };
} // namespace DiFfRG
