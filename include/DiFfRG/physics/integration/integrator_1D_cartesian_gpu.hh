// DiFfRG/physics/integration/integrator_1D_cartesian_gpu.hh -- Integrator1DCartesianGPU<NT, KERNEL>
// (reference integrator_1D_cartesian_gpu.hh:27-203) on the tile engine of dfrg/integrate.hh:
//   get(k, t...) = KERNEL::constant(k, t...) + (2 pi)^-1 int_{qx_min}^{qx_max} dqx KERNEL::kernel(qx, k, t...)
// x_extent and max_block_size are accepted and unused, as in the reference.  get() is one launch;
// request()s are batched (bitwise the values of get()).
#pragma once

#include <array>
#include <cmath>

#include "dfrg/cartesian_gpu.hh"

namespace DiFfRG
{
  template <typename NT, typename KERNEL> class Integrator1DCartesianGPU : public dfrg::tmpl::CartesianGPU<1, NT, KERNEL>
  {
    using Base = dfrg::tmpl::CartesianGPU<1, NT, KERNEL>;

  public:
    using ctype = typename Base::ctype;

    Integrator1DCartesianGPU(QuadratureProvider &quadrature_provider, const std::array<unsigned, 1> grid_size,
                             const ctype x_extent = 0., const unsigned max_block_size = 0,
                             const ctype qx_min = -M_PI, const ctype qx_max = M_PI)
        : Integrator1DCartesianGPU(quadrature_provider, grid_size[0], x_extent, qx_min, qx_max)
    {
      (void)max_block_size;
    }
    template <dfrg::JsonLike JSON>
    Integrator1DCartesianGPU(QuadratureProvider &quadrature_provider, const std::array<unsigned, 1> grid_size,
                             const JSON &json)
        : Integrator1DCartesianGPU(quadrature_provider, grid_size[0], 0.,
                                   json.get_double("/discretization/integration/qx_min", -M_PI),
                                   json.get_double("/discretization/integration/qx_max", M_PI))
    {
    }
    template <dfrg::JsonLike JSON>
    Integrator1DCartesianGPU(QuadratureProvider &quadrature_provider, const std::array<unsigned, 1> grid_size,
                             const ctype x_extent, const JSON &json)
        : Integrator1DCartesianGPU(quadrature_provider, grid_size[0], x_extent,
                                   json.get_double("/discretization/integration/qx_min", -M_PI),
                                   json.get_double("/discretization/integration/qx_max", M_PI))
    {
    }
    Integrator1DCartesianGPU(QuadratureProvider &quadrature_provider, const unsigned grid_size,
                             const ctype x_extent = 0., const ctype qx_min = -M_PI, const ctype qx_max = M_PI)
        : Base(quadrature_provider, {{grid_size}}, {{qx_min}}, {{qx_max}})
    {
      (void)x_extent;
    }

    void set_qx_min(const ctype qx_min) { this->set_min(0, qx_min); }
    void set_qx_max(const ctype qx_max) { this->set_max(0, qx_max); }
  };
} // namespace DiFfRG
