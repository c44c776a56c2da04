// DiFfRG/physics/integration/integrator_2D_cartesian_gpu.hh -- Integrator2DCartesianGPU<NT, KERNEL>
// (reference integrator_2D_cartesian_gpu.hh:54-240) on the tile engine of dfrg/integrate.hh:
//   get(k, t...) = KERNEL::constant(k, t...) + (2 pi)^-2 int dqx dqy KERNEL::kernel(qx, qy, k, t...)
// over [qx_min, qx_max] x [qy_min, qy_max].  As in the reference, the JSON constructor reads only
// the block size (the range keys are read by the 1D and 3D forms and by the host form); x_extent
// and max_block_size are accepted and unused.  get() is one launch; request()s are batched.
#pragma once

#include <array>
#include <cmath>

#include "dfrg/cartesian_gpu.hh"

namespace DiFfRG
{
  template <typename NT, typename KERNEL> class Integrator2DCartesianGPU : public dfrg::tmpl::CartesianGPU<2, NT, KERNEL>
  {
    using Base = dfrg::tmpl::CartesianGPU<2, NT, KERNEL>;

  public:
    using ctype = typename Base::ctype;

    template <dfrg::JsonLike JSON>
    Integrator2DCartesianGPU(QuadratureProvider &quadrature_provider, const std::array<unsigned, 2> grid_sizes,
                             const ctype x_extent, const JSON &json)
        : Integrator2DCartesianGPU(quadrature_provider, grid_sizes, x_extent,
                                   json.get_uint("/integration/cudathreadsperblock", 256))
    {
    }
    Integrator2DCartesianGPU(QuadratureProvider &quadrature_provider, std::array<unsigned, 2> grid_sizes,
                             const ctype x_extent = 0., const unsigned max_block_size = 256,
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
