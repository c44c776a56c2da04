// DiFfRG/physics/integration/integrator_3D_gpu.hh -- Integrator3DGPU<NT, KERNEL> (reference
// integrator_3D_gpu.hh:45-164) on the tile engine of dfrg/integrate.hh:
//   get(k, t...) = KERNEL::constant(k, t...)
//                  + (2pi)^-3 int dq q^2 int dcos int_0^{2pi} dphi KERNEL::kernel(q, cos, phi, k, t...)
// with q^2 in [0, x_extent k^2].  The angular orders must be equal.  get() is one launch;
// request()s are batched (bitwise the values of get()).
#pragma once

#include <array>

#include "dfrg/frequency_gpu.hh"

namespace DiFfRG
{
  template <typename NT, typename KERNEL>
  class Integrator3DGPU : public dfrg::tmpl::FrequencyGPU<dfrg::tmpl::Angle1Phi, true, NT, KERNEL>
  {
    using Base = dfrg::tmpl::FrequencyGPU<dfrg::tmpl::Angle1Phi, true, NT, KERNEL>;

  public:
    using ctype = typename Base::ctype;

    template <dfrg::JsonLike JSON>
    Integrator3DGPU(QuadratureProvider &quadrature_provider, const std::array<unsigned, 3> grid_sizes,
                    const ctype x_extent, const JSON &json)
        : Integrator3DGPU(quadrature_provider, grid_sizes, x_extent,
                          json.get_uint("/integration/cudathreadsperblock", 256))
    {
    }
    Integrator3DGPU(QuadratureProvider &quadrature_provider, std::array<unsigned, 3> grid_sizes, const ctype x_extent,
                    const unsigned max_block_size = 256)
        : Base((dfrg::tmpl::require_equal_angles(grid_sizes[2], grid_sizes[1]), quadrature_provider), grid_sizes[0],
               grid_sizes[1], 0, x_extent, ctype(0))
    {
      (void)max_block_size;
    }
  };
} // namespace DiFfRG
