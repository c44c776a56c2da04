// DiFfRG/physics/integration/integrator_3Dpx0_gpu.hh -- Integrator3Dpx0GPU<NT, KERNEL> (reference
// integrator_3Dpx0_gpu.hh:49-177) on the tile engine of dfrg/integrate.hh, 3+1 dimensions:
//   get(k, t...) = KERNEL::constant(k, t...)
//                  + (2pi)^-4 int dq q^2 int dcos int_0^{2pi} dphi int_0 dq0
//                      (KERNEL::kernel(q, cos, phi, q0, k, t...) + KERNEL::kernel(q, cos, phi, -q0, k, t...))
// with q^2 in [0, x_extent k^2] and q0 = k x0 with x0 in [-x0_extent, x0_extent]; grid_sizes = {q, cos, phi, q0}, the
// angular orders equal.  The q0 axis has its own order grid_sizes[3]; the reference's flat
// (phi, q0) index reads the q0 rule correctly only when that order equals the angular one, the
// case its tests use, and there the two agree.  get() is one launch; request()s are batched.
#pragma once

#include <array>

#include "dfrg/frequency_gpu.hh"

namespace DiFfRG
{
  template <typename NT, typename KERNEL>
  class Integrator3Dpx0GPU : public dfrg::tmpl::FrequencyGPU<dfrg::tmpl::Angle1PhiQ0<true>, true, NT, KERNEL>
  {
    using Base = dfrg::tmpl::FrequencyGPU<dfrg::tmpl::Angle1PhiQ0<true>, true, NT, KERNEL>;

  public:
    using ctype = typename Base::ctype;

    template <dfrg::JsonLike JSON>
    Integrator3Dpx0GPU(QuadratureProvider &quadrature_provider, const std::array<unsigned, 4> grid_sizes,
                       const ctype x_extent, const ctype x0_extent, const JSON &json)
        : Integrator3Dpx0GPU(quadrature_provider, grid_sizes, x_extent, x0_extent,
                             json.get_uint("/integration/cudathreadsperblock", 256))
    {
    }
    Integrator3Dpx0GPU(QuadratureProvider &quadrature_provider, std::array<unsigned, 4> grid_sizes, const ctype x_extent,
                       const ctype x0_extent, const unsigned max_block_size = 256)
        : Base((dfrg::tmpl::require_equal_angles(grid_sizes[2], grid_sizes[1]), quadrature_provider), grid_sizes[0],
               grid_sizes[1], grid_sizes[3], x_extent, x0_extent)
    {
      (void)max_block_size;
    }
  };
} // namespace DiFfRG
