// DiFfRG/physics/integration/integrator_2Dpq0_gpu.hh -- Integrator2Dpq0GPU<NT, KERNEL> (reference
// integrator_2Dpq0_gpu.hh:58-210) on the tile engine of dfrg/integrate.hh, 2+1 dimensions:
//   get(k, t...) = KERNEL::constant(k, t...)
//                  + S_2/2 (2pi)^-3 int dq q int dcos int_0 dq0 (KERNEL::kernel(q, cos, q0, k, t...)
//                                                            + KERNEL::kernel(q, cos, -q0, k, t...))
// with q^2 in [0, x_extent k^2] and q0 in [-q0_extent, q0_extent].  get() is one launch; request()s are batched.
#pragma once

#include <array>

#include "dfrg/frequency_gpu.hh"

namespace DiFfRG
{
  template <typename NT, typename KERNEL>
  class Integrator2Dpq0GPU : public dfrg::tmpl::FrequencyGPU<dfrg::tmpl::Angle1Q0<false>, false, NT, KERNEL>
  {
    using Base = dfrg::tmpl::FrequencyGPU<dfrg::tmpl::Angle1Q0<false>, false, NT, KERNEL>;

  public:
    using ctype = typename Base::ctype;

    template <dfrg::JsonLike JSON>
    Integrator2Dpq0GPU(QuadratureProvider &quadrature_provider, const std::array<unsigned, 3> grid_sizes,
                       const ctype x_extent, const ctype q0_extent, const JSON &json)
        : Integrator2Dpq0GPU(quadrature_provider, grid_sizes, x_extent, q0_extent,
                             json.get_uint("/integration/cudathreadsperblock", 256))
    {
    }
    Integrator2Dpq0GPU(QuadratureProvider &quadrature_provider, std::array<unsigned, 3> grid_sizes, const ctype x_extent,
                       const ctype q0_extent, const unsigned max_block_size = 256)
        : Base(quadrature_provider, grid_sizes[0], grid_sizes[1], grid_sizes[2], x_extent, q0_extent)
    {
      (void)max_block_size;
    }
  };
} // namespace DiFfRG
