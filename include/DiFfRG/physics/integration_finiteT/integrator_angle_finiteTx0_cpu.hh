// DiFfRG/physics/integration_finiteT/integrator_angle_finiteTx0_cpu.hh -- IntegratorAngleFiniteTx0TBB<d,
// NT, KERNEL> (reference integrator_angle_finiteTx0_cpu.hh:14-146): (q, cos, q0) in d-1 spatial dimensions with the
// frequency axis in x0 = q0/k, the first x0_summands Matsubara frequencies summed and the rest
// integrated up to x0_extent (dfrg/finiteTx0.hh); grid_sizes = {q, cos, x0 integral nodes}.
#pragma once

#include <array>

#include "dfrg/finiteTx0.hh"

namespace DiFfRG
{
  template <int d, typename NT, typename KERNEL>
  class IntegratorAngleFiniteTx0TBB : public dfrg::tmpl::FiniteTx0Host<d, 1, NT, KERNEL>
  {
    using Base = dfrg::tmpl::FiniteTx0Host<d, 1, NT, KERNEL>;

  public:
    using ctype = typename Base::ctype;

    template <dfrg::JsonLike JSON>
    IntegratorAngleFiniteTx0TBB(QuadratureProvider &quadrature_provider, const std::array<unsigned, 3> grid_sizes,
                                const ctype x_extent, const ctype x0_extent, const unsigned x0_summands,
                                const JSON &json)
        : IntegratorAngleFiniteTx0TBB(quadrature_provider, grid_sizes, x_extent, x0_extent, x0_summands,
                                      json.get_double("/physical/T"))
    {
    }
    IntegratorAngleFiniteTx0TBB(QuadratureProvider &quadrature_provider, const std::array<unsigned, 3> grid_sizes,
                                const ctype x_extent, const ctype x0_extent, const unsigned x0_summands, const ctype T,
                                const unsigned max_block_size = 0)
        : Base(quadrature_provider, grid_sizes[0], grid_sizes[1], grid_sizes[2], x_extent, x0_extent, x0_summands, T)
    {
      (void)max_block_size;
    }
  };
} // namespace DiFfRG
