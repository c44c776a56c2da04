// DiFfRG/physics/integration_finiteT/integrator_finiteTx0_cpu.hh -- IntegratorFiniteTx0TBB<d, NT,
// KERNEL>, the host form with the reference interface (integrator_finiteTx0_cpu.hh:14-131); the
// measure and nodes are those of IntegratorFiniteTx0GPU (dfrg/finiteTx0.hh).
#pragma once

#include <array>

#include "dfrg/finiteTx0.hh"

namespace DiFfRG
{
  template <int d, typename NT, typename KERNEL>
  class IntegratorFiniteTx0TBB : public dfrg::tmpl::FiniteTx0Host<d, 0, NT, KERNEL>
  {
    using Base = dfrg::tmpl::FiniteTx0Host<d, 0, NT, KERNEL>;

  public:
    using ctype = typename Base::ctype;

    template <dfrg::JsonLike JSON>
    IntegratorFiniteTx0TBB(QuadratureProvider &quadrature_provider, const std::array<unsigned, 2> grid_sizes,
                           const ctype x_extent, const ctype x0_extent, const unsigned x0_summands, const JSON &json)
        : IntegratorFiniteTx0TBB(quadrature_provider, grid_sizes, x_extent, x0_extent, x0_summands,
                                 json.get_double("/physical/T"))
    {
    }
    IntegratorFiniteTx0TBB(QuadratureProvider &quadrature_provider, const std::array<unsigned, 2> grid_sizes,
                           const ctype x_extent, const ctype x0_extent, const unsigned x0_summands, const ctype T,
                           const unsigned max_block_size = 0)
        : Base(quadrature_provider, grid_sizes[0], 0, grid_sizes[1], x_extent, x0_extent, x0_summands, T)
    {
      (void)max_block_size;
    }
  };
} // namespace DiFfRG
