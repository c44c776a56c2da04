// DiFfRG/physics/integration_finiteT/integrator_4D_finiteTx0_cpu.hh -- Integrator4DFiniteTx0TBB<NT,
// KERNEL> (reference integrator_4D_finiteTx0_cpu.hh:14-148): (q, cos, phi, q0) with the frequency axis in x0 = q0/k,
// the first x0_summands Matsubara frequencies summed and the rest integrated up to x0_extent
// (dfrg/finiteTx0.hh); grid_sizes = {q, cos, phi, x0 integral nodes}, the angular orders equal.
#pragma once

#include <array>
#include <stdexcept>

#include "dfrg/finiteTx0.hh"

namespace DiFfRG
{
  template <typename NT, typename KERNEL>
  class Integrator4DFiniteTx0TBB : public dfrg::tmpl::FiniteTx0Host<4, 2, NT, KERNEL>
  {
    using Base = dfrg::tmpl::FiniteTx0Host<4, 2, NT, KERNEL>;

  public:
    using ctype = typename Base::ctype;

    template <dfrg::JsonLike JSON>
    Integrator4DFiniteTx0TBB(QuadratureProvider &quadrature_provider, const std::array<unsigned, 4> grid_sizes,
                             const ctype x_extent, const ctype x0_extent, const unsigned x0_summands, const JSON &json)
        : Integrator4DFiniteTx0TBB(quadrature_provider, grid_sizes, x_extent, x0_extent, x0_summands,
                                   json.get_double("/physical/T"), json.get_uint("/integration/cudathreadsperblock", 256))
    {
    }
    Integrator4DFiniteTx0TBB(QuadratureProvider &quadrature_provider, const std::array<unsigned, 4> grid_sizes,
                             const ctype x_extent, const ctype x0_extent, const unsigned x0_summands, const ctype T,
                             const unsigned max_block_size = 0)
        : Base((check_angles(grid_sizes), quadrature_provider), grid_sizes[0], grid_sizes[1], grid_sizes[3], x_extent,
               x0_extent, x0_summands, T)
    {
      (void)max_block_size;
    }

  private:
    static int check_angles(const std::array<unsigned, 4> &g)
    {
      if (g[2] != g[1]) throw std::runtime_error("Grid sizes must be currently equal for all angular dimensions!");
      return 0;
    }
  };
} // namespace DiFfRG
