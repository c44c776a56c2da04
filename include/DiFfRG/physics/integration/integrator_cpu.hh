// DiFfRG/physics/integration/integrator_cpu.hh -- IntegratorTBB<d, NT, KERNEL>: the host-side
// radial integrator with the reference interface (integrator_cpu.hh:14-95).  The loop is the
// reference's (element, weight, kernel per node), parallelised with OpenMP where available;
// partial sums per node are added in node order.
#pragma once

#include <array>
#include <future>
#include <vector>

#include "dfrg/math.hh"
#include <DiFfRG/common/quadrature/quadrature_provider.hh>

namespace DiFfRG
{
  template <int d, typename NT, typename KERNEL> class IntegratorTBB
  {
  public:
    using ctype = typename dfrg::ctype_t<NT>;
    IntegratorTBB(QuadratureProvider &qp, const std::array<unsigned, 1> grid_size, const ctype x_extent,
                  const unsigned max_block_size = 0)
        : IntegratorTBB(qp, grid_size[0], x_extent, max_block_size)
    {
    }
    template <dfrg::JsonLike JSON>
    IntegratorTBB(QuadratureProvider &qp, const std::array<unsigned, 1> grid_size, const ctype x_extent, const JSON &)
        : IntegratorTBB(qp, grid_size[0], x_extent)
    {
    }
    IntegratorTBB(QuadratureProvider &qp, const unsigned grid_size, const ctype x_extent, const unsigned max_block_size = 0)
        : grid_size(grid_size), x_extent(x_extent), x_quadrature_p(qp.get_points<ctype>(grid_size)),
          x_quadrature_w(qp.get_weights<ctype>(grid_size))
    {
      (void)max_block_size;
    }

    template <typename... T> NT get(const ctype k, const T &...t) const
    {
      constexpr ctype S_d = dfrg::S_d_prec<ctype>(d);
      std::vector<NT> part(grid_size, NT(0));
#pragma omp parallel for schedule(static)
      for (int idx_x = 0; idx_x < (int)grid_size; ++idx_x) {
        const ctype q = k * std::sqrt(x_quadrature_p[idx_x] * x_extent);
        const ctype int_element = S_d * (dfrg::powr<d - 2>(q) / (ctype)2 * dfrg::powr<2>(k)) / dfrg::powr<d>(2 * (ctype)M_PI);
        const ctype weight = x_quadrature_w[idx_x] * x_extent;
        part[idx_x] = int_element * weight * KERNEL::kernel(q, k, t...);
      }
      NT value = NT(0);
      for (const auto &v : part) value += v;
      return KERNEL::constant(k, t...) + value;
    }

    template <typename... T> std::future<NT> request(const ctype k, const T &...t) const
    {
      return std::async(std::launch::deferred, [=, this]() { return get(k, t...); });
    }

  private:
    const unsigned grid_size;
    const ctype x_extent;
    const std::vector<ctype> &x_quadrature_p;
    const std::vector<ctype> &x_quadrature_w;
  };
} // namespace DiFfRG
