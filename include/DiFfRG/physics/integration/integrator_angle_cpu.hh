// DiFfRG/physics/integration/integrator_angle_cpu.hh -- IntegratorAngleTBB<d, NT, KERNEL>, the
// host-side angle integrator with the reference interface (integrator_angle_cpu.hh:24-108).
#pragma once

#include <array>
#include <cmath>
#include <future>
#include <vector>

#include "dfrg/math.hh"
#include <DiFfRG/common/quadrature/quadrature_provider.hh>

namespace DiFfRG
{
  template <int d, typename NT, typename KERNEL> class IntegratorAngleTBB
  {
    static_assert(d > 1, "IntegratorAngleTBB: d must be greater than 1, otherwise angles are not needed");

  public:
    using ctype = typename dfrg::ctype_t<NT>;

    template <dfrg::JsonLike JSON>
    IntegratorAngleTBB(QuadratureProvider &quadrature_provider, const std::array<unsigned, 2> grid_sizes,
                       const ctype x_extent, const JSON &)
        : IntegratorAngleTBB(quadrature_provider, grid_sizes, x_extent)
    {
    }
    IntegratorAngleTBB(QuadratureProvider &quadrature_provider, std::array<unsigned, 2> grid_sizes, const ctype x_extent,
                       const unsigned max_block_size = 0)
        : grid_sizes(grid_sizes), x_extent(x_extent),
          x_quadrature_p(quadrature_provider.get_points<ctype>(grid_sizes[0])),
          x_quadrature_w(quadrature_provider.get_weights<ctype>(grid_sizes[0])),
          ang_quadrature_p(quadrature_provider.get_points<ctype>(grid_sizes[1])),
          ang_quadrature_w(quadrature_provider.get_weights<ctype>(grid_sizes[1]))
    {
      (void)max_block_size;
    }

    template <typename... T> NT get(const ctype k, const T &...t) const
    {
      constexpr ctype S_d = dfrg::S_d_prec<ctype>(d);
      std::vector<NT> part(grid_sizes[0], NT(0));
#pragma omp parallel for schedule(static)
      for (int idx_x = 0; idx_x < (int)grid_sizes[0]; ++idx_x) {
        const ctype q = k * std::sqrt(x_quadrature_p[idx_x] * x_extent);
        const ctype int_element = S_d * (dfrg::powr<d - 2>(q) / (ctype)2 * dfrg::powr<2>(k)) * (1 / (ctype)2) /
                                  dfrg::powr<d>(2 * (ctype)M_PI);
        NT value = NT(0);
        for (unsigned idx_y = 0; idx_y < grid_sizes[1]; ++idx_y) {
          const ctype cos = 2 * (ang_quadrature_p[idx_y] - (ctype)0.5);
          const ctype weight = 2 * ang_quadrature_w[idx_y] * x_quadrature_w[idx_x] * x_extent;
          value += int_element * weight * KERNEL::kernel(q, cos, k, t...);
        }
        part[idx_x] = value;
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
    std::array<unsigned, 2> grid_sizes;
    const ctype x_extent;
    const std::vector<ctype> &x_quadrature_p, &x_quadrature_w, &ang_quadrature_p, &ang_quadrature_w;
  };
} // namespace DiFfRG
