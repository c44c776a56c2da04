// DiFfRG/physics/integration/integrator_4D_2ang_cpu.hh -- Integrator4D2AngTBB<NT, KERNEL>, the
// host-side two-angle integrator with the reference interface (integrator_4D_2ang_cpu.hh:25-119).
#pragma once

#include <array>
#include <future>
#include <vector>

#include "dfrg/math.hh"
#include <DiFfRG/common/quadrature/quadrature_provider.hh>

namespace DiFfRG
{
  template <typename NT, typename KERNEL> class Integrator4D2AngTBB
  {
  public:
    using ctype = typename dfrg::ctype_t<NT>;
    template <dfrg::JsonLike JSON>
    Integrator4D2AngTBB(QuadratureProvider &qp, const std::array<unsigned, 3> grid_sizes, const ctype x_extent,
                        const JSON &)
        : Integrator4D2AngTBB(qp, grid_sizes, x_extent)
    {
    }
    Integrator4D2AngTBB(QuadratureProvider &qp, std::array<unsigned, 3> grid_sizes, const ctype x_extent,
                        const unsigned max_block_size = 0)
        : grid_sizes(grid_sizes), x_extent(x_extent), x_quadrature_p(qp.get_points<ctype>(grid_sizes[0])),
          x_quadrature_w(qp.get_weights<ctype>(grid_sizes[0])), ang_quadrature_p1(qp.get_points<ctype>(grid_sizes[1])),
          ang_quadrature_w1(qp.get_weights<ctype>(grid_sizes[1])), ang_quadrature_p2(qp.get_points<ctype>(grid_sizes[2])),
          ang_quadrature_w2(qp.get_weights<ctype>(grid_sizes[2]))
    {
      (void)max_block_size;
    }

    template <typename... T> NT get(const ctype k, const T &...t) const
    {
      std::vector<NT> part(grid_sizes[0], NT(0));
#pragma omp parallel for schedule(static)
      for (int idx_x = 0; idx_x < (int)grid_sizes[0]; ++idx_x) {
        const ctype q = k * std::sqrt(x_quadrature_p[idx_x] * x_extent);
        NT value = NT(0);
        for (unsigned idx_y = 0; idx_y < grid_sizes[1]; ++idx_y) {
          const ctype cos1 = 2 * (ang_quadrature_p1[idx_y] - (ctype)0.5);
          const ctype int_element = 2 * (ctype)M_PI * (dfrg::powr<2>(q) * (ctype)0.5 * dfrg::powr<2>(k)) *
                                    std::sqrt(1. - dfrg::powr<2>(cos1)) / dfrg::powr<4>(2 * (ctype)M_PI);
          for (unsigned idx_z = 0; idx_z < grid_sizes[2]; ++idx_z) {
            const ctype cos2 = 2 * (ang_quadrature_p2[idx_z] - (ctype)0.5);
            const ctype weight =
                2 * ang_quadrature_w2[idx_z] * 2 * ang_quadrature_w1[idx_y] * x_quadrature_w[idx_x] * x_extent;
            value += int_element * weight * KERNEL::kernel(q, cos1, cos2, k, t...);
          }
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
    const std::array<unsigned, 3> grid_sizes;
    const ctype x_extent;
    const std::vector<ctype> &x_quadrature_p, &x_quadrature_w, &ang_quadrature_p1, &ang_quadrature_w1,
        &ang_quadrature_p2, &ang_quadrature_w2;
  };
} // namespace DiFfRG
