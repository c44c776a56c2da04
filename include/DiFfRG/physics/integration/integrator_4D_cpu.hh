// DiFfRG/physics/integration/integrator_4D_cpu.hh -- Integrator4DTBB<NT, KERNEL>, the host-side
// three-angle integrator with the reference interface (integrator_4D_cpu.hh:25-128).  The
// reference's flat (cos2, phi) index is only right for equal orders; this loop indexes each
// angle separately.
#pragma once

#include <array>
#include <future>
#include <vector>

#include "dfrg/math.hh"
#include <DiFfRG/common/quadrature/quadrature_provider.hh>

namespace DiFfRG
{
  template <typename NT, typename KERNEL> class Integrator4DTBB
  {
  public:
    using ctype = typename dfrg::ctype_t<NT>;
    template <dfrg::JsonLike JSON>
    Integrator4DTBB(QuadratureProvider &qp, const std::array<unsigned, 4> grid_sizes, const ctype x_extent,
                    const JSON &)
        : Integrator4DTBB(qp, grid_sizes, x_extent)
    {
    }
    Integrator4DTBB(QuadratureProvider &qp, std::array<unsigned, 4> grid_sizes, const ctype x_extent,
                    const unsigned max_block_size = 0)
        : grid_sizes(grid_sizes), x_extent(x_extent), xp(qp.get_points<ctype>(grid_sizes[0])),
          xw(qp.get_weights<ctype>(grid_sizes[0])), p1(qp.get_points<ctype>(grid_sizes[1])),
          w1(qp.get_weights<ctype>(grid_sizes[1])), p2(qp.get_points<ctype>(grid_sizes[2])),
          w2(qp.get_weights<ctype>(grid_sizes[2])), p3(qp.get_points<ctype>(grid_sizes[3])),
          w3(qp.get_weights<ctype>(grid_sizes[3]))
    {
      (void)max_block_size;
    }

    template <typename... T> NT get(const ctype k, const T &...t) const
    {
      std::vector<NT> part(grid_sizes[0], NT(0));
#pragma omp parallel for schedule(static)
      for (int idx_x = 0; idx_x < (int)grid_sizes[0]; ++idx_x) {
        const ctype q = k * std::sqrt(xp[idx_x] * x_extent);
        NT value = NT(0);
        for (unsigned i1 = 0; i1 < grid_sizes[1]; ++i1) {
          const ctype cos1 = 2 * (p1[i1] - (ctype)0.5);
          const ctype int_element =
              (dfrg::powr<2>(q) * (ctype)0.5 * dfrg::powr<2>(k)) * std::sqrt(1. - dfrg::powr<2>(cos1)) / dfrg::powr<4>(2 * (ctype)M_PI);
          for (unsigned i2 = 0; i2 < grid_sizes[2]; ++i2) {
            const ctype cos2 = 2 * (p2[i2] - (ctype)0.5);
            for (unsigned i3 = 0; i3 < grid_sizes[3]; ++i3) {
              const ctype phi = 2 * (ctype)M_PI * p3[i3];
              const ctype weight = 2 * (ctype)M_PI * w3[i3] * 2 * w2[i2] * 2 * w1[i1] * xw[idx_x] * x_extent;
              value += int_element * weight * KERNEL::kernel(q, cos1, cos2, phi, k, t...);
            }
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
    const std::array<unsigned, 4> grid_sizes;
    const ctype x_extent;
    const std::vector<ctype> &xp, &xw, &p1, &w1, &p2, &w2, &p3, &w3;
  };
} // namespace DiFfRG
