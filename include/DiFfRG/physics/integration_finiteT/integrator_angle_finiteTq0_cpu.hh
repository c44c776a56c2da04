// DiFfRG/physics/integration_finiteT/integrator_angle_finiteTq0_cpu.hh --
// IntegratorAngleFiniteTq0TBB<d, NT, KERNEL>, the host-side form with the reference interface
// (integrator_angle_finiteTq0_cpu.hh:14-140).
#pragma once

#include <array>
#include <cmath>
#include <future>
#include <vector>

#include "dfrg/math.hh"
#include <DiFfRG/common/quadrature/quadrature_provider.hh>

#include "dfrg/matsubara_state.hh"

namespace DiFfRG
{
  template <int d, typename NT, typename KERNEL> class IntegratorAngleFiniteTq0TBB
  {
    static_assert(d >= 3, "dimension must be at least 3");

  public:
    using ctype = typename dfrg::ctype_t<NT>;

    template <dfrg::JsonLike JSON>
    IntegratorAngleFiniteTq0TBB(QuadratureProvider &quadrature_provider, const std::array<unsigned, 2> grid_sizes,
                                const ctype x_extent, const JSON &json)
        : IntegratorAngleFiniteTq0TBB(quadrature_provider, grid_sizes, x_extent, json.get_double("/physical/T"))
    {
    }
    IntegratorAngleFiniteTq0TBB(QuadratureProvider &quadrature_provider, std::array<unsigned, 2> grid_sizes,
                                const ctype x_extent, const ctype T, const unsigned max_block_size = 0)
        : grid_sizes(grid_sizes), x_extent(x_extent)
    {
      (void)max_block_size;
      xq = quadrature_provider.get_points<ctype>(grid_sizes[0]).data();
      wq = quadrature_provider.get_weights<ctype>(grid_sizes[0]).data();
      xc = quadrature_provider.get_points<ctype>(grid_sizes[1]).data();
      wc = quadrature_provider.get_weights<ctype>(grid_sizes[1]).data();
      mats.provider = &quadrature_provider;
      mats.set_T(T);
    }

    void set_T(const ctype T, const ctype E = 0) { mats.set_T(T, E); }
    void set_E(const ctype E) { mats.set_E(E); }

    template <typename... T> NT get(const ctype k, const T &...t)
    {
      mats.track(k);
      constexpr ctype S_dm1 = dfrg::S_d_prec<ctype>(d - 1);
      std::vector<NT> part(grid_sizes[0], NT(0));
#pragma omp parallel for schedule(static)
      for (int idx_x = 0; idx_x < (int)grid_sizes[0]; ++idx_x) {
        const ctype q = k * std::sqrt(xq[idx_x] * x_extent);
        const ctype int_element = S_dm1 * (dfrg::powr<d - 3>(q) / (ctype)2 * dfrg::powr<2>(k)) * (1 / (ctype)2) /
                                  dfrg::powr<d - 1>(2 * (ctype)M_PI);
        NT value = NT(0);
        for (unsigned idx_y = 0; idx_y < grid_sizes[1]; ++idx_y) {
          const ctype cos = 2 * (xc[idx_y] - (ctype)0.5);
          const ctype x_weight = 2 * wc[idx_y] * wq[idx_x] * x_extent;
          for (unsigned idx_z = 0; idx_z < mats.size; ++idx_z) {
            const ctype q0 = mats.x[idx_z];
            const ctype weight = x_weight * mats.w[idx_z];
            value += int_element * weight * (KERNEL::kernel(q, cos, q0, k, t...) + KERNEL::kernel(q, cos, -q0, k, t...));
            if (mats.T > 0 && idx_z == 0) value += int_element * x_weight * mats.T * KERNEL::kernel(q, cos, (ctype)0, k, t...);
          }
        }
        part[idx_x] = value;
      }
      NT value = NT(0);
      for (const auto &v : part) value += v;
      return KERNEL::constant(k, t...) + value;
    }

    template <typename... T> std::future<NT> request(const ctype k, const T &...t)
    {
      return std::async(std::launch::deferred, [=, this]() { return get(k, t...); });
    }

  private:
    std::array<unsigned, 2> grid_sizes;
    const ctype x_extent;
    const ctype *xq, *wq, *xc, *wc;
    dfrg::tmpl::MatsubaraState<ctype, false> mats;
  };
} // namespace DiFfRG
