// DiFfRG/physics/integration_finiteT/integrator_finiteTq0_cpu.hh -- IntegratorFiniteTq0TBB<d, NT,
// KERNEL>, the host-side finite-T integrator with the reference interface
// (integrator_finiteTq0_cpu.hh:14-134).
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <future>
#include <limits>
#include <vector>

#include "dfrg/math.hh"
#include <DiFfRG/common/quadrature/quadrature_provider.hh>

namespace DiFfRG
{
  template <int d, typename NT, typename KERNEL> class IntegratorFiniteTq0TBB
  {
    static_assert(d >= 2, "dimension must be at least 2");

  public:
    using ctype = typename dfrg::ctype_t<NT>;

    template <dfrg::JsonLike JSON>
    IntegratorFiniteTq0TBB(QuadratureProvider &qp, const std::array<unsigned, 1> grid_sizes, const ctype x_extent,
                           const JSON &json)
        : IntegratorFiniteTq0TBB(qp, grid_sizes, x_extent, json.get_double("/physical/T"))
    {
    }
    IntegratorFiniteTq0TBB(QuadratureProvider &qp, const std::array<unsigned, 1> _grid_sizes, const ctype x_extent,
                           const ctype T, const unsigned max_block_size = 0)
        : quadrature_provider(qp), grid_sizes{{_grid_sizes[0], 0}}, x_extent(x_extent),
          m_T(std::numeric_limits<ctype>::quiet_NaN()), m_E(0), manual_E(false)
    {
      (void)max_block_size;
      ptr_x_quadrature_p = qp.get_points<ctype>(grid_sizes[0]).data();
      ptr_x_quadrature_w = qp.get_weights<ctype>(grid_sizes[0]).data();
      set_T(T);
    }

    void set_T(const ctype T, const ctype E = 0)
    {
      if (dfrg::is_close(T, m_T) && (std::abs(E - m_E) / std::max(E, m_E) < 2.5e-2)) return;
      m_T = T;
      m_E = dfrg::is_close(E, (ctype)0) ? 10 * m_T : E;
      manual_E = !dfrg::is_close(E, (ctype)0);
      const auto &mp = quadrature_provider.get_matsubara_points<ctype>(m_T, m_E);
      grid_sizes[1] = (unsigned)mp.size();
      ptr_matsubara_quadrature_p = mp.data();
      ptr_matsubara_quadrature_w = quadrature_provider.get_matsubara_weights<ctype>(m_T, m_E).data();
    }
    void set_E(const ctype E) { set_T(m_T, E); }

    template <typename... T> NT get(const ctype k, const T &...t)
    {
      if (!manual_E && (std::abs(k - m_E) / std::max(k, m_E) > 2.5e-2)) {
        set_T(m_T, k);
        manual_E = false;
      }
      constexpr ctype S_dm1 = dfrg::S_d_prec<ctype>(d - 1);
      std::vector<NT> part(grid_sizes[0], NT(0));
#pragma omp parallel for schedule(static)
      for (int idx_x = 0; idx_x < (int)grid_sizes[0]; ++idx_x) {
        const ctype q = k * std::sqrt(ptr_x_quadrature_p[idx_x] * x_extent);
        const ctype x_weight = ptr_x_quadrature_w[idx_x] * x_extent;
        const ctype int_element = S_dm1 * (dfrg::powr<d - 3>(q) / (ctype)2 * dfrg::powr<2>(k)) / dfrg::powr<d - 1>(2 * (ctype)M_PI);
        NT value = NT(0);
        for (unsigned idx_y = 0; idx_y < grid_sizes[1]; ++idx_y) {
          const ctype q0 = ptr_matsubara_quadrature_p[idx_y];
          const ctype weight = x_weight * ptr_matsubara_quadrature_w[idx_y];
          value += int_element * weight * (KERNEL::kernel(q, q0, k, t...) + KERNEL::kernel(q, -q0, k, t...));
          if (m_T > 0 && idx_y == 0) value += int_element * x_weight * m_T * KERNEL::kernel(q, (ctype)0, k, t...);
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
    QuadratureProvider &quadrature_provider;
    std::array<unsigned, 2> grid_sizes;
    const ctype x_extent;
    ctype m_T, m_E;
    bool manual_E;
    const ctype *ptr_x_quadrature_p, *ptr_x_quadrature_w;
    const ctype *ptr_matsubara_quadrature_p = nullptr, *ptr_matsubara_quadrature_w = nullptr;
  };
} // namespace DiFfRG
