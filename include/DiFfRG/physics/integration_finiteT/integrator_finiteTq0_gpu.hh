// DiFfRG/physics/integration_finiteT/integrator_finiteTq0_gpu.hh -- IntegratorFiniteTq0GPU<d, NT,
// KERNEL> (reference integrator_finiteTq0_gpu.hh:46-340): (d-1)-dimensional radial integral and
// the Matsubara sum  T sum_n f(q, 2 pi n T) ~ sum_i w_i (f(q, x_i) + f(q, -x_i)) + T f(q, 0), on the
// tile engine of dfrg/integrate.hh.  The Matsubara rule follows the typical energy scale E, which
// tracks k with a 2.5% hysteresis unless set by set_E / set_T(T, E) (integrator_finiteTq0_gpu.hh:
// 105-126, 276-281); a new rule starts a new batch of requests.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <future>
#include <limits>
#include <memory>

#include <DiFfRG/common/quadrature/quadrature_provider.hh>

#include "dfrg/integrate.hh"
#include "dfrg/math.hh"

namespace DiFfRG
{
  template <int d, typename NT, typename KERNEL> class IntegratorFiniteTq0GPU
  {
    static_assert(d >= 2, "dimension must be at least 2");
    using Shape = dfrg::tmpl::FiniteT<d>;

  public:
    using ctype = dfrg::ctype_t<NT>;

    template <dfrg::JsonLike JSON>
    IntegratorFiniteTq0GPU(QuadratureProvider &quadrature_provider, const std::array<unsigned, 1> grid_sizes,
                           const ctype x_extent, const JSON &json)
        : IntegratorFiniteTq0GPU(quadrature_provider, grid_sizes, x_extent, json.get_double("/physical/T"),
                                 json.get_uint("/integration/cudathreadsperblock", 256))
    {
    }
    IntegratorFiniteTq0GPU(QuadratureProvider &quadrature_provider, const std::array<unsigned, 1> grid_sizes,
                           const ctype x_extent, const ctype T, const unsigned max_block_size = 256)
        : quadrature_provider(quadrature_provider), m_T(std::numeric_limits<ctype>::quiet_NaN()), m_E(0),
          manual_E(false)
    {
      (void)max_block_size;
      base.xq = quadrature_provider.get_device_points<ctype>(grid_sizes[0]);
      base.wq = quadrature_provider.get_device_weights<ctype>(grid_sizes[0]);
      base.nx = (int)grid_sizes[0];
      base.x_extent = x_extent;
      set_T(T);
    }

    // Set T and the typical energy scale E (E = 0: default 10 T) and fetch the Matsubara rule.
    void set_T(const ctype T, const ctype E = 0)
    {
      if (dfrg::is_close(T, m_T) && E != 0 && (std::abs(E - m_E) / std::max(E, m_E) < 2.5e-2)) return;
      m_T = T;
      m_E = dfrg::is_close(E, (ctype)0) ? 10 * m_T : E;
      manual_E = !dfrg::is_close(E, (ctype)0);
      base.nm = (int)quadrature_provider.get_matsubara_points<ctype>(m_T, m_E).size();
      base.mx = quadrature_provider.get_device_matsubara_points<ctype>(m_T, m_E);
      base.mw = quadrature_provider.get_device_matsubara_weights<ctype>(m_T, m_E);
      base.m_T = m_T;
    }
    void set_E(const ctype E) { set_T(m_T, E); }

    template <typename... T> NT get(const ctype k, const T &...t)
    {
      track_E(k);
      return dfrg::tmpl::get<Shape, NT, KERNEL>(base, k, t...);
    }
    template <typename... T> std::future<NT> request(const ctype k, const T &...t)
    {
      track_E(k);
      return dfrg::tmpl::request<Shape, NT, KERNEL>(*batcher, (unsigned long long)(uintptr_t)base.mx, base, k, t...);
    }

    unsigned matsubara_size() const { return (unsigned)base.nm; }
    ctype typical_E() const { return m_E; }

  private:
    // E follows k with a 2.5% hysteresis unless set explicitly (integrator_finiteTq0_gpu.hh:276-281)
    void track_E(const ctype k)
    {
      if (!manual_E && (std::abs(k - m_E) / std::max(k, m_E) > 2.5e-2)) {
        set_T(m_T, k);
        manual_E = false;
      }
    }

    QuadratureProvider &quadrature_provider;
    dfrg::tmpl::Nodes<ctype> base;
    ctype m_T, m_E;
    bool manual_E;
    std::shared_ptr<dfrg::tmpl::RequestBatcher<NT>> batcher = std::make_shared<dfrg::tmpl::RequestBatcher<NT>>();
  };
} // namespace DiFfRG
