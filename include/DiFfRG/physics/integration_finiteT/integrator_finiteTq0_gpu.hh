// DiFfRG/physics/integration_finiteT/integrator_finiteTq0_gpu.hh -- IntegratorFiniteTq0GPU<d, NT,
// KERNEL> (reference integrator_finiteTq0_gpu.hh:46-340): (d-1)-dimensional radial integral and
// the Matsubara sum  T sum_n f(q, 2 pi n T) ~ sum_i w_i (f(q, x_i) + f(q, -x_i)) + T f(q, 0).
// The Matsubara rule follows the typical energy scale E, which tracks k with a 2.5% hysteresis
// unless set by set_E / set_T(T, E) (integrator_finiteTq0_gpu.hh:105-126, 276-281).
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <future>
#include <limits>
#include <memory>
#include <tuple>

#include <DiFfRG/common/json.hh>
#include <DiFfRG/common/math.hh>
#include <DiFfRG/common/quadrature/quadrature_provider.hh>
#include <DiFfRG/common/types.hh>

#include "dfrg/integrate.hh"

namespace DiFfRG
{
  template <int d, typename NT, typename KERNEL> class IntegratorFiniteTq0GPU
  {
    static_assert(d >= 2, "dimension must be at least 2");

  public:
    using ctype = typename get_type::ctype<NT>;

    IntegratorFiniteTq0GPU(QuadratureProvider &quadrature_provider, const std::array<uint, 1> grid_sizes,
                           const ctype x_extent, const JSONValue &json)
        : IntegratorFiniteTq0GPU(quadrature_provider, grid_sizes, x_extent, json.get_double("/physical/T"),
                                 json.get_uint("/integration/cudathreadsperblock", 256))
    {
    }
    IntegratorFiniteTq0GPU(QuadratureProvider &quadrature_provider, const std::array<uint, 1> _grid_sizes,
                           const ctype x_extent, const ctype T, const uint max_block_size = 256)
        : quadrature_provider(quadrature_provider), grid_sizes{{_grid_sizes[0], 0}}, x_extent(x_extent),
          m_T(std::numeric_limits<ctype>::quiet_NaN()), m_E(0), manual_E(false)
    {
      (void)max_block_size;
      ptr_x_quadrature_p = quadrature_provider.get_device_points<ctype>(grid_sizes[0]);
      ptr_x_quadrature_w = quadrature_provider.get_device_weights<ctype>(grid_sizes[0]);
      set_T(T);
    }

    // Set T and the typical energy scale E (E = 0: default 10 T) and fetch the Matsubara rule.
    void set_T(const ctype T, const ctype E = 0)
    {
      if (is_close(T, m_T) && E != 0 && (std::abs(E - m_E) / std::max(E, m_E) < 2.5e-2)) return;
      m_T = T;
      m_E = is_close(E, (ctype)0) ? 10 * m_T : E;
      manual_E = !is_close(E, (ctype)0);
      grid_sizes[1] = (uint)quadrature_provider.get_matsubara_points<ctype>(m_T, m_E).size();
      ptr_matsubara_quadrature_p = quadrature_provider.get_device_matsubara_points<ctype>(m_T, m_E);
      ptr_matsubara_quadrature_w = quadrature_provider.get_device_matsubara_weights<ctype>(m_T, m_E);
    }
    void set_E(const ctype E) { set_T(m_T, E); }

    template <typename... T> NT get(const ctype k, const T &...t) { return launch(k, t...).wait(); }

    // Deferred: requests are batched, and reading one future launches the whole batch
    // (dfrg::tmpl::RequestBatcher); a change of the Matsubara rule (T, or E following k) starts a
    // new batch.  Results are bitwise those of get().
    template <typename... T> std::future<NT> request(const ctype k, const T &...t)
    {
      track_E(k);
      using Args = std::tuple<ctype, T...>;
      const long long nodes = (long long)grid_sizes[0] * grid_sizes[1];
      const unsigned long long state = (unsigned long long)(uintptr_t)ptr_matsubara_quadrature_p;
      return batcher->enqueue(state, dfrg::tmpl::grid_for(nodes), KERNEL::constant(k, t...), Args(k, t...), [this]() {
        return [xq = ptr_x_quadrature_p, wq = ptr_x_quadrature_w, mx = ptr_matsubara_quadrature_p,
                mw = ptr_matsubara_quadrature_w, nx = (int)grid_sizes[0], nm = (int)grid_sizes[1], xe = x_extent,
                temp = m_T](const Args *a, int r, int nb, NT *part, unsigned *cnt, NT *res, hipStream_t s) {
          hipLaunchKernelGGL((dfrg::tmpl::finiteTq0_batch_kernel<d, ctype, NT, KERNEL, Args>), dim3(r * nb),
                             dim3(dfrg::tmpl::kBlock), 0, s, xq, wq, mx, mw, nx, nm, xe, temp, a, nb, part, cnt, res);
        };
      });
    }

    uint matsubara_size() const { return grid_sizes[1]; }
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

    template <typename... T> dfrg::tmpl::Pending<NT> launch(const ctype k, const T &...t)
    {
      track_E(k);
      auto p = dfrg::tmpl::Pending<NT>::make();
      p.constant = KERNEL::constant(k, t...);
      const long long nodes = (long long)grid_sizes[0] * grid_sizes[1];
      hipLaunchKernelGGL((dfrg::tmpl::finiteTq0_kernel<d, ctype, NT, KERNEL, T...>),
                         dim3(dfrg::tmpl::grid_for(nodes)), dim3(dfrg::tmpl::kBlock), 0, p.stream, ptr_x_quadrature_p,
                         ptr_x_quadrature_w, ptr_matsubara_quadrature_p, ptr_matsubara_quadrature_w,
                         (int)grid_sizes[0], (int)grid_sizes[1], x_extent, m_T, k, p.partials(), p.counter(),
                         p.result(), t...);
      p.enqueue_result();
      return p;
    }

    QuadratureProvider &quadrature_provider;
    std::array<uint, 2> grid_sizes;
    ctype x_extent;
    ctype m_T, m_E;
    bool manual_E;
    const ctype *ptr_x_quadrature_p, *ptr_x_quadrature_w;
    const ctype *ptr_matsubara_quadrature_p = nullptr, *ptr_matsubara_quadrature_w = nullptr;
    std::shared_ptr<dfrg::tmpl::RequestBatcher<NT>> batcher = std::make_shared<dfrg::tmpl::RequestBatcher<NT>>();
  };
} // namespace DiFfRG
