// dfrg/finiteTx0.hh -- the finite-temperature integrators in x0 = q0/k (reference
// integration_finiteT/integrator_{,angle_,4D_}finiteTx0_{gpu,cpu}.hh): a Matsubara sum over the
// first x0_summands frequencies and an integral over the rest of the frequency axis, log-mapped up
// to x0_extent (dfrg::tmpl::FiniteTx0 in dfrg/integrate.hh has the measure).  The frequency axis
// holds grid_sizes.back() + x0_summands nodes; set_T(T) with T = 0 turns the sum off and gives the
// integral all of them (integrator_finiteTx0_gpu.hh:110-119).  FiniteTx0GPU runs on the tile
// engine (a new T or x0_extent starts a new batch of requests); FiniteTx0Host is the host form.
#pragma once

#include <array>
#include <cmath>
#include <future>
#include <initializer_list>
#include <memory>
#include <stdexcept>
#include <vector>

#include <DiFfRG/common/quadrature/quadrature_provider.hh>

#include "dfrg/integrate.hh"
#include "dfrg/math.hh"

namespace dfrg
{
  namespace tmpl
  {
    // KIND 0: (q, q0); 1: (q, cos, q0); 2: (q, cos, phi, q0) with d = 4
    template <int d, int KIND, typename NT, typename KERNEL> class FiniteTx0GPU
    {
      using Shape = FiniteTx0<d, KIND>;

    public:
      using ctype = ctype_t<NT>;

      FiniteTx0GPU(DiFfRG::QuadratureProvider &quadrature_provider, unsigned nx, unsigned na, unsigned n_x0,
                   ctype x_extent, ctype x0_extent, unsigned x0_summands, ctype T)
          : provider(quadrature_provider), total(n_x0 + x0_summands), summands(x0_summands)
      {
        nodes.xq = quadrature_provider.get_device_points<ctype>(nx);
        nodes.wq = quadrature_provider.get_device_weights<ctype>(nx);
        nodes.nx = (int)nx;
        nodes.x_extent = x_extent;
        nodes.q_ext = x0_extent;
        if constexpr (KIND >= 1) {
          nodes.na = (int)na;
          const auto &xa = quadrature_provider.get_points<ctype>(na);
          const auto &wa = quadrature_provider.get_weights<ctype>(na);
          ang = upload_angles(KIND == 2 ? angle_tables(xa, wa, {AngKind::cos, AngKind::phi})
                                        : angle_tables(xa, wa, {AngKind::cos}));
          nodes.ang = ang.get();
        }
        set_T(T);
      }

      void set_T(const ctype T)
      {
        const unsigned nm = is_close(T, ctype(0)) ? 0u : summands;
        nodes.m_T = T;
        nodes.nm = (int)nm;
        nodes.nb = (int)(total - nm);
        nodes.mx = provider.get_device_points<ctype>(total - nm);
        nodes.mw = provider.get_device_weights<ctype>(total - nm);
        state = next_node_state();
      }
      void set_x0_extent(const ctype val)
      {
        nodes.q_ext = val;
        state = next_node_state();
      }

      template <typename... T> NT get(const ctype k, const T &...t) const
      {
        return tmpl::get<Shape, NT, KERNEL>(nodes, k, t...);
      }
      template <typename... T> std::future<NT> request(const ctype k, const T &...t) const
      {
        return tmpl::request<Shape, NT, KERNEL>(*batcher, state, nodes, k, t...);
      }

    private:
      DiFfRG::QuadratureProvider &provider;
      unsigned total, summands;
      Nodes<ctype> nodes;
      std::shared_ptr<AngNode<ctype>> ang;
      unsigned long long state = 0;
      std::shared_ptr<RequestBatcher<NT>> batcher = std::make_shared<RequestBatcher<NT>>();
    };

    // The host form: radial nodes over OpenMP threads, partial sums added in node order.
    template <int d, int KIND, typename NT, typename KERNEL> class FiniteTx0Host
    {
    public:
      using ctype = ctype_t<NT>;

      FiniteTx0Host(DiFfRG::QuadratureProvider &quadrature_provider, unsigned nx, unsigned na, unsigned n_x0,
                    ctype x_extent, ctype x0_extent, unsigned x0_summands, ctype T)
          : provider(quadrature_provider), total(n_x0 + x0_summands), summands(x0_summands), x_extent(x_extent),
            x0_extent(x0_extent), xq(quadrature_provider.get_points<ctype>(nx)),
            wq(quadrature_provider.get_weights<ctype>(nx)), xa(quadrature_provider.get_points<ctype>(KIND ? na : nx)),
            wa(quadrature_provider.get_weights<ctype>(KIND ? na : nx))
      {
        set_T(T);
      }

      void set_T(const ctype T)
      {
        m_T = T;
        nm = is_close(T, ctype(0)) ? 0u : summands;
        x0 = &provider.get_points<ctype>(total - nm);
        w0 = &provider.get_weights<ctype>(total - nm);
      }
      void set_x0_extent(const ctype val) { x0_extent = val; }

      template <typename... T> NT get(const ctype k, const T &...t) const
      {
        using std::exp, std::log, std::sqrt;
        const ctype two_pi = 2 * (ctype)M_PI;
        const ctype z = m_T == 0 ? ctype(1e-3) : ctype(0);
        const ctype start = (2 * nm * (ctype)M_PI * m_T) / k;
        const ctype ls = log(start + z), le = log(x0_extent / (start + z));
        const int nx = (int)xq.size(), nc = KIND >= 1 ? (int)xa.size() : 1, nphi = KIND == 2 ? (int)xa.size() : 1;
        const int nint = (int)x0->size();
        std::vector<NT> part(nx, NT(0));
#pragma omp parallel for schedule(static)
        for (int ix = 0; ix < nx; ++ix) {
          const ctype q = k * sqrt(xq[ix] * x_extent);
          ctype el = powr<d - 3>(q) / (ctype)2 * powr<2>(k);
          if constexpr (KIND < 2) el *= S_d_prec<ctype>(d - 1);
          if constexpr (KIND == 1) el *= (ctype)1 / (ctype)2;
          const ctype wx = wq[ix] * x_extent;
          NT v = NT(0);
          for (int ic = 0; ic < nc; ++ic) {
            const ctype cos = 2 * (xa[ic] - (ctype)0.5), wc = KIND >= 1 ? 2 * wa[ic] : (ctype)1;
            for (int ip = 0; ip < nphi; ++ip) {
              const ctype phi = two_pi * xa[ip], wp = KIND == 2 ? two_pi * wa[ip] : (ctype)1;
              auto K = [&](ctype q0) {
                if constexpr (KIND == 0)
                  return KERNEL::kernel(q, q0, k, t...);
                else if constexpr (KIND == 1)
                  return KERNEL::kernel(q, cos, q0, k, t...);
                else
                  return KERNEL::kernel(q, cos, phi, q0, k, t...);
              };
              const ctype w = el * (wp * wc * wx);
              for (unsigned n = 0; n < nm; ++n) {
                const ctype f = w * (m_T / powr<d - 1>(two_pi));
                if (n == 0)
                  v += f * K(ctype(0));
                else {
                  const ctype q0 = two_pi * m_T * ctype(n);
                  v += f * (K(q0) + K(-q0));
                }
              }
              for (int j = 0; j < nint; ++j) {
                const ctype q0 = k * (exp(ls + le * (*x0)[j]) - z);
                const ctype wj = (*w0)[j] * le * q0 / k;
                v += (w * (k / powr<d>(two_pi)) * wj) * (K(q0) + K(-q0));
              }
            }
          }
          part[ix] = v;
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
      DiFfRG::QuadratureProvider &provider;
      unsigned total, summands, nm = 0;
      ctype x_extent, x0_extent, m_T = 0;
      const std::vector<ctype> &xq, &wq, &xa, &wa;
      const std::vector<ctype> *x0 = nullptr, *w0 = nullptr;
    };
  } // namespace tmpl
} // namespace dfrg
