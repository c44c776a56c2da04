// dfrg/frequency_host.hh -- the host side of the integrators over (q, cos[, phi][, q0]) at zero
// temperature (reference integrator_3D_cpu.hh, integrator_{2,3}D{pq0,px0}_cpu.hh): the same node
// sets and measures as the GPU forms of dfrg/frequency_gpu.hh.  Radial nodes are spread over
// OpenMP threads; their partial sums are added in node order.
#pragma once

#include <array>
#include <cmath>
#include <future>
#include <stdexcept>
#include <vector>

#include <DiFfRG/common/quadrature/quadrature_provider.hh>

#include "dfrg/math.hh"

namespace dfrg
{
  namespace tmpl
  {
    // DIM: 2 (q, cos, q0) or 3 (q, cos, phi[, q0]); Q0: a frequency axis of nb nodes on [0, ext];
    // PX0: that axis is x0 = q0/k (q0 = k x0, measure times k).
    template <int DIM, bool Q0, bool PX0, typename NT, typename KERNEL> class FrequencyHost
    {
      static_assert(DIM == 2 || DIM == 3, "two or three spatial dimensions");
      static_assert(Q0 || DIM == 3, "the 2D form has a frequency axis");

    public:
      using ctype = ctype_t<NT>;

      FrequencyHost(DiFfRG::QuadratureProvider &quadrature_provider, unsigned nx, unsigned na, unsigned nb,
                    ctype x_extent, ctype ext)
          : x_extent(x_extent), ext(ext), xq(quadrature_provider.get_points<ctype>(nx)),
            wq(quadrature_provider.get_weights<ctype>(nx)), xa(quadrature_provider.get_points<ctype>(na)),
            wa(quadrature_provider.get_weights<ctype>(na)),
            xb(quadrature_provider.get_points<ctype>(Q0 ? nb : na)),
            wb(quadrature_provider.get_weights<ctype>(Q0 ? nb : na)) // unused without a frequency axis
      {
      }

      template <typename... T> NT get(const ctype k, const T &...t) const
      {
        const ctype two_pi = 2 * (ctype)M_PI;
        const int nx = (int)xq.size(), na = (int)xa.size(), nphi = DIM == 3 ? na : 1, nz = Q0 ? (int)xb.size() : 1;
        std::vector<NT> part(nx, NT(0));
#pragma omp parallel for schedule(static)
        for (int ix = 0; ix < nx; ++ix) {
          const ctype q = k * std::sqrt(xq[ix] * x_extent);
          // measure of the radial node: x = q^2/k^2, then S_d-type and Fourier factors
          ctype el;
          if constexpr (DIM == 2)
            el = (powr<2>(k) / (ctype)2) * two_pi * ((ctype)1 / (ctype)2) / powr<3>(two_pi);
          else
            el = (q / (ctype)2 * powr<2>(k)) / powr<Q0 ? 4 : 3>(two_pi);
          if constexpr (PX0) el *= k;
          const ctype wx = wq[ix] * x_extent;
          NT v = NT(0);
          for (int ic = 0; ic < na; ++ic) {
            const ctype cos = 2 * (xa[ic] - (ctype)0.5), wc = 2 * wa[ic];
            for (int ip = 0; ip < nphi; ++ip) {
              const ctype phi = two_pi * xa[ip], wp = DIM == 3 ? two_pi * wa[ip] : (ctype)1;
              for (int iz = 0; iz < nz; ++iz) {
                if constexpr (Q0) {
                  const ctype q0 = (PX0 ? k : (ctype)1) * (xb[iz] * ext), wz = wb[iz] * ext;
                  const ctype w = el * (wz * wp * wc * wx);
                  if constexpr (DIM == 2)
                    v += w * (KERNEL::kernel(q, cos, q0, k, t...) + KERNEL::kernel(q, cos, -q0, k, t...));
                  else
                    v += w * (KERNEL::kernel(q, cos, phi, q0, k, t...) + KERNEL::kernel(q, cos, phi, -q0, k, t...));
                } else {
                  v += el * (wp * wc * wx) * KERNEL::kernel(q, cos, phi, k, t...);
                }
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
      ctype x_extent, ext;
      const std::vector<ctype> &xq, &wq, &xa, &wa, &xb, &wb;
    };
  } // namespace tmpl
} // namespace dfrg
