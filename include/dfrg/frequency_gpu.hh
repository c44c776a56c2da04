// dfrg/frequency_gpu.hh -- the GPU side of the integrators over (q, cos[, phi][, q0]) at zero
// temperature on the tile engine of dfrg/integrate.hh: Integrator3DGPU (q, cos, phi) and the
// frequency forms Integrator{2,3}D{pq0,px0}GPU, whose q0 axis is a Gauss-Legendre rule on [0,1]
// mapped to [0, q0_extent] (pq0) or to x0 = q0/k in [0, x0_extent] (px0), summed as
// f(q0) + f(-q0).  Inner axes are device tables of (coordinate, weight): cos, then phi, then q0.
#pragma once

#include <array>
#include <future>
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
    // Shape S over nx radial nodes (x_extent), na angular nodes (cos; phi too when PHI) and nb
    // frequency nodes on [0, ext] (none when nb = 0).
    template <class S, bool PHI, typename NT, typename KERNEL> class FrequencyGPU
    {
    public:
      using ctype = ctype_t<NT>;

      FrequencyGPU(DiFfRG::QuadratureProvider &quadrature_provider, unsigned nx, unsigned na, unsigned nb,
                   ctype x_extent, ctype ext)
      {
        nodes.xq = quadrature_provider.get_device_points<ctype>(nx);
        nodes.wq = quadrature_provider.get_device_weights<ctype>(nx);
        nodes.nx = (int)nx;
        nodes.na = (int)na;
        nodes.nb = (int)nb;
        nodes.x_extent = x_extent;
        const auto &xa = quadrature_provider.get_points<ctype>(na);
        const auto &wa = quadrature_provider.get_weights<ctype>(na);
        auto t = PHI ? angle_tables(xa, wa, {AngKind::cos, AngKind::phi}) : angle_tables(xa, wa, {AngKind::cos});
        if (nb > 0) {
          const auto z = cartesian_table(quadrature_provider.get_points<ctype>(nb),
                                         quadrature_provider.get_weights<ctype>(nb), ctype(0), ext);
          t.insert(t.end(), z.begin(), z.end());
        }
        ang = upload_angles(t);
        nodes.ang = ang.get();
      }

      template <typename... T> NT get(const ctype k, const T &...t) const
      {
        return tmpl::get<S, NT, KERNEL>(nodes, k, t...);
      }
      template <typename... T> std::future<NT> request(const ctype k, const T &...t) const
      {
        return tmpl::request<S, NT, KERNEL>(*batcher, 0ULL, nodes, k, t...);
      }

    private:
      Nodes<ctype> nodes;
      std::shared_ptr<AngNode<ctype>> ang;
      std::shared_ptr<RequestBatcher<NT>> batcher = std::make_shared<RequestBatcher<NT>>();
    };

    inline void require_equal_angles(unsigned a, unsigned b)
    {
      if (a != b) throw std::runtime_error("Grid sizes must be currently equal for all angular dimensions!");
    }
  } // namespace tmpl
} // namespace dfrg
