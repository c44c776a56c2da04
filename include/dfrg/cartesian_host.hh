// dfrg/cartesian_host.hh -- the host side of the Cartesian integrators (reference
// integrator_{1,2,3}D_cartesian_cpu.hh): int d^Dq/(2 pi)^D KERNEL::kernel(qx[, qy[, qz]], k, t...)
// + constant over the box [q_min, q_max]^D, every axis a Gauss-Legendre rule on [0,1] mapped
// linearly (q = q_min + q_ext x, weight q_ext w).  The qx nodes are spread over OpenMP threads;
// their partial sums are added in node order.
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
    template <int D, typename NT, typename KERNEL> class CartesianHost
    {
      static_assert(D >= 1 && D <= 3, "Cartesian integrators have one to three axes");

    public:
      using ctype = ctype_t<NT>;

      CartesianHost(DiFfRG::QuadratureProvider &quadrature_provider, const std::array<unsigned, D> &grid_sizes,
                    const std::array<ctype, D> &q_min, const std::array<ctype, D> &q_max)
      {
        for (int i = 0; i < D; ++i) {
          if (grid_sizes[i] < 1) throw std::invalid_argument("Cartesian integrator: an axis has no nodes");
          lo[i] = q_min[i];
          ext[i] = q_max[i] - q_min[i];
          x[i] = &quadrature_provider.get_points<ctype>(grid_sizes[i]);
          w[i] = &quadrature_provider.get_weights<ctype>(grid_sizes[i]);
        }
      }

      // integrator_2D_cartesian_cpu.hh:100-125
      void set_min(int axis, ctype v)
      {
        ext[axis] = ext[axis] - v + lo[axis];
        lo[axis] = v;
      }
      void set_max(int axis, ctype v) { ext[axis] = v - lo[axis]; }

      template <typename... T> NT get(const ctype k, const T &...t) const
      {
        const ctype el = ctype(1) / powr<D>(2 * (ctype)M_PI);
        const int nx = (int)x[0]->size();
        std::vector<NT> part(nx, NT(0));
#pragma omp parallel for schedule(static)
        for (int ix = 0; ix < nx; ++ix) {
          const ctype qx = lo[0] + ext[0] * (*x[0])[ix], wx = ext[0] * (*w[0])[ix];
          NT v = NT(0);
          if constexpr (D == 1) {
            v += el * wx * KERNEL::kernel(qx, k, t...);
          } else {
            for (size_t iy = 0; iy < x[1]->size(); ++iy) {
              const ctype qy = lo[1] + ext[1] * (*x[1])[iy], wy = ext[1] * (*w[1])[iy];
              if constexpr (D == 2) {
                v += el * (wy * wx) * KERNEL::kernel(qx, qy, k, t...);
              } else {
                for (size_t iz = 0; iz < x[2]->size(); ++iz) {
                  const ctype qz = lo[2] + ext[2] * (*x[2])[iz], wz = ext[2] * (*w[2])[iz];
                  v += el * (wz * wy * wx) * KERNEL::kernel(qx, qy, qz, k, t...);
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
      std::array<ctype, D> lo{}, ext{};
      std::array<const std::vector<ctype> *, D> x{}, w{};
    };
  } // namespace tmpl
} // namespace dfrg
