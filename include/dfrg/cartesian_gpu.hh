// dfrg/cartesian_gpu.hh -- the GPU side of the Cartesian integrators on the tile engine of
// dfrg/integrate.hh: int d^Dq/(2 pi)^D KERNEL::kernel(qx[, qy[, qz]], k, t...) + constant over the
// box [q_min, q_max]^D, every axis a Gauss-Legendre rule on [0,1] mapped linearly (reference
// integrator_{1,2,3}D_cartesian_gpu.hh).  The first axis is the engine's radial digit (its range
// travels in the kernel arguments); qy and qz are device tables of (coordinate, weight).  A range
// change builds new tables (copies of the integrator keep theirs) and starts a new batch of
// requests; a batch holds the tables it was enqueued with until it is released.
#pragma once

#include <array>
#include <atomic>
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
    template <int D> struct CartesianShape;
    template <> struct CartesianShape<1> {
      using type = Cartesian1;
    };
    template <> struct CartesianShape<2> {
      using type = Cartesian2;
    };
    template <> struct CartesianShape<3> {
      using type = Cartesian3;
    };

    template <int D, typename NT, typename KERNEL> class CartesianGPU
    {
      static_assert(D >= 1 && D <= 3, "Cartesian integrators have one to three axes");
      using Shape = typename CartesianShape<D>::type;

    public:
      using ctype = ctype_t<NT>;

      CartesianGPU(DiFfRG::QuadratureProvider &quadrature_provider, const std::array<unsigned, D> &grid_sizes,
                   const std::array<ctype, D> &q_min, const std::array<ctype, D> &q_max)
      {
        for (int i = 0; i < D; ++i) {
          if (grid_sizes[i] < 1) throw std::invalid_argument("Cartesian integrator: an axis has no nodes");
          lo[i] = q_min[i];
          ext[i] = q_max[i] - q_min[i];
          if (i > 0) {
            x[i] = &quadrature_provider.get_points<ctype>(grid_sizes[i]);
            w[i] = &quadrature_provider.get_weights<ctype>(grid_sizes[i]);
          }
        }
        nodes.xq = quadrature_provider.get_device_points<ctype>(grid_sizes[0]);
        nodes.wq = quadrature_provider.get_device_weights<ctype>(grid_sizes[0]);
        nodes.nx = (int)grid_sizes[0];
        if constexpr (D > 1) nodes.na = (int)grid_sizes[1];
        if constexpr (D > 2) nodes.nb = (int)grid_sizes[2];
        refresh(true);
      }

      // set_q?_min keeps q_max (the extent changes with the lower end), set_q?_max keeps q_min
      // (integrator_2D_cartesian_gpu.hh:144-171)
      void set_min(int axis, ctype v)
      {
        ext[axis] = ext[axis] - v + lo[axis];
        lo[axis] = v;
        refresh(axis > 0);
      }
      void set_max(int axis, ctype v)
      {
        ext[axis] = v - lo[axis];
        refresh(axis > 0);
      }

      template <typename... T> NT get(const ctype k, const T &...t) const
      {
        return tmpl::get<Shape, NT, KERNEL>(nodes, k, t...);
      }
      template <typename... T> std::future<NT> request(const ctype k, const T &...t) const
      {
        return tmpl::request_holding<Shape, NT, KERNEL>(*batcher, state, tab, nodes, k, t...);
      }

    private:
      void refresh(bool tables)
      {
        nodes.q_min = lo[0];
        nodes.q_ext = ext[0];
        if (D > 1 && tables) {
          std::vector<AngNode<ctype>> t;
          for (int i = 1; i < D; ++i) {
            const auto one = cartesian_table(*x[i], *w[i], lo[i], ext[i]);
            t.insert(t.end(), one.begin(), one.end());
          }
          tab = upload_angles(t);
          nodes.ang = tab.get();
        }
        state = next_node_state();
      }

      std::array<ctype, D> lo{}, ext{};
      std::array<const std::vector<ctype> *, D> x{}, w{};
      Nodes<ctype> nodes;
      std::shared_ptr<AngNode<ctype>> tab;
      unsigned long long state = 0;
      std::shared_ptr<RequestBatcher<NT>> batcher = std::make_shared<RequestBatcher<NT>>();
    };
  } // namespace tmpl
} // namespace dfrg
