// DiFfRG/physics/integration_finiteT/integrator_4D_finiteTq0_gpu.hh -- Integrator4DFiniteTq0GPU<NT,
// KERNEL> (reference integrator_4D_finiteTq0_gpu.hh:57-232): three spatial dimensions in (q, cos,
// phi) and the Matsubara sum in q0,
//   int dq dcos dphi q^2/(2pi)^3 [ sum_i w_i (f(q,cos,phi,x_i) + f(q,cos,phi,-x_i)) + T f(q,cos,phi,0) ]
// with KERNEL::kernel(q, cos, phi, q0, k, t...).  Equal angular orders, as in the reference; the
// rule follows E (dfrg::tmpl::MatsubaraState).  One launch per get() and per batch of request()s.
#pragma once

#include <array>
#include <future>
#include <memory>
#include <stdexcept>
#include <tuple>

#include <DiFfRG/common/json.hh>
#include <DiFfRG/common/math.hh>
#include <DiFfRG/common/quadrature/quadrature_provider.hh>
#include <DiFfRG/common/types.hh>

#include "dfrg/integrate.hh"
#include "dfrg/matsubara_state.hh"

namespace DiFfRG
{
  template <typename NT, typename KERNEL> class Integrator4DFiniteTq0GPU
  {
  public:
    using ctype = typename get_type::ctype<NT>;

    Integrator4DFiniteTq0GPU(QuadratureProvider &quadrature_provider, const std::array<uint, 3> grid_sizes,
                             const ctype x_extent, const JSONValue &json)
        : Integrator4DFiniteTq0GPU(quadrature_provider, grid_sizes, x_extent, json.get_double("/physical/T"),
                                   json.get_uint("/integration/cudathreadsperblock", 256))
    {
    }
    Integrator4DFiniteTq0GPU(QuadratureProvider &quadrature_provider, const std::array<uint, 3> grid_sizes,
                             const ctype x_extent, const ctype T, const uint max_block_size = 256)
        : grid_sizes(grid_sizes), x_extent(x_extent)
    {
      (void)max_block_size;
      if (grid_sizes[2] != grid_sizes[1])
        throw std::runtime_error("Grid sizes must be currently equal for all angular dimensions!");
      ptr_x_quadrature_p = quadrature_provider.get_device_points<ctype>(grid_sizes[0]);
      ptr_x_quadrature_w = quadrature_provider.get_device_weights<ctype>(grid_sizes[0]);
      ptr_ang_quadrature_p = quadrature_provider.get_device_points<ctype>(grid_sizes[1]);
      ptr_ang_quadrature_w = quadrature_provider.get_device_weights<ctype>(grid_sizes[1]);
      mats.provider = &quadrature_provider;
      mats.set_T(T);
    }

    void set_T(const ctype T, const ctype E = 0) { mats.set_T(T, E); }
    void set_E(const ctype E) { mats.set_E(E); }
    uint matsubara_size() const { return mats.size; }

    template <typename... T> NT get(const ctype k, const T &...t)
    {
      mats.track(k);
      auto p = dfrg::tmpl::Pending<NT>::make();
      p.constant = KERNEL::constant(k, t...);
      hipLaunchKernelGGL((dfrg::tmpl::angle2_finiteT_kernel<ctype, NT, KERNEL, T...>),
                         dim3(dfrg::tmpl::grid_for(nodes())), dim3(dfrg::tmpl::kBlock), 0, p.stream,
                         ptr_x_quadrature_p, ptr_x_quadrature_w, ptr_ang_quadrature_p, ptr_ang_quadrature_w, mats.x,
                         mats.w, (int)grid_sizes[0], (int)grid_sizes[1], (int)mats.size, x_extent, mats.T, k,
                         p.partials(), p.counter(), p.result(), t...);
      p.enqueue_result();
      return p.wait();
    }

    // Deferred and batched (dfrg::tmpl::RequestBatcher); a new Matsubara rule starts a new batch.
    template <typename... T> std::future<NT> request(const ctype k, const T &...t)
    {
      mats.track(k);
      using Args = std::tuple<ctype, T...>;
      const unsigned long long state = (unsigned long long)(uintptr_t)mats.x;
      return batcher->enqueue(state, dfrg::tmpl::grid_for(nodes()), KERNEL::constant(k, t...), Args(k, t...), [this]() {
        return [xq = ptr_x_quadrature_p, wq = ptr_x_quadrature_w, xa = ptr_ang_quadrature_p,
                wa = ptr_ang_quadrature_w, mx = mats.x, mw = mats.w, nx = (int)grid_sizes[0], na = (int)grid_sizes[1],
                nm = (int)mats.size, xe = x_extent,
                temp = mats.T](const Args *a, int r, int nb, NT *part, unsigned *cnt, NT *res, hipStream_t s) {
          hipLaunchKernelGGL((dfrg::tmpl::angle2_finiteT_batch_kernel<ctype, NT, KERNEL, Args>), dim3(r * nb),
                             dim3(dfrg::tmpl::kBlock), 0, s, xq, wq, xa, wa, mx, mw, nx, na, nm, xe, temp, a, nb, part,
                             cnt, res);
        };
      });
    }

  private:
    long long nodes() const
    {
      return (long long)grid_sizes[0] * grid_sizes[1] * grid_sizes[1] * mats.size;
    }

    std::array<uint, 3> grid_sizes;
    ctype x_extent;
    const ctype *ptr_x_quadrature_p, *ptr_x_quadrature_w, *ptr_ang_quadrature_p, *ptr_ang_quadrature_w;
    dfrg::tmpl::MatsubaraState<ctype, true> mats;
    std::shared_ptr<dfrg::tmpl::RequestBatcher<NT>> batcher = std::make_shared<dfrg::tmpl::RequestBatcher<NT>>();
  };
} // namespace DiFfRG
