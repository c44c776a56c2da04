// DiFfRG/physics/integration/integrator_angle_gpu.hh -- IntegratorAngleGPU<d, NT, KERNEL>
// (reference integrator_angle_gpu.hh:52-185):
//   get(k, t...) = KERNEL::constant(k, t...) + S_d/(2pi)^d int dq q^{d-1} (1/2) int dcos KERNEL::kernel(q, cos, k, t...)
// with q^2 in [0, x_extent k^2].  One launch per get() and per batch of request()s.
#pragma once

#include <array>
#include <future>
#include <memory>
#include <tuple>

#include <DiFfRG/common/json.hh>
#include <DiFfRG/common/math.hh>
#include <DiFfRG/common/quadrature/quadrature_provider.hh>
#include <DiFfRG/common/types.hh>

#include "dfrg/integrate.hh"

namespace DiFfRG
{
  template <int d, typename NT, typename KERNEL> class IntegratorAngleGPU
  {
    static_assert(d > 1, "IntegratorAngleGPU: d must be greater than 1, otherwise angles are not needed");

  public:
    using ctype = typename get_type::ctype<NT>;

    IntegratorAngleGPU(QuadratureProvider &quadrature_provider, const std::array<uint, 2> grid_sizes,
                       const ctype x_extent, const JSONValue &json)
        : IntegratorAngleGPU(quadrature_provider, grid_sizes, x_extent,
                             json.get_uint("/integration/cudathreadsperblock", 256))
    {
    }
    IntegratorAngleGPU(QuadratureProvider &quadrature_provider, std::array<uint, 2> grid_sizes, const ctype x_extent,
                       const uint max_block_size = 256)
        : grid_sizes(grid_sizes), x_extent(x_extent)
    {
      (void)max_block_size; // launch shape is fixed by the node count
      ptr_x_quadrature_p = quadrature_provider.get_device_points<ctype>(grid_sizes[0]);
      ptr_x_quadrature_w = quadrature_provider.get_device_weights<ctype>(grid_sizes[0]);
      ptr_ang_quadrature_p = quadrature_provider.get_device_points<ctype>(grid_sizes[1]);
      ptr_ang_quadrature_w = quadrature_provider.get_device_weights<ctype>(grid_sizes[1]);
    }

    template <typename... T> NT get(const ctype k, const T &...t) const { return launch(k, t...).wait(); }

    // Deferred: requests are batched, and reading one future launches the whole batch
    // (dfrg::tmpl::RequestBatcher); results are bitwise those of get().
    template <typename... T> std::future<NT> request(const ctype k, const T &...t) const
    {
      using Args = std::tuple<ctype, T...>;
      return batcher->enqueue(0ULL, dfrg::tmpl::grid_for(nodes()), KERNEL::constant(k, t...), Args(k, t...), [this]() {
        return [xq = ptr_x_quadrature_p, wq = ptr_x_quadrature_w, xa = ptr_ang_quadrature_p,
                wa = ptr_ang_quadrature_w, nx = (int)grid_sizes[0], na = (int)grid_sizes[1],
                xe = x_extent](const Args *a, int r, int nb, NT *part, unsigned *cnt, NT *res, hipStream_t s) {
          hipLaunchKernelGGL((dfrg::tmpl::angle1_batch_kernel<d, ctype, NT, KERNEL, Args>), dim3(r * nb),
                             dim3(dfrg::tmpl::kBlock), 0, s, xq, wq, xa, wa, nx, na, xe, a, nb, part, cnt, res);
        };
      });
    }

  private:
    long long nodes() const { return (long long)grid_sizes[0] * grid_sizes[1]; }

    template <typename... T> dfrg::tmpl::Pending<NT> launch(const ctype k, const T &...t) const
    {
      auto p = dfrg::tmpl::Pending<NT>::make();
      p.constant = KERNEL::constant(k, t...);
      hipLaunchKernelGGL((dfrg::tmpl::angle1_kernel<d, ctype, NT, KERNEL, T...>), dim3(dfrg::tmpl::grid_for(nodes())),
                         dim3(dfrg::tmpl::kBlock), 0, p.stream, ptr_x_quadrature_p, ptr_x_quadrature_w,
                         ptr_ang_quadrature_p, ptr_ang_quadrature_w, (int)grid_sizes[0], (int)grid_sizes[1], x_extent,
                         k, p.partials(), p.counter(), p.result(), t...);
      p.enqueue_result();
      return p;
    }

    std::array<uint, 2> grid_sizes;
    ctype x_extent;
    const ctype *ptr_x_quadrature_p, *ptr_x_quadrature_w, *ptr_ang_quadrature_p, *ptr_ang_quadrature_w;
    std::shared_ptr<dfrg::tmpl::RequestBatcher<NT>> batcher = std::make_shared<dfrg::tmpl::RequestBatcher<NT>>();
  };
} // namespace DiFfRG
