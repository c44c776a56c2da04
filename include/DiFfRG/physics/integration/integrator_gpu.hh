// DiFfRG/physics/integration/integrator_gpu.hh -- IntegratorGPU<d, NT, KERNEL>: radial d-dim
// momentum integral (reference integrator_gpu.hh:21-176): one HIP launch per get(), one per batch
// of request()s.
//   get(k, t...) = KERNEL::constant(k, t...) + sum_j S_d q_j^{d-2}/2 k^2/(2pi)^d w_j x_ext KERNEL::kernel(q_j, k, t...)
// with q_j = k sqrt(x_j x_ext) on the provider's Gauss-Legendre nodes x_j in [0,1].
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
  template <int d, typename NT, typename KERNEL> class IntegratorGPU
  {
  public:
    using ctype = typename get_type::ctype<NT>;

    IntegratorGPU(QuadratureProvider &quadrature_provider, const std::array<uint, 1> grid_size, const ctype x_extent,
                  const uint max_block_size = 256)
        : IntegratorGPU(quadrature_provider, grid_size[0], x_extent, max_block_size)
    {
    }
    IntegratorGPU(QuadratureProvider &quadrature_provider, const std::array<uint, 1> grid_size, const ctype x_extent,
                  const JSONValue &json)
        : IntegratorGPU(quadrature_provider, grid_size[0], x_extent,
                        json.get_uint("/integration/cudathreadsperblock", 256))
    {
    }
    IntegratorGPU(QuadratureProvider &quadrature_provider, const uint grid_size, const ctype x_extent,
                  const uint max_block_size = 256)
        : grid_size(grid_size), x_extent(x_extent)
    {
      (void)max_block_size; // launch shape is fixed by the node count
      ptr_x_quadrature_p = quadrature_provider.get_device_points<ctype>(grid_size);
      ptr_x_quadrature_w = quadrature_provider.get_device_weights<ctype>(grid_size);
    }

    template <typename... T> NT get(const ctype k, const T &...t) const { return launch(k, t...).wait(); }

    // Deferred: requests are batched, and reading one future launches the whole batch
    // (dfrg::tmpl::RequestBatcher); results are bitwise those of get().
    template <typename... T> std::future<NT> request(const ctype k, const T &...t) const
    {
      using Args = std::tuple<ctype, T...>;
      const int nb = dfrg::tmpl::grid_for(grid_size);
      return batcher->enqueue(0ULL, nb, KERNEL::constant(k, t...), Args(k, t...), [this]() {
        return [xq = ptr_x_quadrature_p, wq = ptr_x_quadrature_w, n = (int)grid_size, xe = x_extent](
                   const Args *a, int r, int nb, NT *part, unsigned *cnt, NT *res, hipStream_t s) {
          hipLaunchKernelGGL((dfrg::tmpl::radial_batch_kernel<d, ctype, NT, KERNEL, Args>), dim3(r * nb),
                             dim3(dfrg::tmpl::kBlock), 0, s, xq, wq, n, xe, a, nb, part, cnt, res);
        };
      });
    }

  private:
    template <typename... T> dfrg::tmpl::Pending<NT> launch(const ctype k, const T &...t) const
    {
      auto p = dfrg::tmpl::Pending<NT>::make();
      p.constant = KERNEL::constant(k, t...);
      const int grid = dfrg::tmpl::grid_for(grid_size);
      hipLaunchKernelGGL((dfrg::tmpl::radial_kernel<d, ctype, NT, KERNEL, T...>), dim3(grid),
                         dim3(dfrg::tmpl::kBlock), 0, p.stream, ptr_x_quadrature_p, ptr_x_quadrature_w,
                         (int)grid_size, x_extent, k, p.partials(), p.counter(), p.result(), t...);
      p.enqueue_result();
      return p;
    }

    uint grid_size;
    ctype x_extent;
    const ctype *ptr_x_quadrature_p;
    const ctype *ptr_x_quadrature_w;
    std::shared_ptr<dfrg::tmpl::RequestBatcher<NT>> batcher = std::make_shared<dfrg::tmpl::RequestBatcher<NT>>();
  };
} // namespace DiFfRG
