// DiFfRG/physics/integration/integrator_gpu.hh -- IntegratorGPU<d, NT, KERNEL>: radial d-dim
// momentum integral (reference integrator_gpu.hh:21-176) on the tile engine of dfrg/integrate.hh.
//   get(k, t...) = KERNEL::constant(k, t...) + sum_j S_d q_j^{d-2}/2 k^2/(2pi)^d w_j x_ext KERNEL::kernel(q_j, k, t...)
// with q_j = k sqrt(x_j x_ext) on the provider's Gauss-Legendre nodes x_j in [0,1].  get() is one
// launch; request()s are batched (one launch per batch, bitwise the values of get()).
#pragma once

#include <array>
#include <future>
#include <memory>

#include <DiFfRG/common/quadrature/quadrature_provider.hh>

#include "dfrg/integrate.hh"
#include "dfrg/math.hh"

namespace DiFfRG
{
  template <int d, typename NT, typename KERNEL> class IntegratorGPU
  {
    using Shape = dfrg::tmpl::Radial<d>;

  public:
    using ctype = dfrg::ctype_t<NT>;

    IntegratorGPU(QuadratureProvider &quadrature_provider, const std::array<unsigned, 1> grid_size,
                  const ctype x_extent, const unsigned max_block_size = 256)
        : IntegratorGPU(quadrature_provider, grid_size[0], x_extent, max_block_size)
    {
    }
    template <dfrg::JsonLike JSON>
    IntegratorGPU(QuadratureProvider &quadrature_provider, const std::array<unsigned, 1> grid_size,
                  const ctype x_extent, const JSON &json)
        : IntegratorGPU(quadrature_provider, grid_size[0], x_extent,
                        json.get_uint("/integration/cudathreadsperblock", 256))
    {
    }
    IntegratorGPU(QuadratureProvider &quadrature_provider, const unsigned grid_size, const ctype x_extent,
                  const unsigned max_block_size = 256)
    {
      (void)max_block_size; // the tile plan follows the node count
      nodes.xq = quadrature_provider.get_device_points<ctype>(grid_size);
      nodes.wq = quadrature_provider.get_device_weights<ctype>(grid_size);
      nodes.nx = (int)grid_size;
      nodes.x_extent = x_extent;
    }

    template <typename... T> NT get(const ctype k, const T &...t) const
    {
      return dfrg::tmpl::get<Shape, NT, KERNEL>(nodes, k, t...);
    }
    template <typename... T> std::future<NT> request(const ctype k, const T &...t) const
    {
      return dfrg::tmpl::request<Shape, NT, KERNEL>(*batcher, 0ULL, nodes, k, t...);
    }

  private:
    dfrg::tmpl::Nodes<ctype> nodes;
    std::shared_ptr<dfrg::tmpl::RequestBatcher<NT>> batcher = std::make_shared<dfrg::tmpl::RequestBatcher<NT>>();
  };
} // namespace DiFfRG
