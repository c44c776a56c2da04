// DiFfRG/physics/integration/integrator_angle_gpu.hh -- IntegratorAngleGPU<d, NT, KERNEL>
// (reference integrator_angle_gpu.hh:52-185) on the tile engine of dfrg/integrate.hh:
//   get(k, t...) = KERNEL::constant(k, t...) + S_d/(2pi)^d int dq q^{d-1} (1/2) int dcos KERNEL::kernel(q, cos, k, t...)
// with q^2 in [0, x_extent k^2].  get() is one launch; request()s are batched.
#pragma once

#include <array>
#include <future>
#include <memory>

#include <DiFfRG/common/quadrature/quadrature_provider.hh>

#include "dfrg/integrate.hh"
#include "dfrg/math.hh"

namespace DiFfRG
{
  template <int d, typename NT, typename KERNEL> class IntegratorAngleGPU
  {
    static_assert(d > 1, "IntegratorAngleGPU: d must be greater than 1, otherwise angles are not needed");
    using Shape = dfrg::tmpl::Angle1<d>;

  public:
    using ctype = dfrg::ctype_t<NT>;

    template <dfrg::JsonLike JSON>
    IntegratorAngleGPU(QuadratureProvider &quadrature_provider, const std::array<unsigned, 2> grid_sizes,
                       const ctype x_extent, const JSON &json)
        : IntegratorAngleGPU(quadrature_provider, grid_sizes, x_extent,
                             json.get_uint("/integration/cudathreadsperblock", 256))
    {
    }
    IntegratorAngleGPU(QuadratureProvider &quadrature_provider, std::array<unsigned, 2> grid_sizes,
                       const ctype x_extent, const unsigned max_block_size = 256)
    {
      (void)max_block_size;
      nodes.xq = quadrature_provider.get_device_points<ctype>(grid_sizes[0]);
      nodes.wq = quadrature_provider.get_device_weights<ctype>(grid_sizes[0]);
      nodes.nx = (int)grid_sizes[0];
      nodes.na = (int)grid_sizes[1];
      nodes.x_extent = x_extent;
      using dfrg::tmpl::AngKind;
      ang = dfrg::tmpl::upload_angles(dfrg::tmpl::angle_tables(quadrature_provider.get_points<ctype>(grid_sizes[1]),
                                                               quadrature_provider.get_weights<ctype>(grid_sizes[1]),
                                                               {AngKind::cos}));
      nodes.ang = ang.get();
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
    std::shared_ptr<dfrg::tmpl::AngNode<ctype>> ang;
    std::shared_ptr<dfrg::tmpl::RequestBatcher<NT>> batcher = std::make_shared<dfrg::tmpl::RequestBatcher<NT>>();
  };
} // namespace DiFfRG
