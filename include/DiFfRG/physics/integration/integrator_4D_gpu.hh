// DiFfRG/physics/integration/integrator_4D_gpu.hh -- Integrator4DGPU<NT, KERNEL>
// (reference integrator_4D_gpu.hh:62-352) on the tile engine of dfrg/integrate.hh:
//   int dq dcos1 dcos2 dphi (2pi)^-4 q^2/2 k^2 sqrt(1 - cos1^2) KERNEL::kernel(q, cos1, cos2, phi, k, t...) + constant
// with q^2 in [0, x_extent k^2].  get() is one launch; request()s are batched.  The three angular
// orders must be equal.
#pragma once

#include <array>
#include <future>
#include <memory>
#include <stdexcept>

#include <DiFfRG/common/quadrature/quadrature_provider.hh>

#include "dfrg/integrate.hh"
#include "dfrg/math.hh"

namespace DiFfRG
{
  template <typename NT, typename KERNEL> class Integrator4DGPU
  {
    using Shape = dfrg::tmpl::Angle3;

  public:
    using ctype = dfrg::ctype_t<NT>;

    template <dfrg::JsonLike JSON>
    Integrator4DGPU(QuadratureProvider &quadrature_provider, const std::array<unsigned, 4> grid_sizes,
                    const ctype x_extent, const JSON &json)
        : Integrator4DGPU(quadrature_provider, grid_sizes, x_extent,
                          json.get_uint("/integration/cudathreadsperblock", 256))
    {
    }
    Integrator4DGPU(QuadratureProvider &quadrature_provider, std::array<unsigned, 4> grid_sizes, const ctype x_extent,
                    const unsigned max_block_size = 256)
    {
      (void)max_block_size;
      if (grid_sizes[2] != grid_sizes[1] || grid_sizes[3] != grid_sizes[1])
        throw std::runtime_error("Grid sizes must be currently equal for all angular dimensions!");
      nodes.xq = quadrature_provider.get_device_points<ctype>(grid_sizes[0]);
      nodes.wq = quadrature_provider.get_device_weights<ctype>(grid_sizes[0]);
      nodes.nx = (int)grid_sizes[0];
      nodes.na = (int)grid_sizes[1];
      nodes.x_extent = x_extent;
      using dfrg::tmpl::AngKind;
      // digits (q, cos1, phi, cos2), cos2 fastest (integrator_4D_gpu.hh:20-49)
      ang = dfrg::tmpl::upload_angles(dfrg::tmpl::angle_tables(quadrature_provider.get_points<ctype>(grid_sizes[1]),
                                                               quadrature_provider.get_weights<ctype>(grid_sizes[1]),
                                                               {AngKind::cos_jacobian, AngKind::phi, AngKind::cos}));
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
