// DiFfRG/physics/integration_finiteT/integrator_4D_finiteTq0_gpu.hh -- Integrator4DFiniteTq0GPU<NT,
// KERNEL> (reference integrator_4D_finiteTq0_gpu.hh:57-232): three spatial dimensions in (q, cos,
// phi) and the Matsubara sum in q0,
//   int dq dcos dphi q^2/(2pi)^3 [ sum_i w_i (f(q,cos,phi,x_i) + f(q,cos,phi,-x_i)) + T f(q,cos,phi,0) ]
// with KERNEL::kernel(q, cos, phi, q0, k, t...), on the tile engine of dfrg/integrate.hh.  Equal
// angular orders, as in the reference; the rule follows E (dfrg::tmpl::MatsubaraState).
#pragma once

#include <array>
#include <cstdint>
#include <future>
#include <memory>
#include <stdexcept>

#include <DiFfRG/common/quadrature/quadrature_provider.hh>

#include "dfrg/integrate.hh"
#include "dfrg/math.hh"
#include "dfrg/matsubara_state.hh"

namespace DiFfRG
{
  template <typename NT, typename KERNEL> class Integrator4DFiniteTq0GPU
  {
    using Shape = dfrg::tmpl::Angle2FiniteT;

  public:
    using ctype = dfrg::ctype_t<NT>;

    template <dfrg::JsonLike JSON>
    Integrator4DFiniteTq0GPU(QuadratureProvider &quadrature_provider, const std::array<unsigned, 3> grid_sizes,
                             const ctype x_extent, const JSON &json)
        : Integrator4DFiniteTq0GPU(quadrature_provider, grid_sizes, x_extent, json.get_double("/physical/T"),
                                   json.get_uint("/integration/cudathreadsperblock", 256))
    {
    }
    Integrator4DFiniteTq0GPU(QuadratureProvider &quadrature_provider, const std::array<unsigned, 3> grid_sizes,
                             const ctype x_extent, const ctype T, const unsigned max_block_size = 256)
    {
      (void)max_block_size;
      if (grid_sizes[2] != grid_sizes[1])
        throw std::runtime_error("Grid sizes must be currently equal for all angular dimensions!");
      base.xq = quadrature_provider.get_device_points<ctype>(grid_sizes[0]);
      base.wq = quadrature_provider.get_device_weights<ctype>(grid_sizes[0]);
      base.nx = (int)grid_sizes[0];
      base.na = (int)grid_sizes[1];
      base.x_extent = x_extent;
      using dfrg::tmpl::AngKind;
      // digits (q, cos, phi, n) (integrator_4D_finiteTq0_gpu.hh:20-55)
      ang = dfrg::tmpl::upload_angles(dfrg::tmpl::angle_tables(quadrature_provider.get_points<ctype>(grid_sizes[1]),
                                                               quadrature_provider.get_weights<ctype>(grid_sizes[1]),
                                                               {AngKind::cos, AngKind::phi}));
      base.ang = ang.get();
      mats.provider = &quadrature_provider;
      mats.set_T(T);
    }

    void set_T(const ctype T, const ctype E = 0) { mats.set_T(T, E); }
    void set_E(const ctype E) { mats.set_E(E); }
    unsigned matsubara_size() const { return mats.size; }

    template <typename... T> NT get(const ctype k, const T &...t)
    {
      mats.track(k);
      return dfrg::tmpl::get<Shape, NT, KERNEL>(nodes(), k, t...);
    }
    template <typename... T> std::future<NT> request(const ctype k, const T &...t)
    {
      mats.track(k);
      return dfrg::tmpl::request<Shape, NT, KERNEL>(*batcher, (unsigned long long)(uintptr_t)mats.x, nodes(), k,
                                                    t...);
    }

  private:
    dfrg::tmpl::Nodes<ctype> nodes() const
    {
      dfrg::tmpl::Nodes<ctype> n = base;
      n.mx = mats.x;
      n.mw = mats.w;
      n.nm = (int)mats.size;
      n.m_T = mats.T;
      return n;
    }

    dfrg::tmpl::Nodes<ctype> base;
    std::shared_ptr<dfrg::tmpl::AngNode<ctype>> ang;
    dfrg::tmpl::MatsubaraState<ctype, true> mats;
    std::shared_ptr<dfrg::tmpl::RequestBatcher<NT>> batcher = std::make_shared<dfrg::tmpl::RequestBatcher<NT>>();
  };
} // namespace DiFfRG
