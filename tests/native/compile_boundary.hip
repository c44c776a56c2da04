// compile_boundary.hip -- compile-only checks of the drop-in boundary (built with -c, never
// linked or run; tests/test_boundary_compile.py drives it).
//
// 1. Every GPU and host integrator, the QuadratureProvider and request_data are instantiated with
//    a parameter object that has exactly the getters of the reference's JSONValue
//    (DiFfRG/include/DiFfRG/common/json.hh:62-139) and nothing else, declared only: the JSON
//    constructors bind to the reference's own JSONValue, whatever backs it.
// 2. With DFRG_REFERENCE_KERNELS, the reference's generated kernels are included unchanged by
//    include path -- Examples/ONfiniteT/flows/V/V.kernel and Examples/YangMills/SP/flows/AA/ZA.kernel
//    -- and instantiated in the integrators their wrappers use (V.cu: IntegratorTBB<3>; ZA.cu:
//    IntegratorAngleGPU<4>), with the argument types of those wrappers.
#include <string>

#include <DiFfRG/common/math.hh>
#include <DiFfRG/discretization/grid/coordinates.hh>
#include <DiFfRG/physics/regulators.hh>
#include <DiFfRG/physics/thermodynamics.hh>

#include <DiFfRG/common/quadrature/quadrature_provider.hh>
#include <DiFfRG/physics/integration/integrator_1D_cartesian_cpu.hh>
#include <DiFfRG/physics/integration/integrator_1D_cartesian_gpu.hh>
#include <DiFfRG/physics/integration/integrator_2D_cartesian_cpu.hh>
#include <DiFfRG/physics/integration/integrator_2D_cartesian_gpu.hh>
#include <DiFfRG/physics/integration/integrator_3D_cartesian_cpu.hh>
#include <DiFfRG/physics/integration/integrator_3D_cartesian_gpu.hh>
#include <DiFfRG/physics/integration/integrator_2Dpq0_cpu.hh>
#include <DiFfRG/physics/integration/integrator_2Dpq0_gpu.hh>
#include <DiFfRG/physics/integration/integrator_2Dpx0_cpu.hh>
#include <DiFfRG/physics/integration/integrator_2Dpx0_gpu.hh>
#include <DiFfRG/physics/integration/integrator_3D_cpu.hh>
#include <DiFfRG/physics/integration/integrator_3D_gpu.hh>
#include <DiFfRG/physics/integration/integrator_3Dpq0_cpu.hh>
#include <DiFfRG/physics/integration/integrator_3Dpq0_gpu.hh>
#include <DiFfRG/physics/integration/integrator_3Dpx0_cpu.hh>
#include <DiFfRG/physics/integration/integrator_3Dpx0_gpu.hh>
#include <DiFfRG/physics/integration/integrator_4D_2ang_cpu.hh>
#include <DiFfRG/physics/integration/integrator_4D_2ang_gpu.hh>
#include <DiFfRG/physics/integration/integrator_4D_cpu.hh>
#include <DiFfRG/physics/integration/integrator_4D_gpu.hh>
#include <DiFfRG/physics/integration/integrator_angle_cpu.hh>
#include <DiFfRG/physics/integration/integrator_angle_gpu.hh>
#include <DiFfRG/physics/integration/integrator_constant.hh>
#include <DiFfRG/physics/integration/integrator_cpu.hh>
#include <DiFfRG/physics/integration/integrator_gpu.hh>
#include <DiFfRG/physics/integration_finiteT/integrator_4D_finiteTq0_cpu.hh>
#include <DiFfRG/physics/integration_finiteT/integrator_4D_finiteTx0_cpu.hh>
#include <DiFfRG/physics/integration_finiteT/integrator_4D_finiteTx0_gpu.hh>
#include <DiFfRG/physics/integration_finiteT/integrator_angle_finiteTx0_cpu.hh>
#include <DiFfRG/physics/integration_finiteT/integrator_angle_finiteTx0_gpu.hh>
#include <DiFfRG/physics/integration_finiteT/integrator_finiteTx0_cpu.hh>
#include <DiFfRG/physics/integration_finiteT/integrator_finiteTx0_gpu.hh>
#include <DiFfRG/physics/integration_finiteT/integrator_4D_finiteTq0_gpu.hh>
#include <DiFfRG/physics/integration_finiteT/integrator_angle_finiteTq0_cpu.hh>
#include <DiFfRG/physics/integration_finiteT/integrator_angle_finiteTq0_gpu.hh>
#include <DiFfRG/physics/integration_finiteT/integrator_finiteTq0_cpu.hh>
#include <DiFfRG/physics/integration_finiteT/integrator_finiteTq0_gpu.hh>
#include <DiFfRG/physics/interpolation/linear_interpolation_1D.hh>
#include <DiFfRG/physics/interpolation/tex_linear_interpolation_1D.hh>
#include <DiFfRG/physics/utils.hh>

using namespace ::DiFfRG;

// the getters of the reference's JSONValue, declarations only
struct ReferenceJSONValue {
  double get_double(const std::string &key) const;
  int get_int(const std::string &key) const;
  unsigned get_uint(const std::string &key) const;
  std::string get_string(const std::string &key) const;
  bool get_bool(const std::string &key) const;
  double get_double(const std::string &key, const double def) const;
  int get_int(const std::string &key, const int def) const;
  unsigned get_uint(const std::string &key, const unsigned def) const;
  std::string get_string(const std::string &key, const std::string &def) const;
  bool get_bool(const std::string &key, const bool def) const;
};

#ifdef DFRG_REFERENCE_KERNELS
#include "V.kernel"
#include "ZA.kernel"
#endif

namespace
{
  struct K1 {
    static __host__ __device__ double kernel(double q, double k, double a) { return a * q * k; }
    static __host__ __device__ double constant(double, double) { return 0.; }
  };
  struct K1p {
    static __host__ __device__ double kernel(double q, double k, double p, double a) { return a * q * k / (p + q); }
    static __host__ __device__ double constant(double, double, double) { return 0.; }
  };
  struct KAngle {
    static __host__ __device__ double kernel(double q, double c, double k, double a) { return a * q * c * k; }
    static __host__ __device__ double constant(double, double) { return 0.; }
  };
  struct K2Ang {
    static __host__ __device__ double kernel(double q, double c1, double c2, double k, double a) { return a * q * c1 * c2 * k; }
    static __host__ __device__ double constant(double, double) { return 0.; }
  };
  struct K4D {
    static __host__ __device__ double kernel(double q, double c1, double c2, double ph, double k, double a)
    {
      return a * q * c1 * c2 * ph * k;
    }
    static __host__ __device__ double constant(double, double) { return 0.; }
  };
  struct KT {
    static __host__ __device__ double kernel(double q, double q0, double k, double a) { return a / (q * q + q0 * q0 + k); }
    static __host__ __device__ double constant(double, double) { return 0.; }
  };
  struct KAngleT {
    static __host__ __device__ double kernel(double q, double c, double q0, double k, double a)
    {
      return a * c / (q * q + q0 * q0 + k);
    }
    static __host__ __device__ double constant(double, double) { return 0.; }
  };
  struct K4DT {
    static __host__ __device__ double kernel(double q, double c, double ph, double q0, double k, double a)
    {
      return a * c * ph / (q * q + q0 * q0 + k);
    }
    static __host__ __device__ double constant(double, double) { return 0.; }
  };

  struct KC0 {
    static __host__ __device__ double kernel(double k, double a) { return a * k; }
    static __host__ __device__ double constant(double, double) { return 0.; }
  };
  struct KC1 {
    static __host__ __device__ double kernel(double qx, double k, double a) { return a * qx * k; }
    static __host__ __device__ double constant(double, double) { return 0.; }
  };
  struct KC2 {
    static __host__ __device__ double kernel(double qx, double qy, double k, double a) { return a * qx * qy * k; }
    static __host__ __device__ double constant(double, double) { return 0.; }
  };
  struct KC3 {
    static __host__ __device__ double kernel(double qx, double qy, double qz, double k, double a)
    {
      return a * qx * qy * qz * k;
    }
    static __host__ __device__ double constant(double, double) { return 0.; }
  };

  struct K3D {
    static __host__ __device__ double kernel(double q, double c, double ph, double k, double a) { return a * q * c * ph * k; }
    static __host__ __device__ double constant(double, double) { return 0.; }
  };
  struct KQ0 {
    static __host__ __device__ double kernel(double q, double c, double q0, double k, double a) { return a * q * c / (q0 * q0 + k); }
    static __host__ __device__ double constant(double, double) { return 0.; }
  };
  struct K3Q0 {
    static __host__ __device__ double kernel(double q, double c, double ph, double q0, double k, double a)
    {
      return a * q * c * ph / (q0 * q0 + k);
    }
    static __host__ __device__ double constant(double, double) { return 0.; }
  };

  // a wrapper of the shape the generated X_integrator classes have (templated request<NT>)
  template <class I> struct Wrapper {
    I &integ;
    template <typename NT, typename... T> std::future<NT> request(T &&...t) { return integ.request(t...); }
  };
} // namespace

double instantiate_with_reference_json(const ReferenceJSONValue &json)
{
  QuadratureProvider qp(json);
  double s = 0.;
  IntegratorGPU<4, double, K1> a(qp, {{64}}, 20., json);
  IntegratorAngleGPU<4, double, KAngle> b(qp, {{64, 16}}, 20., json);
  Integrator4D2AngGPU<double, K2Ang> c(qp, {{64, 32, 32}}, 20., json);
  Integrator4DGPU<double, K4D> d(qp, {{32, 8, 8, 8}}, 20., json);
  IntegratorFiniteTq0GPU<4, double, KT> e(qp, {{32}}, 20., json);
  IntegratorAngleFiniteTq0GPU<4, double, KAngleT> f(qp, {{32, 8}}, 20., json);
  Integrator4DFiniteTq0GPU<double, K4DT> g(qp, {{32, 8, 8}}, 20., json);
  s += a.get(1., 2.) + b.get(1., 2.) + c.get(1., 2.) + d.get(1., 2.) + e.get(1., 2.) + f.get(1., 2.) + g.get(1., 2.);
  s += a.request(1., 2.).get() + c.request(1., 2.).get() + e.request(1., 2.).get();
  IntegratorTBB<4, double, K1> at(qp, {{64}}, 20., json);
  IntegratorAngleTBB<4, double, KAngle> bt(qp, {{64, 16}}, 20., json);
  Integrator4D2AngTBB<double, K2Ang> ct(qp, {{64, 32, 32}}, 20., json);
  Integrator4DTBB<double, K4D> dt(qp, {{32, 8, 8, 8}}, 20., json);
  IntegratorFiniteTq0TBB<4, double, KT> et(qp, {{32}}, 20., json);
  s += at.get(1., 2.) + bt.get(1., 2.) + ct.get(1., 2.) + dt.get(1., 2.) + et.get(1., 2.);
  // Cartesian forms: JSON constructors (ranges from /discretization/integration/q?_{min,max}) and
  // the range mutators
  Integrator1DCartesianGPU<double, KC1> h1(qp, {{32}}, json);
  Integrator1DCartesianGPU<double, KC1> h1x(qp, {{32}}, 0., json);
  Integrator2DCartesianGPU<double, KC2> h2(qp, {{32, 16}}, 0., json);
  Integrator3DCartesianGPU<double, KC3> h3(qp, {{16, 16, 8}}, 0., json);
  Integrator1DCartesianTBB<double, KC1> h1t(qp, {{32}}, json);
  Integrator2DCartesianTBB<double, KC2> h2t(qp, {{32, 16}}, 0., json);
  Integrator3DCartesianTBB<double, KC3> h3t(qp, {{16, 16, 8}}, 0., json);
  h1.set_qx_min(-1.);
  h2.set_qy_max(1.);
  h3.set_qz_min(-2.);
  h3t.set_qz_max(2.);
  s += h1.get(1., 2.) + h1x.request(1., 2.).get() + h2.get(1., 2.) + h3.request(1., 2.).get();
  s += h1t.get(1., 2.) + h2t.request(1., 2.).get() + h3t.get(1., 2.);
  // zero-temperature (q, cos[, phi][, q0]) forms with JSON constructors
  Integrator3DGPU<double, K3D> f3(qp, {{32, 8, 8}}, 20., json);
  Integrator2Dpq0GPU<double, KQ0> f2q(qp, {{32, 8, 8}}, 20., 2., json);
  Integrator2Dpx0GPU<double, KQ0> f2x(qp, {{32, 8, 8}}, 20., 2., json);
  Integrator3Dpq0GPU<double, K3Q0> f3q(qp, {{32, 8, 8, 8}}, 20., 2., json);
  Integrator3Dpx0GPU<double, K3Q0> f3x(qp, {{32, 8, 8, 8}}, 20., 2., json);
  Integrator3DTBB<double, K3D> f3t(qp, {{32, 8, 8}}, 20., json);
  Integrator2Dpq0TBB<double, KQ0> f2qt(qp, {{32, 8, 8}}, 20., 2., json);
  Integrator2Dpx0TBB<double, KQ0> f2xt(qp, {{32, 8, 8}}, 20., 2., json);
  Integrator3Dpq0TBB<double, K3Q0> f3qt(qp, {{32, 8, 8, 8}}, 20., 2., json);
  Integrator3Dpx0TBB<double, K3Q0> f3xt(qp, {{32, 8, 8, 8}}, 20., 2., json);
  s += f3.get(1., 2.) + f2q.request(1., 2.).get() + f2x.get(1., 2.) + f3q.request(1., 2.).get() + f3x.get(1., 2.);
  s += f3t.get(1., 2.) + f2qt.get(1., 2.) + f2xt.request(1., 2.).get() + f3qt.get(1., 2.) + f3xt.get(1., 2.);
  // finite-T forms in x0 = q0/k (JSON: /physical/T)
  IntegratorFiniteTx0GPU<4, double, KT> x1(qp, {{32, 16}}, 20., 100., 8, json);
  IntegratorAngleFiniteTx0GPU<4, double, KAngleT> x2(qp, {{32, 8, 16}}, 20., 100., 8, json);
  Integrator4DFiniteTx0GPU<double, K3Q0> x3(qp, {{32, 8, 8, 16}}, 20., 100., 8, json);
  IntegratorFiniteTx0TBB<4, double, KT> x1t(qp, {{32, 16}}, 20., 100., 8, json);
  IntegratorAngleFiniteTx0TBB<4, double, KAngleT> x2t(qp, {{32, 8, 16}}, 20., 100., 8, json);
  Integrator4DFiniteTx0TBB<double, K3Q0> x3t(qp, {{32, 8, 8, 16}}, 20., 100., 8, json);
  x1.set_T(0.1);
  x2.set_x0_extent(50.);
  x3t.set_T(0.);
  s += x1.get(1., 2.) + x2.request(1., 2.).get() + x3.get(1., 2.) + x1t.get(1., 2.) + x2t.request(1., 2.).get() +
       x3t.get(1., 2.);
  IntegratorConstant<4, double, KC0> cst(qp, {{64}}, 20., json);
  s += cst.get(1., 2.) + cst.request(1., 2.).get();
  // FP32 node types (FP32 evaluation, FP64 accumulation) on the newer shapes
  IntegratorFiniteTx0GPU<4, float, KT> xf(qp, {{32, 16}}, 20.f, 100.f, 8, 0.1f);
  Integrator3Dpq0GPU<float, K3Q0> f3qf(qp, {{32, 8, 8, 8}}, 20.f, 2.f);
  Integrator2DCartesianGPU<float, KC2> h2f(qp, {{32, 16}}, 0.f, 256, -1.f, -1.f, 1.f, 1.f);
  s += xf.get(1.f, 2.) + f3qf.request(1.f, 2.).get() + h2f.get(1.f, 2.);
  // request_data over a coordinate grid, through a generated-wrapper-shaped object
  LogarithmicCoordinates1D<double> grid(16, 1e-3, 10., 3.);
  IntegratorGPU<4, double, K1p> ap(qp, {{64}}, 20., json);
  Wrapper<IntegratorGPU<4, double, K1p>> w{ap};
  std::vector<double> out(16);
  update_data(request_data<double>(w, grid, 1., 2.), out.data());
  return s + out[0];
}

#ifdef DFRG_REFERENCE_KERNELS
double instantiate_reference_kernels(const ReferenceJSONValue &json)
{
  QuadratureProvider qp(json);
  using REG = PolynomialExpRegulator<>;
  // ONfiniteT V.cu: IntegratorTBB<3, double, V_kernel<REG>>, request(k, N, T, rhoPhi, m2Pi, m2Sigma)
  IntegratorTBB<3, double, V_kernel<REG>> vt(qp, {{64}}, 20., json);
  IntegratorGPU<3, double, V_kernel<REG>> vg(qp, {{64}}, 20., json);
  double s = vt.get(1., 4., 0.1, 0.2, 0.3, 0.4) + vt.request(1., 4., 0.1, 0.2, 0.3, 0.4).get();
  s += vg.get(1., 4., 0.1, 0.2, 0.3, 0.4) + vg.request(1., 4., 0.1, 0.2, 0.3, 0.4).get();
  // YangMills SP ZA.cu: IntegratorAngleGPU<4, double, ZA_kernel<REG>>, seven dressing functions
  using Tex = TexLinearInterpolator1D<double, LogarithmicCoordinates1D<float>>;
  LogarithmicCoordinates1D<float> coords(64, 1e-3f, 40.f, 5.f);
  Tex Z(coords);
  IntegratorAngleGPU<4, double, ZA_kernel<REG>> za(qp, {{64, 16}}, 20., json);
  s += za.get(1., 0.5, Z, Z, Z, Z, Z, Z, Z, 0.1) + za.request(1., 0.5, Z, Z, Z, Z, Z, Z, Z, 0.1).get();
  return s;
}
#endif
