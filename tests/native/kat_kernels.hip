// kat_kernels.hip -- the DiFfRG-compatible integrator templates instantiated with user KERNEL
// structs, exported through extern "C" for the GPU tests: the polynomial integrands of the
// reference tests (tests/physics/integration/integrator_*_cpu.cc, integration_finiteT/
// integrator_finiteTq0_cpu.cc) and the three built-in integrands written as ordinary KERNELs
// with regulators and LinearInterpolator1D.  Each function returns 0, or -1 (invalid argument) /
// -2 (runtime error) with the exception message in kat_last_error().
#include <thread>

#include <DiFfRG/common/math.hh>
#include <DiFfRG/common/quadrature/quadrature_provider.hh>
#include <DiFfRG/discretization/grid/coordinates.hh>
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
#include <DiFfRG/discretization/grid/combined_coordinates.hh>
#include <DiFfRG/discretization/grid/stack_coordinates.hh>
#include <DiFfRG/physics/interpolation/linear_interpolation_1D.hh>
#include <DiFfRG/physics/interpolation/linear_interpolation_2D.hh>
#include <DiFfRG/physics/interpolation/linear_interpolation_3D.hh>
#include <DiFfRG/physics/interpolation/tex_linear_interpolation_1D.hh>
#include <DiFfRG/physics/interpolation/tex_linear_interpolation_1D_stack.hh>
#include <DiFfRG/physics/interpolation/tex_linear_interpolation_2D.hh>
#include <DiFfRG/physics/interpolation/tex_linear_interpolation_3D.hh>
#include <DiFfRG/physics/regulators.hh>
#include <DiFfRG/physics/thermodynamics.hh>
#include <DiFfRG/physics/utils.hh>

// A minimal forward-mode dual number with the interface autodiff::real offers (value, derivative,
// construction from both, arithmetic with doubles): the derivative channel of the FP32 tables, the
// generic-NT path of LinearInterpolator1D and dual-valued integrals, without the autodiff library.
struct Dual1 {
  double v = 0.0, d = 0.0;
  __host__ __device__ Dual1() {}
  __host__ __device__ Dual1(double x) : v(x) {}
  __host__ __device__ Dual1(double x, double y) : v(x), d(y) {}
  __host__ __device__ Dual1 &operator+=(const Dual1 &o)
  {
    v += o.v;
    d += o.d;
    return *this;
  }
};
__host__ __device__ inline Dual1 operator*(const Dual1 &a, double s) { return Dual1(a.v * s, a.d * s); }
__host__ __device__ inline Dual1 operator*(double s, const Dual1 &a) { return Dual1(s * a.v, s * a.d); }
__host__ __device__ inline Dual1 operator+(const Dual1 &a, const Dual1 &b) { return Dual1(a.v + b.v, a.d + b.d); }

namespace dfrg
{
  template <> struct DualTraits<Dual1> {
    static constexpr bool is_dual = true;
    __host__ __device__ static double value(const Dual1 &x) { return x.v; }
    __host__ __device__ static double derivative(const Dual1 &x) { return x.d; }
    __host__ __device__ static Dual1 make(double v, double d) { return Dual1(v, d); }
  };
} // namespace dfrg
#include <dfrg/explicit_rk.hh>
#include <dfrg/x_extent.hh>

#include <algorithm>
#include <chrono>
#include <memory>
#include <string>
#include <vector>

using namespace DiFfRG;

namespace
{
  std::string g_err;
  QuadratureProvider &provider()
  {
    static std::unique_ptr<QuadratureProvider> qp;
    if (!qp) qp = std::make_unique<QuadratureProvider>();
    return *qp;
  }
  template <class F> int guard(F &&f)
  {
    try {
      f();
      return 0;
    } catch (const std::invalid_argument &e) {
      g_err = e.what();
      return -1;
    } catch (const std::exception &e) {
      g_err = e.what();
      return -2;
    }
  }

  // integrator_cpu.cc:15-32
  struct PolyRadial {
    static __forceinline__ __host__ __device__ auto kernel(const double q, const double, const double, const double x0,
                                                           const double x1, const double x2, const double x3,
                                                           const double x4, const double x5)
    {
      return x0 + x1 * powr<1>(q) + x2 * powr<2>(q) + x3 * powr<3>(q) + x4 * powr<4>(q) + x5 * powr<5>(q);
    }
    static __forceinline__ __host__ __device__ auto constant(const double, const double c, const double, const double,
                                                             const double, const double, const double, const double)
    {
      return c;
    }
  };
  // complex-valued variant: i * poly(q) + poly(q)
  struct PolyRadialC {
    static __forceinline__ __host__ __device__ complex<double> kernel(const double q, const double k, const double c,
                                                                      const double x0, const double x1, const double x2,
                                                                      const double x3, const double x4,
                                                                      const double x5)
    {
      const double v = PolyRadial::kernel(q, k, c, x0, x1, x2, x3, x4, x5);
      return complex<double>(v, 2.0 * v);
    }
    static __forceinline__ __host__ __device__ complex<double> constant(const double, const double c, const double,
                                                                        const double, const double, const double,
                                                                        const double, const double)
    {
      return complex<double>(c, 0.0);
    }
  };
  // complex<float> variant: poly(q) (1 + 2i) on FP32 nodes
  struct PolyRadialCF {
    static __forceinline__ __host__ __device__ complex<float> kernel(const float q, const float, const float,
                                                                     const float x0, const float x1, const float x2,
                                                                     const float x3, const float x4, const float x5)
    {
      const float v = x0 + x1 * powr<1>(q) + x2 * powr<2>(q) + x3 * powr<3>(q) + x4 * powr<4>(q) + x5 * powr<5>(q);
      return complex<float>(v, 2.0f * v);
    }
    static __forceinline__ __host__ __device__ complex<float> constant(const float, const float c, const float,
                                                                       const float, const float, const float,
                                                                       const float, const float)
    {
      return complex<float>(c, 0.0f);
    }
  };
  // FP32 variant of PolyRadial (NT = ctype = float)
  struct PolyRadialF {
    static __forceinline__ __host__ __device__ float kernel(const float q, const float, const float, const float x0,
                                                            const float x1, const float x2, const float x3,
                                                            const float x4, const float x5)
    {
      return x0 + x1 * powr<1>(q) + x2 * powr<2>(q) + x3 * powr<3>(q) + x4 * powr<4>(q) + x5 * powr<5>(q);
    }
    static __forceinline__ __host__ __device__ float constant(const float, const float c, const float, const float,
                                                              const float, const float, const float, const float)
    {
      return c;
    }
  };
  // dual-valued variant: (poly_f(q), poly_g(q)) with constant (f0, g0)
  struct PolyRadialDual {
    static __forceinline__ __host__ __device__ Dual1 kernel(const double q, const double, const double f0,
                                                            const double f1, const double f2, const double f3,
                                                            const double f4, const double f5, const double g0,
                                                            const double g1, const double g2, const double g3,
                                                            const double g4, const double g5)
    {
      return Dual1(f0 + f1 * powr<1>(q) + f2 * powr<2>(q) + f3 * powr<3>(q) + f4 * powr<4>(q) + f5 * powr<5>(q),
                   g0 + g1 * powr<1>(q) + g2 * powr<2>(q) + g3 * powr<3>(q) + g4 * powr<4>(q) + g5 * powr<5>(q));
    }
    static __forceinline__ __host__ __device__ Dual1 constant(const double, const double f0, const double,
                                                              const double, const double, const double, const double,
                                                              const double g0, const double, const double, const double,
                                                              const double, const double)
    {
      return Dual1(f0, g0);
    }
  };
  // integrator_4D_2ang_cpu.cc:14-36
  struct Poly2Ang {
    static __forceinline__ __host__ __device__ auto kernel(const double q, const double cos1, const double cos2,
                                                           const double, const double, const double x0, const double x1,
                                                           const double x2, const double x3, const double x4,
                                                           const double x5, const double a0, const double a1,
                                                           const double a2, const double a3, const double b0,
                                                           const double b1, const double b2, const double b3)
    {
      return (x0 + x1 * powr<1>(q) + x2 * powr<2>(q) + x3 * powr<3>(q) + x4 * powr<4>(q) + x5 * powr<5>(q)) *
             (a0 + a1 * powr<1>(cos1) + a2 * powr<2>(cos1) + a3 * powr<3>(cos1)) *
             (b0 + b1 * powr<1>(cos2) + b2 * powr<2>(cos2) + b3 * powr<3>(cos2));
    }
    static __forceinline__ __host__ __device__ auto constant(const double, const double c, const double, const double,
                                                             const double, const double, const double, const double,
                                                             const double, const double, const double, const double,
                                                             const double, const double, const double, const double)
    {
      return c;
    }
  };
  // integrator_4D_cpu.cc:14-39
  struct Poly3Ang {
    static __forceinline__ __host__ __device__ auto
    kernel(const double q, const double cos1, const double cos2, const double phi, const double, const double,
           const double x0, const double x1, const double x2, const double x3, const double x4, const double x5,
           const double a0, const double a1, const double a2, const double a3, const double b0, const double b1,
           const double b2, const double b3, const double p0, const double p1, const double p2, const double p3)
    {
      return (x0 + x1 * powr<1>(q) + x2 * powr<2>(q) + x3 * powr<3>(q) + x4 * powr<4>(q) + x5 * powr<5>(q)) *
             (a0 + a1 * powr<1>(cos1) + a2 * powr<2>(cos1) + a3 * powr<3>(cos1)) *
             (b0 + b1 * powr<1>(cos2) + b2 * powr<2>(cos2) + b3 * powr<3>(cos2)) *
             (p0 + p1 * powr<1>(phi) + p2 * powr<2>(phi) + p3 * powr<3>(phi));
    }
    static __forceinline__ __host__ __device__ auto constant(const double, const double c, const double, const double,
                                                             const double, const double, const double, const double,
                                                             const double, const double, const double, const double,
                                                             const double, const double, const double, const double,
                                                             const double, const double, const double, const double)
    {
      return c;
    }
  };
  // integrator_finiteTq0_cpu.cc:14-34
  struct PolyFiniteT {
    static __forceinline__ __host__ __device__ auto kernel(const double q, const double q0, const double,
                                                           const double, const double x0, const double x1,
                                                           const double x2, const double x3, const double x4,
                                                           const double x5, const double d0, const double d1,
                                                           const double d2, const double d3)
    {
      return (x0 + x1 * powr<1>(q) + x2 * powr<2>(q) + x3 * powr<3>(q) + x4 * powr<4>(q) + x5 * powr<5>(q)) /
             (d0 + d1 * powr<1>(q0) + d2 * powr<2>(q0) + d3 * powr<3>(q0));
    }
    static __forceinline__ __host__ __device__ auto constant(const double, const double c, const double, const double,
                                                             const double, const double, const double, const double,
                                                             const double, const double, const double, const double)
    {
      return c;
    }
  };

  // integrator_angle_gpu.cu:14-32
  struct PolyAngle {
    static __forceinline__ __host__ __device__ auto kernel(const double q, const double cos, const double,
                                                           const double, const double x0, const double x1,
                                                           const double x2, const double x3, const double x4,
                                                           const double x5, const double a0, const double a1,
                                                           const double a2, const double a3)
    {
      return (x0 + x1 * powr<1>(q) + x2 * powr<2>(q) + x3 * powr<3>(q) + x4 * powr<4>(q) + x5 * powr<5>(q)) *
             (a0 + a1 * powr<1>(cos) + a2 * powr<2>(cos) + a3 * powr<3>(cos));
    }
    static __forceinline__ __host__ __device__ auto constant(const double, const double c, const double, const double,
                                                             const double, const double, const double, const double,
                                                             const double, const double, const double, const double)
    {
      return c;
    }
  };
  // integrator_4D_finiteTq0_gpu.cu:16-39
  struct Poly4DFiniteT {
    static __forceinline__ __host__ __device__ auto
    kernel(const double q, const double cos, const double phi, const double q0, const double, const double,
           const double x0, const double x1, const double x2, const double x3, const double x4, const double x5,
           const double a0, const double a1, const double a2, const double a3, const double p0, const double p1,
           const double p2, const double p3, const double d0, const double d1, const double d2, const double d3,
           const double d4)
    {
      return (x0 + x1 * powr<1>(q) + x2 * powr<2>(q) + x3 * powr<3>(q) + x4 * powr<4>(q) + x5 * powr<5>(q)) *
             (a0 + a1 * powr<1>(cos) + a2 * powr<2>(cos) + a3 * powr<3>(cos)) *
             (p0 + p1 * powr<1>(phi) + p2 * powr<2>(phi) + p3 * powr<3>(phi)) /
             (d0 + d1 * powr<1>(q0) + d2 * powr<2>(q0) + d3 * powr<3>(q0) + d4 * powr<4>(q0));
    }
    static __forceinline__ __host__ __device__ auto
    constant(const double, const double c, const double, const double, const double, const double, const double,
             const double, const double, const double, const double, const double, const double, const double,
             const double, const double, const double, const double, const double, const double, const double)
    {
      return c;
    }
  };
  // integrator_angle_finiteTq0_gpu.cu:14-34
  struct PolyAngleFiniteT {
    static __forceinline__ __host__ __device__ auto
    kernel(const double q, const double cos, const double q0, const double, const double, const double x0,
           const double x1, const double x2, const double x3, const double x4, const double x5, const double a0,
           const double a1, const double a2, const double a3, const double d0, const double d1, const double d2,
           const double d3)
    {
      return (x0 + x1 * powr<1>(q) + x2 * powr<2>(q) + x3 * powr<3>(q) + x4 * powr<4>(q) + x5 * powr<5>(q)) *
             (a0 + a1 * powr<1>(cos) + a2 * powr<2>(cos) + a3 * powr<3>(cos)) /
             (d0 + d1 * powr<1>(q0) + d2 * powr<2>(q0) + d3 * powr<3>(q0));
    }
    static __forceinline__ __host__ __device__ auto constant(const double, const double c, const double, const double,
                                                             const double, const double, const double, const double,
                                                             const double, const double, const double, const double,
                                                             const double, const double, const double, const double)
    {
      return c;
    }
  };

  // integrator_1D_cartesian_gpu.cu:16-31: a cubic in qx
  __forceinline__ __host__ __device__ double cubic(double v, double c0, double c1, double c2, double c3)
  {
    return c0 + c1 * powr<1>(v) + c2 * powr<2>(v) + c3 * powr<3>(v);
  }
  struct PolyCart1 {
    static __forceinline__ __host__ __device__ auto kernel(const double qx, const double, const double,
                                                           const double x0, const double x1, const double x2,
                                                           const double x3)
    {
      return cubic(qx, x0, x1, x2, x3);
    }
    static __forceinline__ __host__ __device__ auto constant(const double, const double c, const double, const double,
                                                             const double, const double)
    {
      return c;
    }
  };
  // complex-valued and dual-valued variants: (1 + i/2) cubic and (cubic, cubic/4)
  struct PolyCart1C {
    static __forceinline__ __host__ __device__ complex<double> kernel(const double qx, const double, const double,
                                                                      const double x0, const double x1,
                                                                      const double x2, const double x3)
    {
      const double v = cubic(qx, x0, x1, x2, x3);
      return complex<double>(v, 0.5 * v);
    }
    static __forceinline__ __host__ __device__ complex<double> constant(const double, const double c, const double,
                                                                        const double, const double, const double)
    {
      return complex<double>(c, 0.0);
    }
  };
  struct PolyCart1D {
    static __forceinline__ __host__ __device__ Dual1 kernel(const double qx, const double, const double,
                                                            const double x0, const double x1, const double x2,
                                                            const double x3)
    {
      const double v = cubic(qx, x0, x1, x2, x3);
      return Dual1(v, 0.25 * v);
    }
    static __forceinline__ __host__ __device__ Dual1 constant(const double, const double c, const double,
                                                              const double, const double, const double)
    {
      return Dual1(c, 0.0);
    }
  };
  // integrator_2D_cartesian_gpu.cu:14-31
  struct PolyCart2 {
    static __forceinline__ __host__ __device__ auto kernel(const double qx, const double qy, const double,
                                                           const double, const double x0, const double x1,
                                                           const double x2, const double x3, const double y0,
                                                           const double y1, const double y2, const double y3)
    {
      return cubic(qx, x0, x1, x2, x3) * cubic(qy, y0, y1, y2, y3);
    }
    static __forceinline__ __host__ __device__ auto constant(const double, const double c, const double, const double,
                                                             const double, const double, const double, const double,
                                                             const double, const double)
    {
      return c;
    }
  };
  // integrator_3D_cartesian_gpu.cu:14-36
  struct PolyCart3 {
    static __forceinline__ __host__ __device__ auto
    kernel(const double qx, const double qy, const double qz, const double, const double, const double x0,
           const double x1, const double x2, const double x3, const double y0, const double y1, const double y2,
           const double y3, const double z0, const double z1, const double z2, const double z3)
    {
      return cubic(qx, x0, x1, x2, x3) * cubic(qy, y0, y1, y2, y3) * cubic(qz, z0, z1, z2, z3);
    }
    static __forceinline__ __host__ __device__ auto constant(const double, const double c, const double, const double,
                                                             const double, const double, const double, const double,
                                                             const double, const double, const double, const double,
                                                             const double, const double)
    {
      return c;
    }
  };

  __forceinline__ __host__ __device__ double quintic(double v, double x0, double x1, double x2, double x3, double x4,
                                                     double x5)
  {
    return x0 + x1 * powr<1>(v) + x2 * powr<2>(v) + x3 * powr<3>(v) + x4 * powr<4>(v) + x5 * powr<5>(v);
  }
  // integrator_3D_gpu.cu:16-37: poly(q) cos_poly(cos) phi_poly(phi)
  struct PolyFreq3D {
    static __forceinline__ __host__ __device__ auto
    kernel(const double q, const double cos, const double phi, const double, const double, const double x0,
           const double x1, const double x2, const double x3, const double x4, const double x5, const double c0,
           const double c1, const double c2, const double c3, const double p0, const double p1, const double p2,
           const double p3)
    {
      return quintic(q, x0, x1, x2, x3, x4, x5) * cubic(cos, c0, c1, c2, c3) * cubic(phi, p0, p1, p2, p3);
    }
    static __forceinline__ __host__ __device__ auto constant(const double, const double c, const double, const double,
                                                             const double, const double, const double, const double,
                                                             const double, const double, const double, const double,
                                                             const double, const double, const double, const double)
    {
      return c;
    }
  };
  // integrator_2Dpq0_gpu.cu:16-37 and integrator_2Dpx0_gpu.cu: poly(q) cos_poly(cos) q0_poly(q0)
  struct PolyFreq2Q0 {
    static __forceinline__ __host__ __device__ auto
    kernel(const double q, const double cos, const double q0, const double, const double, const double x0,
           const double x1, const double x2, const double x3, const double x4, const double x5, const double c0,
           const double c1, const double c2, const double c3, const double z0, const double z1, const double z2,
           const double z3)
    {
      return quintic(q, x0, x1, x2, x3, x4, x5) * cubic(cos, c0, c1, c2, c3) * cubic(q0, z0, z1, z2, z3);
    }
    static __forceinline__ __host__ __device__ auto constant(const double, const double c, const double, const double,
                                                             const double, const double, const double, const double,
                                                             const double, const double, const double, const double,
                                                             const double, const double, const double, const double)
    {
      return c;
    }
  };
  // integrator_3Dpq0_gpu.cu:16-38 and integrator_3Dpx0_gpu.cu: poly(q) cos_poly phi_poly q0_poly
  struct PolyFreq3Q0 {
    static __forceinline__ __host__ __device__ auto
    kernel(const double q, const double cos, const double phi, const double q0, const double, const double,
           const double x0, const double x1, const double x2, const double x3, const double x4, const double x5,
           const double c0, const double c1, const double c2, const double c3, const double p0, const double p1,
           const double p2, const double p3, const double z0, const double z1, const double z2, const double z3)
    {
      return quintic(q, x0, x1, x2, x3, x4, x5) * cubic(cos, c0, c1, c2, c3) * cubic(phi, p0, p1, p2, p3) *
             cubic(q0, z0, z1, z2, z3);
    }
    static __forceinline__ __host__ __device__ auto
    constant(const double, const double c, const double, const double, const double, const double, const double,
             const double, const double, const double, const double, const double, const double, const double,
             const double, const double, const double, const double, const double, const double)
    {
      return c;
    }
  };

  // integrator_finiteTx0_gpu.cu:14-33: poly(q) q0_poly(q0)
  struct PolyTx0 {
    static __forceinline__ __host__ __device__ auto kernel(const double q, const double q0, const double, const double,
                                                           const double x0, const double x1, const double x2,
                                                           const double x3, const double x4, const double x5,
                                                           const double z0, const double z1, const double z2,
                                                           const double z3)
    {
      return quintic(q, x0, x1, x2, x3, x4, x5) * cubic(q0, z0, z1, z2, z3);
    }
    static __forceinline__ __host__ __device__ auto constant(const double, const double c, const double, const double,
                                                             const double, const double, const double, const double,
                                                             const double, const double, const double, const double)
    {
      return c;
    }
  };
  // integrator_angle_finiteTx0_gpu.cu: poly(q) cos_poly(cos) q0_poly(q0)
  struct PolyTx0Angle {
    static __forceinline__ __host__ __device__ auto
    kernel(const double q, const double cos, const double q0, const double, const double, const double x0,
           const double x1, const double x2, const double x3, const double x4, const double x5, const double c0,
           const double c1, const double c2, const double c3, const double z0, const double z1, const double z2,
           const double z3)
    {
      return quintic(q, x0, x1, x2, x3, x4, x5) * cubic(cos, c0, c1, c2, c3) * cubic(q0, z0, z1, z2, z3);
    }
    static __forceinline__ __host__ __device__ auto constant(const double, const double c, const double, const double,
                                                             const double, const double, const double, const double,
                                                             const double, const double, const double, const double,
                                                             const double, const double, const double, const double)
    {
      return c;
    }
  };

  // integrator_constant.cc:12-24
  struct PolyConstant {
    static __forceinline__ __host__ __device__ auto kernel(const double, const double, const double x0) { return x0; }
    static __forceinline__ __host__ __device__ auto constant(const double, const double c, const double) { return c; }
  };

  struct B1 {
    static constexpr double b = 1.;
  };
  using Reg = ExponentialRegulator<B1>;
  using ZInterp = LinearInterpolator1D<double, LogarithmicCoordinates1D<double>>;

  // LPA threshold integrand RBdot / (q^2 + RB + m2)^2
  struct ThresholdKernel {
    static __forceinline__ __host__ __device__ double kernel(const double q, const double k, const double m2)
    {
      const double k2 = powr<2>(k), q2 = powr<2>(q);
      return Reg::RBdot(k2, q2) / powr<2>(q2 + Reg::RB(k2, q2) + m2);
    }
    static __forceinline__ __host__ __device__ double constant(const double, const double) { return 0.; }
  };
  // self-energy RBdot(q) G(q)^2 G(|p-q|), G(y) = 1/(Z(y) y^2 + RB(y^2))
  struct SelfEnergyKernel {
    static __forceinline__ __host__ __device__ double G(const double k2, const double y2, const ZInterp &Z)
    {
      return 1.0 / (Z(sqrt(y2)) * y2 + Reg::RB(k2, y2));
    }
    static __forceinline__ __host__ __device__ double kernel(const double q, const double cos1, const double,
                                                             const double k, const double p, const ZInterp &Z)
    {
      const double k2 = powr<2>(k), q2 = powr<2>(q);
      const double Gq = G(k2, q2, Z);
      return Reg::RBdot(k2, q2) * Gq * Gq * G(k2, powr<2>(p) + q2 - 2. * p * q * cos1, Z);
    }
    static __forceinline__ __host__ __device__ double constant(const double, const double, const ZInterp &) { return 0.; }
  };
  // the same self-energy integrand written for FP32 arguments and an FP32 table: with NT = float
  // the integrator evaluates it in FP32 on FP32 nodes and accumulates in FP64
  using ZInterpF = LinearInterpolator1D<float, LogarithmicCoordinates1D<float>>;
  struct SelfEnergyKernelF {
    static __forceinline__ __host__ __device__ float G(const float k2, const float y2, const ZInterpF &Z)
    {
      return 1.0f / (Z(sqrtf(y2)) * y2 + k2 * expf(-y2 / k2));
    }
    static __forceinline__ __host__ __device__ float kernel(const float q, const float cos1, const float,
                                                            const float k, const float p, const ZInterpF &Z)
    {
      const float k2 = k * k, q2 = q * q, x = q2 / k2;
      const float Gq = G(k2, q2, Z);
      return 2.0f * expf(-x) * k2 * (1.0f + x) * Gq * Gq * G(k2, p * p + q2 - 2.0f * p * q * cos1, Z);
    }
    static __forceinline__ __host__ __device__ float constant(const float, const float, const ZInterpF &) { return 0.f; }
  };

  // finite-T two-propagator integrand
  struct FiniteTKernel {
    static __forceinline__ __host__ __device__ double G(const double k2, const double w, const double s, const double m)
    {
      return 1.0 / (powr<2>(w) + s + Reg::RB(k2, s) + m);
    }
    static __forceinline__ __host__ __device__ double kernel(const double q, const double q0, const double k,
                                                             const double p0, const double pabs, const double m1,
                                                             const double m2)
    {
      const double k2 = powr<2>(k), q2 = powr<2>(q);
      const double g1 = G(k2, q0, q2, m1);
      return Reg::RBdot(k2, q2) * g1 * g1 * G(k2, q0 + p0, q2 + powr<2>(pabs), m2);
    }
    static __forceinline__ __host__ __device__ double constant(const double, const double, const double, const double,
                                                               const double)
    {
      return 0.;
    }
  };
} // namespace

extern "C" {

const char *kat_last_error() { return g_err.c_str(); }

// device = 1: GPU template, 0: TBB template
int kat_radial(int device, int d, int nx, double x_extent, double k, const double *c, double constant, double *out)
{
  return guard([&] {
    auto run = [&](auto integ) {
      *out = integ.get(k, constant, c[0], c[1], c[2], c[3], c[4], c[5]);
    };
    switch (d * 2 + device) {
    case 2: run(IntegratorTBB<1, double, PolyRadial>(provider(), nx, x_extent)); break;
    case 3: run(IntegratorGPU<1, double, PolyRadial>(provider(), nx, x_extent)); break;
    case 4: run(IntegratorTBB<2, double, PolyRadial>(provider(), nx, x_extent)); break;
    case 5: run(IntegratorGPU<2, double, PolyRadial>(provider(), nx, x_extent)); break;
    case 6: run(IntegratorTBB<3, double, PolyRadial>(provider(), nx, x_extent)); break;
    case 7: run(IntegratorGPU<3, double, PolyRadial>(provider(), nx, x_extent)); break;
    case 8: run(IntegratorTBB<4, double, PolyRadial>(provider(), nx, x_extent)); break;
    case 9: run(IntegratorGPU<4, double, PolyRadial>(provider(), nx, x_extent)); break;
    default: throw std::invalid_argument("unsupported dimension");
    }
  });
}

int kat_radial_complex(int nx, double x_extent, double k, const double *c, double constant, double *out_re,
                       double *out_im)
{
  return guard([&] {
    IntegratorGPU<4, complex<double>, PolyRadialC> integ(provider(), nx, x_extent);
    const auto v = integ.get(k, constant, c[0], c[1], c[2], c[3], c[4], c[5]);
    *out_re = v.re;
    *out_im = v.im;
  });
}

} // extern "C"

namespace
{
  __global__ void wave_sum8_kernel(const double *in, double *out)
  {
    const int l = threadIdx.x; // one wave
    double v[8];
    for (int i = 0; i < 8; ++i) v[i] = in[i * 64 + l];
    out[l] = dfrg::dev::wave_sum8(v);
  }
  __global__ void wave_sum4_kernel(const double *in, double *out)
  {
    const int l = threadIdx.x; // one wave
    double v[4];
    for (int i = 0; i < 4; ++i) v[i] = in[i * 64 + l];
    out[l] = dfrg::dev::wave_sum4(v);
  }
} // namespace

extern "C" {

// dfrg::dev::wave_sum8 on one wave: in holds 8 vectors of 64 lanes (vector-major), out the 64
// lane results
int kat_wave_sum8(const double *in, double *out)
{
  return guard([&] {
    void *d_in = nullptr, *d_out = nullptr;
    dfrg::check(dfrg_malloc(&d_in, sizeof(double) * 512), "in");
    dfrg::check(dfrg_malloc(&d_out, sizeof(double) * 64), "out");
    void *s = nullptr;
    dfrg::check(dfrg_thread_stream(&s), "thread stream");
    dfrg::check(dfrg_memcpy_h2d(d_in, in, sizeof(double) * 512, s), "in");
    hipLaunchKernelGGL(wave_sum8_kernel, dim3(1), dim3(64), 0, (hipStream_t)s, (const double *)d_in, (double *)d_out);
    dfrg::check_hip("wave_sum8_kernel");
    dfrg::check(dfrg_memcpy_d2h(out, d_out, sizeof(double) * 64, s), "out");
    dfrg::check(dfrg_stream_sync(s), "out");
    dfrg_free(d_in);
    dfrg_free(d_out);
  });
}

// dfrg::dev::wave_sum4 on one wave: in holds 4 vectors of 64 lanes (vector-major), out the 64
// lane results
int kat_wave_sum4(const double *in, double *out)
{
  return guard([&] {
    void *d_in = nullptr, *d_out = nullptr;
    dfrg::check(dfrg_malloc(&d_in, sizeof(double) * 256), "in");
    dfrg::check(dfrg_malloc(&d_out, sizeof(double) * 64), "out");
    void *s = nullptr;
    dfrg::check(dfrg_thread_stream(&s), "thread stream");
    dfrg::check(dfrg_memcpy_h2d(d_in, in, sizeof(double) * 256, s), "in");
    hipLaunchKernelGGL(wave_sum4_kernel, dim3(1), dim3(64), 0, (hipStream_t)s, (const double *)d_in, (double *)d_out);
    dfrg::check_hip("wave_sum4_kernel");
    dfrg::check(dfrg_memcpy_d2h(out, d_out, sizeof(double) * 64, s), "out");
    dfrg::check(dfrg_stream_sync(s), "out");
    dfrg_free(d_in);
    dfrg_free(d_out);
  });
}

// NT = complex<float>: get and a batched request; out = (get re, get im, request re, request im)
int kat_radial_complex_float(int nx, float x_extent, float k, const double *c, float constant, double *out)
{
  return guard([&] {
    IntegratorGPU<4, complex<float>, PolyRadialCF> integ(provider(), nx, x_extent);
    const float cf[6] = {(float)c[0], (float)c[1], (float)c[2], (float)c[3], (float)c[4], (float)c[5]};
    const auto a = integ.get(k, constant, cf[0], cf[1], cf[2], cf[3], cf[4], cf[5]);
    const auto b = integ.request(k, constant, cf[0], cf[1], cf[2], cf[3], cf[4], cf[5]).get();
    out[0] = a.re;
    out[1] = a.im;
    out[2] = b.re;
    out[3] = b.im;
  });
}

// Concurrent callers, as the reference's TBB get_data runs get() (physics/utils.hh:205-220): n_thr
// host threads share one radial and one two-angle integrator; thread t evaluates ks[i] for
// i = t, t + n_thr, ..., alternating get() and request().get().  out_conc[i] (radial) and
// out_conc[n + i] (two-angle) against the same calls made one after another in out_seq.
int kat_concurrent(int n_thr, const double *ks, int n, const double *c, const double *a, const double *bc,
                   double *out_conc, double *out_seq)
{
  return guard([&] {
    IntegratorGPU<4, double, PolyRadial> rad(provider(), 64, 1.5);
    Integrator4D2AngGPU<double, Poly2Ang> ang(provider(), {{48, 12, 12}}, 1.5);
    auto eval_rad = [&](double k, bool req) {
      return req ? rad.request(k, 0.25, c[0], c[1], c[2], c[3], c[4], c[5]).get()
                 : rad.get(k, 0.25, c[0], c[1], c[2], c[3], c[4], c[5]);
    };
    auto eval_ang = [&](double k, bool req) {
      return req ? ang.request(k, 0.5, c[0], c[1], c[2], c[3], c[4], c[5], a[0], a[1], a[2], a[3], bc[0], bc[1],
                               bc[2], bc[3]).get()
                 : ang.get(k, 0.5, c[0], c[1], c[2], c[3], c[4], c[5], a[0], a[1], a[2], a[3], bc[0], bc[1], bc[2],
                           bc[3]);
    };
    for (int i = 0; i < n; ++i) {
      out_seq[i] = eval_rad(ks[i], false);
      out_seq[n + i] = eval_ang(ks[i], false);
    }
    std::vector<std::thread> th;
    std::vector<std::string> errs(n_thr);
    for (int t = 0; t < n_thr; ++t)
      th.emplace_back([&, t] {
        try {
          for (int i = t; i < n; i += n_thr) {
            out_conc[i] = eval_rad(ks[i], (i / n_thr) % 2 == 1);
            out_conc[n + i] = eval_ang(ks[i], (i / n_thr) % 2 == 0);
          }
        } catch (const std::exception &e) {
          errs[t] = e.what();
        }
      });
    for (auto &x : th) x.join();
    for (auto &e : errs)
      if (!e.empty()) throw std::runtime_error(e);
  });
}

// n_req futures issued before any is waited on (request/update_data pattern)
int kat_radial_requests(int nx, double x_extent, const double *ks, int n_req, const double *c, double *out)
{
  return guard([&] {
    IntegratorGPU<4, double, PolyRadial> integ(provider(), nx, x_extent);
    std::vector<std::future<double>> f;
    for (int i = 0; i < n_req; ++i) f.push_back(integ.request(ks[i], 0.0, c[0], c[1], c[2], c[3], c[4], c[5]));
    for (int i = 0; i < n_req; ++i) out[i] = f[i].get();
  });
}

int kat_2ang(int device, int nx, int na1, int na2, double x_extent, double k, const double *c, const double *a,
             const double *b, double constant, double *out)
{
  return guard([&] {
    auto run = [&](auto integ) {
      *out = integ.get(k, constant, c[0], c[1], c[2], c[3], c[4], c[5], a[0], a[1], a[2], a[3], b[0], b[1], b[2],
                       b[3]);
    };
    const std::array<uint, 3> g{{(uint)nx, (uint)na1, (uint)na2}};
    if (device)
      run(Integrator4D2AngGPU<double, Poly2Ang>(provider(), g, x_extent));
    else
      run(Integrator4D2AngTBB<double, Poly2Ang>(provider(), g, x_extent));
  });
}

int kat_3ang(int device, int nx, int na, double x_extent, double k, const double *c, const double *a, const double *b,
             const double *p, double constant, double *out)
{
  return guard([&] {
    auto run = [&](auto integ) {
      *out = integ.get(k, constant, c[0], c[1], c[2], c[3], c[4], c[5], a[0], a[1], a[2], a[3], b[0], b[1], b[2],
                       b[3], p[0], p[1], p[2], p[3]);
    };
    const std::array<uint, 4> g{{(uint)nx, (uint)na, (uint)na, (uint)na}};
    if (device)
      run(Integrator4DGPU<double, Poly3Ang>(provider(), g, x_extent));
    else
      run(Integrator4DTBB<double, Poly3Ang>(provider(), g, x_extent));
  });
}

// *nm receives the Matsubara rule size used for the (last) evaluation
int kat_finiteT(int device, int d, int nx, double x_extent, double T, double k, const double *c, const double *den,
                double constant, double *out, int *nm)
{
  return guard([&] {
    auto run = [&](auto integ) {
      *out = integ.get(k, constant, c[0], c[1], c[2], c[3], c[4], c[5], den[0], den[1], den[2], den[3]);
      (void)nm;
    };
    const std::array<uint, 1> g{{(uint)nx}};
    if (device) {
      switch (d) {
      case 2: { IntegratorFiniteTq0GPU<2, double, PolyFiniteT> i(provider(), g, x_extent, T); run(i); *nm = i.matsubara_size(); break; }
      case 3: { IntegratorFiniteTq0GPU<3, double, PolyFiniteT> i(provider(), g, x_extent, T); run(i); *nm = i.matsubara_size(); break; }
      case 4: { IntegratorFiniteTq0GPU<4, double, PolyFiniteT> i(provider(), g, x_extent, T); run(i); *nm = i.matsubara_size(); break; }
      default: throw std::invalid_argument("unsupported dimension");
      }
    } else {
      switch (d) {
      case 2: run(IntegratorFiniteTq0TBB<2, double, PolyFiniteT>(provider(), g, x_extent, T)); break;
      case 3: run(IntegratorFiniteTq0TBB<3, double, PolyFiniteT>(provider(), g, x_extent, T)); break;
      case 4: run(IntegratorFiniteTq0TBB<4, double, PolyFiniteT>(provider(), g, x_extent, T)); break;
      default: throw std::invalid_argument("unsupported dimension");
      }
    }
  });
}

// E tracking: evaluate at each k in turn; records the rule size and typical E after each call
int kat_finiteT_track(int nx, double x_extent, double T, const double *ks, int n, int *nm, double *E)
{
  return guard([&] {
    IntegratorFiniteTq0GPU<4, double, PolyFiniteT> integ(provider(), {{(uint)nx}}, x_extent, T);
    for (int i = 0; i < n; ++i) {
      (void)integ.get(ks[i], 0.0, 1.0, 0, 0, 0, 0, 0, T * T, 0, 1.0, 0);
      nm[i] = integ.matsubara_size();
      E[i] = integ.typical_E();
    }
  });
}

int kat_threshold(int nx, double x_extent, double k, const double *m2, int n, double *out)
{
  return guard([&] {
    IntegratorGPU<4, double, ThresholdKernel> integ(provider(), nx, x_extent);
    for (int i = 0; i < n; ++i) out[i] = integ.get(k, m2[i]);
  });
}

int kat_selfenergy(int nq, int na, double x_extent, double k, const double *p, int n_p, const double *Z, int n_tab,
                   double start, double stop, double bias, double *out)
{
  return guard([&] {
    LogarithmicCoordinates1D<double> coords((uint)n_tab, start, stop, bias);
    ZInterp interp(Z, coords);
    Integrator4D2AngGPU<double, SelfEnergyKernel> integ(provider(), {{(uint)nq, (uint)na, (uint)na}}, x_extent);
    std::vector<std::future<double>> f;
    for (int i = 0; i < n_p; ++i) f.push_back(integ.request(k, p[i], interp));
    for (int i = 0; i < n_p; ++i) out[i] = f[i].get();
  });
}

// finite-T two-propagator integrand, reference x-mapping, Matsubara rule of the provider at E = k
int kat_finiteT_twoprop(int nx, double x_extent, double k, const double *T, const double *p0, const double *pabs,
                        const double *m1, const double *m2, int n, double *out)
{
  return guard([&] {
    for (int i = 0; i < n; ++i) {
      IntegratorFiniteTq0GPU<4, double, FiniteTKernel> integ(provider(), {{(uint)nx}}, x_extent, T[i]);
      out[i] = integ.get(k, p0[i], pabs[i], m1[i], m2[i]);
    }
  });
}

int kat_unequal_angles()
{
  return guard([&] { Integrator4D2AngGPU<double, Poly2Ang> integ(provider(), {{16, 8, 16}}, 1.0); });
}

} // extern "C"

// ---------------------------------------------------------------------------------------------
// batched requests
// ---------------------------------------------------------------------------------------------
namespace
{
  // the shape of a generated X_integrator wrapper (Examples/*/flows/*/X.hh): templated on the
  // result type, forwarding to the integrator
  struct SelfEnergyWrapper {
    Integrator4D2AngGPU<double, SelfEnergyKernel> integ;
    template <typename NT, typename... T> std::future<NT> request(double k, const T &...t)
    {
      return integ.request(k, t...);
    }
    template <typename NT, typename... T> NT get(double k, const T &...t) { return integ.get(k, t...); }
  };
  // the FP32 form: the wrapper converts k and p to the integrator's ctype
  struct SelfEnergyWrapperF {
    Integrator4D2AngGPU<float, SelfEnergyKernelF> integ;
    template <typename NT, typename... T> std::future<NT> request(double k, double p, const T &...t)
    {
      return integ.request((float)k, (float)p, t...);
    }
  };
} // namespace

extern "C" {

// the self-energy call set with one get() per external momentum (one launch each)
int kat_selfenergy_get(int nq, int na, double x_extent, double k, const double *p, int n_p, const double *Z,
                       int n_tab, double start, double stop, double bias, double *out)
{
  return guard([&] {
    LogarithmicCoordinates1D<double> coords((uint)n_tab, start, stop, bias);
    ZInterp interp(Z, coords);
    Integrator4D2AngGPU<double, SelfEnergyKernel> integ(provider(), {{(uint)nq, (uint)na, (uint)na}}, x_extent);
    for (int i = 0; i < n_p; ++i) out[i] = integ.get(k, p[i], interp);
  });
}

// the same call set, timed: reps rounds of n_p get() calls after one untimed round; *t_call is the
// median over the rounds of the wall time per get() (each is a launch and a synchronisation)
int kat_selfenergy_get_bench(int nq, int na, double x_extent, double k, const double *p, int n_p, const double *Z,
                             int n_tab, double start, double stop, double bias, int reps, double *out,
                             double *t_call)
{
  return guard([&] {
    LogarithmicCoordinates1D<double> coords((uint)n_tab, start, stop, bias);
    ZInterp interp(Z, coords);
    Integrator4D2AngGPU<double, SelfEnergyKernel> integ(provider(), {{(uint)nq, (uint)na, (uint)na}}, x_extent);
    std::vector<double> per;
    for (int r = 0; r <= reps; ++r) {
      const auto t0 = std::chrono::steady_clock::now();
      for (int i = 0; i < n_p; ++i) out[i] = integ.get(k, p[i], interp);
      const auto t1 = std::chrono::steady_clock::now();
      if (r > 0) per.push_back(std::chrono::duration<double>(t1 - t0).count() / n_p);
    }
    std::sort(per.begin(), per.end());
    *t_call = per[per.size() / 2];
  });
}

// request_data / update_data and get_data over the external momenta of a log grid
int kat_request_data(int nq, int na, double x_extent, double k, int n_p, double p_start, double p_stop,
                     double p_bias, const double *Z, int n_tab, double start, double stop, double bias,
                     double *out_req, double *out_get, double *p_out)
{
  return guard([&] {
    LogarithmicCoordinates1D<double> coords((uint)n_tab, start, stop, bias);
    LogarithmicCoordinates1D<double> pgrid((uint)n_p, p_start, p_stop, p_bias);
    ZInterp interp(Z, coords);
    SelfEnergyWrapper w{Integrator4D2AngGPU<double, SelfEnergyKernel>(provider(), {{(uint)nq, (uint)na, (uint)na}},
                                                                       x_extent)};
    auto futures = request_data<double>(w, pgrid, k, interp);
    update_data(futures, out_req);
    get_data<double>(out_get, w, make_grid(pgrid), k, std::tuple<ZInterp>(interp));
    const auto g = make_grid(pgrid);
    for (int i = 0; i < n_p; ++i) p_out[i] = g[i];
  });
}

// finite-T requests over several k (E follows k, so the rule and the batch change on the way)
// against get() at the same k
int kat_finiteT_requests(int nx, double x_extent, double T, const double *ks, int n, const double *c,
                         const double *den, double *out_req, double *out_get)
{
  return guard([&] {
    IntegratorFiniteTq0GPU<4, double, PolyFiniteT> a(provider(), {{(uint)nx}}, x_extent, T);
    IntegratorFiniteTq0GPU<4, double, PolyFiniteT> b(provider(), {{(uint)nx}}, x_extent, T);
    std::vector<std::future<double>> f;
    for (int i = 0; i < n; ++i)
      f.push_back(a.request(ks[i], 0.0, c[0], c[1], c[2], c[3], c[4], c[5], den[0], den[1], den[2], den[3]));
    for (int i = 0; i < n; ++i)
      out_get[i] = b.get(ks[i], 0.0, c[0], c[1], c[2], c[3], c[4], c[5], den[0], den[1], den[2], den[3]);
    for (int i = 0; i < n; ++i) out_req[i] = f[i].get();
  });
}

// wall-clock seconds of n_p get() calls and of n_p requests read back (one batch), repeated reps
// times each after one warm-up round
int kat_request_timing(int nq, int na, double x_extent, double k, const double *p, int n_p, const double *Z,
                       int n_tab, double start, double stop, double bias, int reps, double *t_get, double *t_req)
{
  return guard([&] {
    LogarithmicCoordinates1D<double> coords((uint)n_tab, start, stop, bias);
    ZInterp interp(Z, coords);
    Integrator4D2AngGPU<double, SelfEnergyKernel> integ(provider(), {{(uint)nq, (uint)na, (uint)na}}, x_extent);
    std::vector<double> out(n_p);
    auto gets = [&] {
      for (int i = 0; i < n_p; ++i) out[i] = integ.get(k, p[i], interp);
    };
    auto reqs = [&] {
      std::vector<std::future<double>> f;
      for (int i = 0; i < n_p; ++i) f.push_back(integ.request(k, p[i], interp));
      for (int i = 0; i < n_p; ++i) out[i] = f[i].get();
    };
    gets();
    reqs();
    auto t0 = std::chrono::steady_clock::now();
    for (int r = 0; r < reps; ++r) gets();
    auto t1 = std::chrono::steady_clock::now();
    for (int r = 0; r < reps; ++r) reqs();
    auto t2 = std::chrono::steady_clock::now();
    *t_get = std::chrono::duration<double>(t1 - t0).count() / reps;
    *t_req = std::chrono::duration<double>(t2 - t1).count() / reps;
  });
}

// a large batch of two-angle requests, read back, then get() calls with the same integrator: the
// get() scratch must not inherit the batch's partials or tickets
int kat_selfenergy_after_batch(int nq, int na, double x_extent, double k, const double *p, int n_p, const double *Z,
                               int n_tab, double start, double stop, double bias, int n_batch, double *out_get,
                               double *out_batch)
{
  return guard([&] {
    LogarithmicCoordinates1D<double> coords((unsigned)n_tab, start, stop, bias);
    ZInterp interp(Z, coords);
    Integrator4D2AngGPU<double, SelfEnergyKernel> integ(provider(), {{(unsigned)nq, (unsigned)na, (unsigned)na}},
                                                        x_extent);
    std::vector<std::future<double>> f;
    for (int i = 0; i < n_batch; ++i) f.push_back(integ.request(k, p[i % n_p] * (1.0 + 1e-3 * i), interp));
    for (int i = 0; i < n_batch; ++i) out_batch[i] = f[i].get();
    for (int i = 0; i < n_p; ++i) out_get[i] = integ.get(k, p[i], interp);
  });
}

// request_data over n_p external momenta, host work of sleep_us, then the reads: seconds spent in
// the reads (t_read) against seconds of a request_data + read with no host work between (t_full)
int kat_request_overlap(int nq, int na, double x_extent, double k, int n_p, const double *Z, int n_tab, double start,
                        double stop, double bias, int sleep_us, int reps, double *t_full, double *t_read)
{
  return guard([&] {
    LogarithmicCoordinates1D<double> coords((unsigned)n_tab, start, stop, bias);
    LogarithmicCoordinates1D<double> pgrid((unsigned)n_p, 1e-3, 40.0, 6.0);
    ZInterp interp(Z, coords);
    SelfEnergyWrapper w{Integrator4D2AngGPU<double, SelfEnergyKernel>(
        provider(), {{(unsigned)nq, (unsigned)na, (unsigned)na}}, x_extent)};
    std::vector<double> out(n_p);
    std::vector<double> full, read;
    for (int r = 0; r <= reps; ++r) {
      auto t0 = std::chrono::steady_clock::now();
      update_data(request_data<double>(w, pgrid, k, interp), out.data());
      auto t1 = std::chrono::steady_clock::now();
      auto fut = request_data<double>(w, pgrid, k, interp);
      std::this_thread::sleep_for(std::chrono::microseconds(sleep_us));
      auto t2 = std::chrono::steady_clock::now();
      update_data(fut, out.data());
      auto t3 = std::chrono::steady_clock::now();
      if (r > 0) { // the first round warms up
        full.push_back(std::chrono::duration<double>(t1 - t0).count());
        read.push_back(std::chrono::duration<double>(t3 - t2).count());
      }
    }
    std::sort(full.begin(), full.end());
    std::sort(read.begin(), read.end());
    *t_full = full[full.size() / 2];
    *t_read = read[read.size() / 2];
  });
}

// The self-energy call set as request_data makes it (one request per momentum of a log grid of n_p
// on [1e-3, 40], bias 6, then the reads): results, and the medians over reps of the device time of
// one batch -- HIP events on the thread's stream around its launch and the reads: tuples to HBM,
// the kernel, results back -- and of the host time of the whole call set (requests, launch,
// reads).  fp32 != 0 runs the FP32 form of the integrand (NT = float, FP64 accumulation).
// phases (may be NULL): medians of the host request loop, the host launch (flush), the host reads,
// the device time from before the launch to an event enqueued right after it (copy and kernel plus
// the host's enqueue) and from that event to the one recorded after the reads (the host's wake-up
// and reads), in seconds
int kat_selfenergy_request_bench(int nq, int na, double x_extent, double k, int n_p, const double *Z, int n_tab,
                                 double start, double stop, double bias, int fp32, int reps, double *out,
                                 double *p_out, double *t_dev, double *t_host, double *phases)
{
  return guard([&] {
    LogarithmicCoordinates1D<double> pgrid((unsigned)n_p, 1e-3, 40.0, 6.0);
    const auto grid = make_grid(pgrid);
    for (int i = 0; i < n_p; ++i) p_out[i] = grid[i];
    provider(); // initialises the library on device 0
    void *sv = nullptr;
    dfrg::check(dfrg_thread_stream(&sv), "thread stream");
    hipEvent_t e0, e1, em;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess || hipEventCreate(&em) != hipSuccess)
      throw std::runtime_error("event create failed");
    std::vector<double> dev, host, ph[5];
    auto secs = [](auto a, auto b) { return std::chrono::duration<double>(b - a).count(); };
    auto run = [&](auto &w, auto &grid_c, auto kk, auto &interp, auto *res) {
      using R = std::remove_pointer_t<decltype(res)>;
      const auto g = make_grid(grid_c);
      for (int r = 0; r <= reps; ++r) {
        auto t0 = std::chrono::steady_clock::now();
        // the requests of request_data's loop; the batch is launched between the events
        std::vector<std::future<R>> fut;
        for (unsigned i = 0; i < g.size(); ++i) fut.push_back(w.template request<R>(kk, g[i], interp));
        auto t1 = std::chrono::steady_clock::now();
        (void)hipEventRecord(e0, (hipStream_t)sv);
        dfrg::tmpl::flush_thread();
        (void)hipEventRecord(em, (hipStream_t)sv);
        auto t2 = std::chrono::steady_clock::now();
        for (unsigned i = 0; i < g.size(); ++i) res[i] = fut[i].get();
        (void)hipEventRecord(e1, (hipStream_t)sv);
        (void)hipEventSynchronize(e1);
        auto t3 = std::chrono::steady_clock::now();
        float ms = 0.f, ms0 = 0.f, ms1 = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        (void)hipEventElapsedTime(&ms0, e0, em);
        (void)hipEventElapsedTime(&ms1, em, e1);
        if (r > 0) {
          dev.push_back(ms * 1e-3);
          host.push_back(secs(t0, t3));
          ph[0].push_back(secs(t0, t1));
          ph[1].push_back(secs(t1, t2));
          ph[2].push_back(secs(t2, t3));
          ph[3].push_back(ms0 * 1e-3);
          ph[4].push_back(ms1 * 1e-3);
        }
      }
    };
    if (!fp32) {
      LogarithmicCoordinates1D<double> coords((unsigned)n_tab, start, stop, bias);
      ZInterp interp(Z, coords);
      SelfEnergyWrapper w{Integrator4D2AngGPU<double, SelfEnergyKernel>(
          provider(), {{(unsigned)nq, (unsigned)na, (unsigned)na}}, x_extent)};
      run(w, pgrid, k, interp, out);
    } else {
      LogarithmicCoordinates1D<float> coords((unsigned)n_tab, (float)start, (float)stop, (float)bias);
      std::vector<float> Zf(Z, Z + n_tab);
      ZInterpF interp(Zf, coords);
      SelfEnergyWrapperF w{Integrator4D2AngGPU<float, SelfEnergyKernelF>(
          provider(), {{(unsigned)nq, (unsigned)na, (unsigned)na}}, (float)x_extent)};
      std::vector<float> outf(n_p);
      run(w, pgrid, k, interp, outf.data());
      for (int i = 0; i < n_p; ++i) out[i] = outf[i];
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(em);
    auto median = [](std::vector<double> &v) {
      std::sort(v.begin(), v.end());
      return v[v.size() / 2];
    };
    *t_dev = median(dev);
    *t_host = median(host);
    if (phases)
      for (int i = 0; i < 5; ++i) phases[i] = median(ph[i]);
  });
}

// ---------------------------------------------------------------------------------------------
// angle and finite-T variants: get() on the GPU and host templates, and requests against get()
// ---------------------------------------------------------------------------------------------

int kat_angle(int device, int d, int nx, int na, double x_extent, double k, const double *c, const double *a,
              double constant, double *out)
{
  return guard([&] {
    auto run = [&](auto integ) {
      *out = integ.get(k, constant, c[0], c[1], c[2], c[3], c[4], c[5], a[0], a[1], a[2], a[3]);
    };
    const std::array<uint, 2> g{{(uint)nx, (uint)na}};
    switch (d * 2 + device) {
    case 4: run(IntegratorAngleTBB<2, double, PolyAngle>(provider(), g, x_extent)); break;
    case 5: run(IntegratorAngleGPU<2, double, PolyAngle>(provider(), g, x_extent)); break;
    case 6: run(IntegratorAngleTBB<3, double, PolyAngle>(provider(), g, x_extent)); break;
    case 7: run(IntegratorAngleGPU<3, double, PolyAngle>(provider(), g, x_extent)); break;
    case 8: run(IntegratorAngleTBB<4, double, PolyAngle>(provider(), g, x_extent)); break;
    case 9: run(IntegratorAngleGPU<4, double, PolyAngle>(provider(), g, x_extent)); break;
    default: throw std::invalid_argument("unsupported dimension");
    }
  });
}

// ---------------------------------------------------------------------------------------------
// Cartesian integrators (integrator_{1,2,3}D_cartesian_{cpu,gpu}.hh)
// ---------------------------------------------------------------------------------------------

} // extern "C"

namespace
{
  // I.get / I.request of the cubic product with coefficients c (4 per axis)
  template <int D, class I> auto cart_call(I &integ, bool req, double k, double constant, const double *c)
  {
    if constexpr (D == 1) {
      return req ? integ.request(k, constant, c[0], c[1], c[2], c[3]).get() : integ.get(k, constant, c[0], c[1], c[2], c[3]);
    } else if constexpr (D == 2) {
      return req ? integ.request(k, constant, c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7]).get()
                 : integ.get(k, constant, c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7]);
    } else {
      return req ? integ.request(k, constant, c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7], c[8], c[9], c[10], c[11])
                       .get()
                 : integ.get(k, constant, c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7], c[8], c[9], c[10], c[11]);
    }
  }
  template <int D, class I> auto cart_request(I &integ, double k, double constant, const double *c)
  {
    if constexpr (D == 1)
      return integ.request(k, constant, c[0], c[1], c[2], c[3]);
    else if constexpr (D == 2)
      return integ.request(k, constant, c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7]);
    else
      return integ.request(k, constant, c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7], c[8], c[9], c[10], c[11]);
  }
  // The range of every axis through the mutators: set_q?_min first (max_first: set_q?_max first).
  template <int D, class I> void cart_set(I &integ, const double *lo, const double *hi, bool max_first)
  {
    auto one = [&](auto set_min, auto set_max, int i) {
      if (max_first) {
        set_max(hi[i]);
        set_min(lo[i]);
      } else {
        set_min(lo[i]);
        set_max(hi[i]);
      }
    };
    one([&](double v) { integ.set_qx_min(v); }, [&](double v) { integ.set_qx_max(v); }, 0);
    if constexpr (D > 1) one([&](double v) { integ.set_qy_min(v); }, [&](double v) { integ.set_qy_max(v); }, 1);
    if constexpr (D > 2) one([&](double v) { integ.set_qz_min(v); }, [&](double v) { integ.set_qz_max(v); }, 2);
  }
  template <int D, template <class, class> class G, template <class, class> class H, class K>
  struct CartForms {
    using Gpu = G<double, K>;
    using Host = H<double, K>;
  };
  template <int D> struct CartOf;
  template <> struct CartOf<1> : CartForms<1, Integrator1DCartesianGPU, Integrator1DCartesianTBB, PolyCart1> {
    static Gpu gpu(const int *n, const double *lo, const double *hi)
    {
      return Gpu(provider(), {{(uint)n[0]}}, 0., 256, lo[0], hi[0]);
    }
    static Host host(const int *n, const double *lo, const double *hi)
    {
      return Host(provider(), {{(uint)n[0]}}, 0., 0, lo[0], hi[0]);
    }
  };
  template <> struct CartOf<2> : CartForms<2, Integrator2DCartesianGPU, Integrator2DCartesianTBB, PolyCart2> {
    static Gpu gpu(const int *n, const double *lo, const double *hi)
    {
      return Gpu(provider(), {{(uint)n[0], (uint)n[1]}}, 0., 256, lo[0], lo[1], hi[0], hi[1]);
    }
    static Host host(const int *n, const double *lo, const double *hi)
    {
      return Host(provider(), {{(uint)n[0], (uint)n[1]}}, 0., 0, lo[0], lo[1], hi[0], hi[1]);
    }
  };
  template <> struct CartOf<3> : CartForms<3, Integrator3DCartesianGPU, Integrator3DCartesianTBB, PolyCart3> {
    static Gpu gpu(const int *n, const double *lo, const double *hi)
    {
      return Gpu(provider(), {{(uint)n[0], (uint)n[1], (uint)n[2]}}, 0., 256, lo[0], lo[1], lo[2], hi[0], hi[1],
                 hi[2]);
    }
    static Host host(const int *n, const double *lo, const double *hi)
    {
      return Host(provider(), {{(uint)n[0], (uint)n[1], (uint)n[2]}}, 0., 0, lo[0], lo[1], lo[2], hi[0], hi[1],
                  hi[2]);
    }
  };
  template <class F> void cart_dispatch(int D, F &&f)
  {
    switch (D) {
    case 1: f(std::integral_constant<int, 1>()); break;
    case 2: f(std::integral_constant<int, 2>()); break;
    case 3: f(std::integral_constant<int, 3>()); break;
    default: throw std::invalid_argument("Cartesian integrators have one to three axes");
    }
  }
} // namespace

extern "C" {

// get() (req = 0) or request().get() (req = 1) of the D-axis Cartesian integrator on the GPU
// (device 1) or the host (device 0); n, lo, hi: D entries; c: 4 cubic coefficients per axis
int kat_cartesian(int device, int req, int D, const int *n, const double *lo, const double *hi, double k,
                  double constant, const double *c, double *out)
{
  return guard([&] {
    cart_dispatch(D, [&](auto dc) {
      constexpr int DD = decltype(dc)::value;
      if (device) {
        auto I = CartOf<DD>::gpu(n, lo, hi);
        *out = cart_call<DD>(I, req != 0, k, constant, c);
      } else {
        auto I = CartOf<DD>::host(n, lo, hi);
        *out = cart_call<DD>(I, req != 0, k, constant, c);
      }
    });
  });
}

// Range changes on the GPU integrator built on (lo, hi):
//   out[0] a request made before the change, read after it;  out[1] a request after the change;
//   out[2] get() after the change;  out[3] get() of a new integrator on (lo2, hi2);
//   out[4] get() of a copy made before the change;  out[5] get() of a copy set to (lo2, hi2) with
//   set_q?_max before set_q?_min
int kat_cartesian_ranges(int D, const int *n, const double *lo, const double *hi, const double *lo2,
                         const double *hi2, double k, const double *c, double *out)
{
  return guard([&] {
    cart_dispatch(D, [&](auto dc) {
      constexpr int DD = decltype(dc)::value;
      auto I = CartOf<DD>::gpu(n, lo, hi);
      auto J = I, K = I;
      auto f0 = cart_request<DD>(I, k, 0.0, c);
      cart_set<DD>(I, lo2, hi2, false);
      auto f1 = cart_request<DD>(I, k, 0.0, c);
      out[2] = cart_call<DD>(I, false, k, 0.0, c);
      auto F = CartOf<DD>::gpu(n, lo2, hi2);
      out[3] = cart_call<DD>(F, false, k, 0.0, c);
      out[0] = f0.get();
      out[1] = f1.get();
      out[4] = cart_call<DD>(J, false, k, 0.0, c);
      cart_set<DD>(K, lo2, hi2, true);
      out[5] = cart_call<DD>(K, false, k, 0.0, c);
    });
  });
}

// The 1D integrator over number types nt = 0 double, 1 complex<double>, 2 dual: out = (first,
// second component) of get(), GPU (device 1) or host (device 0)
int kat_cartesian1_nt(int device, int nt, int n, double lo, double hi, double k, double constant, const double *c,
                      double *out)
{
  return guard([&] {
    auto run = [&](auto I) {
      const auto v = I.get(k, constant, c[0], c[1], c[2], c[3]);
      using V = std::decay_t<decltype(v)>;
      if constexpr (std::is_same_v<V, double>) {
        out[0] = v;
        out[1] = 0.0;
      } else if constexpr (std::is_same_v<V, Dual1>) {
        out[0] = v.v;
        out[1] = v.d;
      } else {
        out[0] = v.real();
        out[1] = v.imag();
      }
    };
    const std::array<uint, 1> g{{(uint)n}};
    switch (nt * 2 + device) {
    case 0: run(Integrator1DCartesianTBB<double, PolyCart1>(provider(), g, 0., 0, lo, hi)); break;
    case 1: run(Integrator1DCartesianGPU<double, PolyCart1>(provider(), g, 0., 256, lo, hi)); break;
    case 2: run(Integrator1DCartesianTBB<complex<double>, PolyCart1C>(provider(), g, 0., 0, lo, hi)); break;
    case 3: run(Integrator1DCartesianGPU<complex<double>, PolyCart1C>(provider(), g, 0., 256, lo, hi)); break;
    case 4: run(Integrator1DCartesianTBB<Dual1, PolyCart1D>(provider(), g, 0., 0, lo, hi)); break;
    case 5: run(Integrator1DCartesianGPU<Dual1, PolyCart1D>(provider(), g, 0., 256, lo, hi)); break;
    default: throw std::invalid_argument("unsupported number type");
    }
  });
}

// get() (req = 0) or request().get() (req = 1) of the zero-temperature forms on the GPU (device 1)
// or the host (device 0): form 0 Integrator3D {nx, na, nphi}, 1 Integrator2Dpq0 {nx, na, nb},
// 2 Integrator2Dpx0, 3 Integrator3Dpq0 {nx, na, nphi, nb}, 4 Integrator3Dpx0; ext is q0_extent or
// x0_extent; c: 6 q coefficients, a, p, z: 4 cos, phi and q0 coefficients
int kat_freq(int device, int form, int req, int nx, int na, int nphi, int nb, double x_ext, double ext, double k,
             const double *c, const double *a, const double *p, const double *z, double constant, double *out)
{
  return guard([&] {
    auto run3 = [&](auto &&I) {
      *out = req ? I.request(k, constant, c[0], c[1], c[2], c[3], c[4], c[5], a[0], a[1], a[2], a[3], p[0], p[1], p[2],
                             p[3])
                       .get()
                 : I.get(k, constant, c[0], c[1], c[2], c[3], c[4], c[5], a[0], a[1], a[2], a[3], p[0], p[1], p[2],
                         p[3]);
    };
    auto run2 = [&](auto &&I) {
      *out = req ? I.request(k, constant, c[0], c[1], c[2], c[3], c[4], c[5], a[0], a[1], a[2], a[3], z[0], z[1], z[2],
                             z[3])
                       .get()
                 : I.get(k, constant, c[0], c[1], c[2], c[3], c[4], c[5], a[0], a[1], a[2], a[3], z[0], z[1], z[2],
                         z[3]);
    };
    auto run4 = [&](auto &&I) {
      *out = req ? I.request(k, constant, c[0], c[1], c[2], c[3], c[4], c[5], a[0], a[1], a[2], a[3], p[0], p[1], p[2],
                             p[3], z[0], z[1], z[2], z[3])
                       .get()
                 : I.get(k, constant, c[0], c[1], c[2], c[3], c[4], c[5], a[0], a[1], a[2], a[3], p[0], p[1], p[2],
                         p[3], z[0], z[1], z[2], z[3]);
    };
    const std::array<uint, 3> g3{{(uint)nx, (uint)na, (uint)nphi}}, g2{{(uint)nx, (uint)na, (uint)nb}};
    const std::array<uint, 4> g4{{(uint)nx, (uint)na, (uint)nphi, (uint)nb}};
    switch (form * 2 + device) {
    case 0: run3(Integrator3DTBB<double, PolyFreq3D>(provider(), g3, x_ext)); break;
    case 1: run3(Integrator3DGPU<double, PolyFreq3D>(provider(), g3, x_ext)); break;
    case 2: run2(Integrator2Dpq0TBB<double, PolyFreq2Q0>(provider(), g2, x_ext, ext)); break;
    case 3: run2(Integrator2Dpq0GPU<double, PolyFreq2Q0>(provider(), g2, x_ext, ext)); break;
    case 4: run2(Integrator2Dpx0TBB<double, PolyFreq2Q0>(provider(), g2, x_ext, ext)); break;
    case 5: run2(Integrator2Dpx0GPU<double, PolyFreq2Q0>(provider(), g2, x_ext, ext)); break;
    case 6: run4(Integrator3Dpq0TBB<double, PolyFreq3Q0>(provider(), g4, x_ext, ext)); break;
    case 7: run4(Integrator3Dpq0GPU<double, PolyFreq3Q0>(provider(), g4, x_ext, ext)); break;
    case 8: run4(Integrator3Dpx0TBB<double, PolyFreq3Q0>(provider(), g4, x_ext, ext)); break;
    case 9: run4(Integrator3Dpx0GPU<double, PolyFreq3Q0>(provider(), g4, x_ext, ext)); break;
    default: throw std::invalid_argument("unsupported form");
    }
  });
}

// get() (req = 0) or request().get() (req = 1) of the finite-T x0 forms on the GPU (device 1) or
// the host (device 0): kind 0 IntegratorFiniteTx0<d> {nx, nint}, 1 IntegratorAngleFiniteTx0<d>
// {nx, na, nint}, 2 Integrator4DFiniteTx0 {nx, na, na, nint}; built at T_first, then set_T(T)
// when T_first differs; c: 6 q coefficients, a, p, z: 4 cos, phi and q0 coefficients
int kat_finiteTx0(int device, int kind, int d, int req, int nx, int na, int nint, int summands, double x_ext,
                  double x0_ext, double T_first, double T, double k, const double *c, const double *a, const double *p,
                  const double *z, double constant, double *out)
{
  return guard([&] {
    auto go1 = [&](auto &&I) {
      if (T_first != T) I.set_T(T);
      *out = req ? I.request(k, constant, c[0], c[1], c[2], c[3], c[4], c[5], z[0], z[1], z[2], z[3]).get()
                 : I.get(k, constant, c[0], c[1], c[2], c[3], c[4], c[5], z[0], z[1], z[2], z[3]);
    };
    auto go2 = [&](auto &&I) {
      if (T_first != T) I.set_T(T);
      *out = req ? I.request(k, constant, c[0], c[1], c[2], c[3], c[4], c[5], a[0], a[1], a[2], a[3], z[0], z[1], z[2],
                             z[3])
                       .get()
                 : I.get(k, constant, c[0], c[1], c[2], c[3], c[4], c[5], a[0], a[1], a[2], a[3], z[0], z[1], z[2],
                         z[3]);
    };
    auto go4 = [&](auto &&I) {
      if (T_first != T) I.set_T(T);
      *out = req ? I.request(k, constant, c[0], c[1], c[2], c[3], c[4], c[5], a[0], a[1], a[2], a[3], p[0], p[1], p[2],
                             p[3], z[0], z[1], z[2], z[3])
                       .get()
                 : I.get(k, constant, c[0], c[1], c[2], c[3], c[4], c[5], a[0], a[1], a[2], a[3], p[0], p[1], p[2],
                         p[3], z[0], z[1], z[2], z[3]);
    };
    const unsigned S = (unsigned)summands;
    const std::array<uint, 2> g1{{(uint)nx, (uint)nint}};
    const std::array<uint, 3> g2{{(uint)nx, (uint)na, (uint)nint}};
    const std::array<uint, 4> g4{{(uint)nx, (uint)na, (uint)na, (uint)nint}};
    const int key = kind * 100 + d * 2 + device;
    switch (key) {
    case 4: go1(IntegratorFiniteTx0TBB<2, double, PolyTx0>(provider(), g1, x_ext, x0_ext, S, T_first)); break;
    case 5: go1(IntegratorFiniteTx0GPU<2, double, PolyTx0>(provider(), g1, x_ext, x0_ext, S, T_first)); break;
    case 6: go1(IntegratorFiniteTx0TBB<3, double, PolyTx0>(provider(), g1, x_ext, x0_ext, S, T_first)); break;
    case 7: go1(IntegratorFiniteTx0GPU<3, double, PolyTx0>(provider(), g1, x_ext, x0_ext, S, T_first)); break;
    case 8: go1(IntegratorFiniteTx0TBB<4, double, PolyTx0>(provider(), g1, x_ext, x0_ext, S, T_first)); break;
    case 9: go1(IntegratorFiniteTx0GPU<4, double, PolyTx0>(provider(), g1, x_ext, x0_ext, S, T_first)); break;
    case 106: go2(IntegratorAngleFiniteTx0TBB<3, double, PolyTx0Angle>(provider(), g2, x_ext, x0_ext, S, T_first)); break;
    case 107: go2(IntegratorAngleFiniteTx0GPU<3, double, PolyTx0Angle>(provider(), g2, x_ext, x0_ext, S, T_first)); break;
    case 108: go2(IntegratorAngleFiniteTx0TBB<4, double, PolyTx0Angle>(provider(), g2, x_ext, x0_ext, S, T_first)); break;
    case 109: go2(IntegratorAngleFiniteTx0GPU<4, double, PolyTx0Angle>(provider(), g2, x_ext, x0_ext, S, T_first)); break;
    case 208: go4(Integrator4DFiniteTx0TBB<double, PolyFreq3Q0>(provider(), g4, x_ext, x0_ext, S, T_first)); break;
    case 209: go4(Integrator4DFiniteTx0GPU<double, PolyFreq3Q0>(provider(), g4, x_ext, x0_ext, S, T_first)); break;
    default: throw std::invalid_argument("unsupported form");
    }
  });
}

// IntegratorConstant<d> (integrator_constant.cc): out[0] = get(), out[1] = request().get()
int kat_constant(int d, double k, double constant, double x0, double *out)
{
  return guard([&] {
    auto run = [&](auto I) {
      out[0] = I.get(k, constant, x0);
      out[1] = I.request(k, constant, x0).get();
    };
    switch (d) {
    case 1: run(IntegratorConstant<1, double, PolyConstant>(provider(), 64, 1.)); break;
    case 2: run(IntegratorConstant<2, double, PolyConstant>(provider(), 64, 1.)); break;
    case 3: run(IntegratorConstant<3, double, PolyConstant>(provider(), 64, 1.)); break;
    case 4: run(IntegratorConstant<4, double, PolyConstant>(provider(), {{64}}, 1.)); break;
    default: throw std::invalid_argument("unsupported dimension");
    }
  });
}

// request() launch mode (dfrg::tmpl::set_launch_at_request): 1 launches every request at once
int kat_set_launch_at_request(int on)
{
  dfrg::tmpl::set_launch_at_request(on != 0);
  return dfrg::tmpl::launch_at_request() ? 1 : 0;
}

// *nm receives the Matsubara rule size (GPU template only)
int kat_4d_finiteT(int device, int nx, int na, double x_extent, double T, double k, const double *c, const double *a,
                   const double *p, const double *den, double constant, double *out, int *nm)
{
  return guard([&] {
    auto run = [&](auto &integ) {
      *out = integ.get(k, constant, c[0], c[1], c[2], c[3], c[4], c[5], a[0], a[1], a[2], a[3], p[0], p[1], p[2],
                       p[3], den[0], den[1], den[2], den[3], den[4]);
    };
    if (device) {
      Integrator4DFiniteTq0GPU<double, Poly4DFiniteT> i(provider(), {{(uint)nx, (uint)na, (uint)na}}, x_extent, T);
      run(i);
      *nm = (int)i.matsubara_size();
    } else {
      Integrator4DFiniteTq0TBB<double, Poly4DFiniteT> i(provider(), {{(uint)nx, (uint)na, (uint)na, 0u}}, x_extent, T);
      run(i);
    }
  });
}

int kat_angle_finiteT(int device, int d, int nx, int na, double x_extent, double T, double k, const double *c,
                      const double *a, const double *den, double constant, double *out, int *nm)
{
  return guard([&] {
    auto run = [&](auto &integ) {
      *out = integ.get(k, constant, c[0], c[1], c[2], c[3], c[4], c[5], a[0], a[1], a[2], a[3], den[0], den[1], den[2],
                       den[3]);
      if constexpr (requires { integ.matsubara_size(); }) *nm = (int)integ.matsubara_size();
    };
    const std::array<uint, 2> g{{(uint)nx, (uint)na}};
    switch (d * 2 + device) {
    case 6: { IntegratorAngleFiniteTq0TBB<3, double, PolyAngleFiniteT> i(provider(), g, x_extent, T); run(i); break; }
    case 7: { IntegratorAngleFiniteTq0GPU<3, double, PolyAngleFiniteT> i(provider(), g, x_extent, T); run(i); break; }
    case 8: { IntegratorAngleFiniteTq0TBB<4, double, PolyAngleFiniteT> i(provider(), g, x_extent, T); run(i); break; }
    case 9: { IntegratorAngleFiniteTq0GPU<4, double, PolyAngleFiniteT> i(provider(), g, x_extent, T); run(i); break; }
    default: throw std::invalid_argument("unsupported dimension");
    }
  });
}

// n requests over ks (one batch per Matsubara rule for the finite-T forms) against get() on a
// second integrator at the same k; which = 0: angle d=4, 1: 4D finite-T, 2: angle finite-T d=4
int kat_variant_requests(int which, int nx, int na, double x_extent, double T, const double *ks, int n,
                         double *out_req, double *out_get)
{
  return guard([&] {
    auto both = [&](auto &ra, auto &gb, auto &&call_req, auto &&call_get) {
      std::vector<std::future<double>> f;
      for (int i = 0; i < n; ++i) f.push_back(call_req(ra, ks[i]));
      for (int i = 0; i < n; ++i) out_get[i] = call_get(gb, ks[i]);
      for (int i = 0; i < n; ++i) out_req[i] = f[i].get();
    };
    const double c0 = 1.0, c1 = 0.3, c2 = -0.2, a0 = 1.0, a1 = 0.25, a2 = 0.1;
    if (which == 0) {
      IntegratorAngleGPU<4, double, PolyAngle> A(provider(), {{(uint)nx, (uint)na}}, x_extent);
      IntegratorAngleGPU<4, double, PolyAngle> B(provider(), {{(uint)nx, (uint)na}}, x_extent);
      both(A, B, [&](auto &I, double k) { return I.request(k, 0.0, c0, c1, c2, 0., 0., 0., a0, a1, a2, 0.); },
           [&](auto &I, double k) { return I.get(k, 0.0, c0, c1, c2, 0., 0., 0., a0, a1, a2, 0.); });
    } else if (which == 1) {
      Integrator4DFiniteTq0GPU<double, Poly4DFiniteT> A(provider(), {{(uint)nx, (uint)na, (uint)na}}, x_extent, T);
      Integrator4DFiniteTq0GPU<double, Poly4DFiniteT> B(provider(), {{(uint)nx, (uint)na, (uint)na}}, x_extent, T);
      both(A, B,
           [&](auto &I, double k) {
             return I.request(k, 0.0, c0, c1, c2, 0., 0., 0., a0, a1, a2, 0., 1., 0.05, 0., 0., T * T, 0., 1.1, 0.,
                              0.);
           },
           [&](auto &I, double k) {
             return I.get(k, 0.0, c0, c1, c2, 0., 0., 0., a0, a1, a2, 0., 1., 0.05, 0., 0., T * T, 0., 1.1, 0., 0.);
           });
    } else if (which == 2) {
      IntegratorAngleFiniteTq0GPU<4, double, PolyAngleFiniteT> A(provider(), {{(uint)nx, (uint)na}}, x_extent, T);
      IntegratorAngleFiniteTq0GPU<4, double, PolyAngleFiniteT> B(provider(), {{(uint)nx, (uint)na}}, x_extent, T);
      both(A, B,
           [&](auto &I, double k) { return I.request(k, 0.0, c0, c1, c2, 0., 0., 0., a0, a1, a2, 0., T * T, 0., 1.1, 0.); },
           [&](auto &I, double k) { return I.get(k, 0.0, c0, c1, c2, 0., 0., 0., a0, a1, a2, 0., T * T, 0., 1.1, 0.); });
    } else
      throw std::invalid_argument("unknown variant");
  });
}

// ---------------------------------------------------------------------------------------------
// interpolators: one evaluation per point inside a kernel (HBM table through the by-value copy)
// and on the host
// ---------------------------------------------------------------------------------------------
} // extern "C"

namespace
{
  template <int DIM, class I>
  __global__ void interp_eval_kernel(I interp, const double *pts, int n, double *out)
  {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *x = pts + (size_t)i * DIM;
    if constexpr (DIM == 1)
      out[i] = double(interp(x[0]));
    else if constexpr (DIM == 2)
      out[i] = double(interp(x[0], x[1]));
    else
      out[i] = double(interp(x[0], x[1], x[2]));
  }

  template <int DIM, class I> void interp_eval(const I &interp, const double *pts, int n, double *out_dev, double *out_host)
  {
    void *d_pts = nullptr, *d_out = nullptr;
    dfrg::check(dfrg_malloc(&d_pts, sizeof(double) * DIM * n), "pts");
    dfrg::check(dfrg_malloc(&d_out, sizeof(double) * n), "out");
    // the library's copies run on this thread's stream: launch there too
    void *s = nullptr;
    dfrg::check(dfrg_thread_stream(&s), "thread stream");
    dfrg::check(dfrg_memcpy_h2d(d_pts, pts, sizeof(double) * DIM * n, s), "pts");
    hipLaunchKernelGGL((interp_eval_kernel<DIM, I>), dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)s, interp,
                       (const double *)d_pts, n, (double *)d_out);
    dfrg::check_hip("interp_eval_kernel");
    dfrg::check(dfrg_memcpy_d2h(out_dev, d_out, sizeof(double) * n, s), "out");
    dfrg::check(dfrg_stream_sync(s), "out");
    dfrg_free(d_pts);
    dfrg_free(d_out);
    for (int i = 0; i < n; ++i) {
      const double *x = pts + (size_t)i * DIM;
      if constexpr (DIM == 1)
        out_host[i] = double(interp(x[0]));
      else if constexpr (DIM == 2)
        out_host[i] = double(interp(x[0], x[1]));
      else
        out_host[i] = double(interp(x[0], x[1], x[2]));
    }
  }
} // namespace

extern "C" {

// kind: 0 LinearInterpolator2D<double>, 1 LinearInterpolator3D<double>, 2 TexLinearInterpolator1D
// (log coordinates), 3 TexLinearInterpolator2D, 4 TexLinearInterpolator3D, 5 TexLinearInterpolator1DStack
// (bosonic), 6 TexLinearInterpolator1DStack (fermionic), 7 LinearInterpolator2D<float>,
// 8 LinearInterpolator1D<double> (FP64 log coordinates).
// cp: linear axes (n, start, stop) per axis; log (n, start, stop, bias); stacks (m_start, m_stop, T,
// n, start, stop, bias).  pts: n_pts points of the interpolator's dimension.
int kat_interp(int kind, const double *cp, const double *data, int n_data, const double *pts, int n_pts,
               double *out_dev, double *out_host)
{
  return guard([&] {
    using LD = LinearCoordinates1D<double>;
    using LF = LinearCoordinates1D<float>;
    auto need = [&](unsigned n) {
      if ((unsigned)n_data != n) throw std::invalid_argument("data size does not match the coordinates");
    };
    switch (kind) {
    case 0: {
      CoordinatePackND<LD, LD> c(LD(cp[0], cp[1], cp[2]), LD(cp[3], cp[4], cp[5]));
      need(c.size());
      LinearInterpolator2D<double, decltype(c)> I(c);
      I.update(data);
      interp_eval<2>(I, pts, n_pts, out_dev, out_host);
      break;
    }
    case 1: {
      CoordinatePackND<LD, LD, LD> c(LD(cp[0], cp[1], cp[2]), LD(cp[3], cp[4], cp[5]), LD(cp[6], cp[7], cp[8]));
      need(c.size());
      LinearInterpolator3D<double, decltype(c)> I(std::vector<double>(data, data + n_data), c);
      interp_eval<3>(I, pts, n_pts, out_dev, out_host);
      break;
    }
    case 2: {
      LogarithmicCoordinates1D<float> c((unsigned)cp[0], cp[1], cp[2], cp[3]);
      need(c.size());
      TexLinearInterpolator1D<float, decltype(c)> I(c);
      I.update(data);
      interp_eval<1>(I, pts, n_pts, out_dev, out_host);
      break;
    }
    case 3: {
      CoordinatePackND<LF, LF> c(LF(cp[0], cp[1], cp[2]), LF(cp[3], cp[4], cp[5]));
      need(c.size());
      TexLinearInterpolator2D<double, decltype(c)> I(data, c);
      interp_eval<2>(I, pts, n_pts, out_dev, out_host);
      break;
    }
    case 4: {
      CoordinatePackND<LF, LF, LF> c(LF(cp[0], cp[1], cp[2]), LF(cp[3], cp[4], cp[5]), LF(cp[6], cp[7], cp[8]));
      need(c.size());
      TexLinearInterpolator3D<float, decltype(c)> I(c);
      I.update(data);
      interp_eval<3>(I, pts, n_pts, out_dev, out_host);
      break;
    }
    case 5: {
      BosonicCoordinates1DFiniteT<int, float> c((int)cp[0], (int)cp[1], cp[2], (unsigned)cp[3], cp[4], cp[5], cp[6]);
      need(c.size());
      TexLinearInterpolator1DStack<float, decltype(c)> I(c);
      I.update(data);
      interp_eval<2>(I, pts, n_pts, out_dev, out_host);
      break;
    }
    case 6: {
      FermionicCoordinates1DFiniteT<int, float> c((int)cp[0], (int)cp[1], cp[2], (unsigned)cp[3], cp[4], cp[5], cp[6]);
      need(c.size());
      TexLinearInterpolator1DStack<float, decltype(c)> I(c);
      I.update(data);
      interp_eval<2>(I, pts, n_pts, out_dev, out_host);
      break;
    }
    case 7: {
      CoordinatePackND<LF, LF> c(LF(cp[0], cp[1], cp[2]), LF(cp[3], cp[4], cp[5]));
      need(c.size());
      LinearInterpolator2D<float, decltype(c)> I(std::vector<float>(c.size(), 0.f), c);
      I.update(data);
      interp_eval<2>(I, pts, n_pts, out_dev, out_host);
      break;
    }
    case 8: {
      LogarithmicCoordinates1D<double> c((unsigned)cp[0], cp[1], cp[2], cp[3]);
      need(c.size());
      LinearInterpolator1D<double, decltype(c)> I(data, c);
      interp_eval<1>(I, pts, n_pts, out_dev, out_host);
      break;
    }
    default: throw std::invalid_argument("unknown interpolator kind");
    }
  });
}

} // extern "C"


namespace
{
  template <int DIM, class I> __global__ void interp_dual_kernel(I interp, const double *pts, int n, double *out)
  {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *x = pts + (size_t)i * DIM;
    Dual1 r;
    if constexpr (DIM == 1)
      r = interp(x[0]);
    else if constexpr (DIM == 2)
      r = interp(x[0], x[1]);
    else
      r = interp(x[0], x[1], x[2]);
    out[2 * i] = r.v;
    out[2 * i + 1] = r.d;
  }

  template <int DIM, class I>
  void interp_dual_eval(const I &interp, const double *pts, int n, double *out_dev, double *out_host)
  {
    void *d_pts = nullptr, *d_out = nullptr;
    dfrg::check(dfrg_malloc(&d_pts, sizeof(double) * DIM * n), "pts");
    dfrg::check(dfrg_malloc(&d_out, sizeof(double) * 2 * n), "out");
    void *s = nullptr;
    dfrg::check(dfrg_thread_stream(&s), "thread stream");
    dfrg::check(dfrg_memcpy_h2d(d_pts, pts, sizeof(double) * DIM * n, s), "pts");
    hipLaunchKernelGGL((interp_dual_kernel<DIM, I>), dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)s, interp,
                       (const double *)d_pts, n, (double *)d_out);
    dfrg::check_hip("interp_dual_kernel");
    dfrg::check(dfrg_memcpy_d2h(out_dev, d_out, sizeof(double) * 2 * n, s), "out");
    dfrg::check(dfrg_stream_sync(s), "out");
    dfrg_free(d_pts);
    dfrg_free(d_out);
    for (int i = 0; i < n; ++i) {
      const double *x = pts + (size_t)i * DIM;
      Dual1 r;
      if constexpr (DIM == 1)
        r = interp(x[0]);
      else if constexpr (DIM == 2)
        r = interp(x[0], x[1]);
      else
        r = interp(x[0], x[1], x[2]);
      out_host[2 * i] = r.v;
      out_host[2 * i + 1] = r.d;
    }
  }
} // namespace

extern "C" {

// Dual-number data (value val[i], derivative der[i]) in the interpolators; out_dev / out_host get
// (value, derivative) per point.  kind: 0 TexLinearInterpolator1D (log coordinates), 1
// TexLinearInterpolator2D, 2 TexLinearInterpolator3D, 3 TexLinearInterpolator1DStack (bosonic),
// 4 LinearInterpolator1D<Dual1> (log coordinates, FP64), 5 TexLinearInterpolator1D built from the
// dual data and then updated with plain values 2 val[i] (the derivative table is kept).  cp as in
// kat_interp for the same coordinates.
int kat_interp_dual(int kind, const double *cp, const double *val, const double *der, int n_data, const double *pts,
                    int n_pts, double *out_dev, double *out_host)
{
  return guard([&] {
    using LF = LinearCoordinates1D<float>;
    std::vector<Dual1> data(n_data);
    for (int i = 0; i < n_data; ++i) data[i] = Dual1(val[i], der[i]);
    auto need = [&](unsigned n) {
      if ((unsigned)n_data != n) throw std::invalid_argument("data size does not match the coordinates");
    };
    switch (kind) {
    case 0:
    case 5: {
      LogarithmicCoordinates1D<float> c((unsigned)cp[0], cp[1], cp[2], cp[3]);
      need(c.size());
      TexLinearInterpolator1D<Dual1, decltype(c)> I(data, c);
      if (kind == 5) {
        std::vector<double> twice(val, val + n_data);
        for (double &v : twice) v *= 2.0;
        I.update(twice.data());
      }
      interp_dual_eval<1>(I, pts, n_pts, out_dev, out_host);
      break;
    }
    case 1: {
      CoordinatePackND<LF, LF> c(LF(cp[0], cp[1], cp[2]), LF(cp[3], cp[4], cp[5]));
      need(c.size());
      TexLinearInterpolator2D<Dual1, decltype(c)> I(data.data(), c);
      interp_dual_eval<2>(I, pts, n_pts, out_dev, out_host);
      break;
    }
    case 2: {
      CoordinatePackND<LF, LF, LF> c(LF(cp[0], cp[1], cp[2]), LF(cp[3], cp[4], cp[5]), LF(cp[6], cp[7], cp[8]));
      need(c.size());
      TexLinearInterpolator3D<Dual1, decltype(c)> I(c);
      I.update(data.data());
      interp_dual_eval<3>(I, pts, n_pts, out_dev, out_host);
      break;
    }
    case 3: {
      BosonicCoordinates1DFiniteT<int, float> c((int)cp[0], (int)cp[1], cp[2], (unsigned)cp[3], cp[4], cp[5], cp[6]);
      need(c.size());
      TexLinearInterpolator1DStack<Dual1, decltype(c)> I(data, c);
      interp_dual_eval<2>(I, pts, n_pts, out_dev, out_host);
      break;
    }
    case 4: {
      LogarithmicCoordinates1D<double> c((unsigned)cp[0], cp[1], cp[2], cp[3]);
      need(c.size());
      LinearInterpolator1D<Dual1, decltype(c)> I(data, c);
      interp_dual_eval<1>(I, pts, n_pts, out_dev, out_host);
      break;
    }
    default: throw std::invalid_argument("unknown interpolator kind");
    }
  });
}

// operator[] of a Tex interpolator over dual data throws, as the reference's does
int kat_interp_dual_index(void)
{
  return guard([&] {
    LogarithmicCoordinates1D<float> c(8, 1e-3f, 10.f, 2.f);
    TexLinearInterpolator1D<Dual1, decltype(c)> I(c);
    (void)I[0];
  });
}

// views: update() on a copy throws; stack coordinates reject start > stop
int kat_interp_errors(int which)
{
  return guard([&] {
    using LD = LinearCoordinates1D<double>;
    if (which == 0) {
      CoordinatePackND<LD, LD> c(LD(4, 0., 1.), LD(5, 0., 1.));
      LinearInterpolator2D<double, decltype(c)> I(c);
      auto view = I;
      view.update();
    } else if (which == 1) {
      BosonicMatsubaraValues<int, double> m(3, 1, 0.1);
    } else {
      CoordinatePackND<LD, LD> c(LD(1, 0., 1.), LD(5, 0., 1.));
      LinearInterpolator2D<double, decltype(c)> I(c);
    }
  });
}

// ---------------------------------------------------------------------------------------------
// time stepping of variables systems and the x_extent choice
// ---------------------------------------------------------------------------------------------
} // extern "C"

namespace
{
  // y' = A y for a 2x2 matrix per system (A row-major in a device array), plus a decaying
  // integral term c_s * int d^4q/(2pi)^4 RBdot(q^2)/(q^2 + RB(q^2) + y0^2)^2 on GL nodes when nx > 0
  struct LinearRHS {
    const double *A; // n_systems x 4
    const double *xq, *wq;
    int nx;
    double c;
    __device__ void operator()(double, const double *y, double *f, int s) const
    {
      const double *a = A + 4 * s;
      f[0] = a[0] * y[0] + a[1] * y[1];
      f[1] = a[2] * y[0] + a[3] * y[1];
      if (nx > 0) {
        double I = 0.0;
        for (int i = 0; i < nx; ++i) {
          const double q2 = xq[i] * 20.0;
          I += wq[i] * 20.0 * (0.5 * q2) * Reg::RBdot(1.0, q2) / powr<2>(q2 + Reg::RB(1.0, q2) + y[0] * y[0]);
        }
        f[0] += c * I * (2.0 * M_PI * M_PI / powr<4>(2.0 * M_PI));
      }
    }
  };
} // namespace

extern "C" {

// host stepper: 2-variable linear systems (no integral term); y: n x 2 in/out; snapshots of the
// first system: n_out x 2
int kat_rk_host(int method, const double *A, double *y, int n, double t0, double t1, double dt, double output_dt,
                double *snap, int n_snap, int *n_written)
{
  return guard([&] {
    int w = 0;
    for (int s = 0; s < n; ++s) {
      std::vector<double> v{y[2 * s], y[2 * s + 1]};
      const double *a = A + 4 * s;
      dfrg::tmpl::run_variables(
          v, t0, t1, dt, output_dt, method,
          [&](double, const std::vector<double> &Y, std::vector<double> &f) {
            f[0] = a[0] * Y[0] + a[1] * Y[1];
            f[1] = a[2] * Y[0] + a[3] * Y[1];
          },
          [&](double, const std::vector<double> &Y) {
            if (s == 0 && w < n_snap) {
              snap[2 * w] = Y[0];
              snap[2 * w + 1] = Y[1];
              ++w;
            }
          });
      y[2 * s] = v[0];
      y[2 * s + 1] = v[1];
    }
    *n_written = w;
  });
}

// device ensemble of the same systems; with nx > 0 the RHS adds the GL threshold integral
// (c * I(y0)); snap: n_snap x n x 2; status per system
int kat_rk_device(int method, const double *A, double *y, int n, double t0, double t1, double dt, double output_dt,
                  int nx, double c, double *snap, int n_snap, int *status)
{
  return guard([&] {
    void *s = nullptr;
    dfrg::check(dfrg_thread_stream(&s), "thread stream");
    void *dA, *dy, *dsnap = nullptr, *dst;
    dfrg::check(dfrg_malloc(&dA, sizeof(double) * 4 * n), "A");
    dfrg::check(dfrg_malloc(&dy, sizeof(double) * 2 * n), "y");
    dfrg::check(dfrg_malloc(&dst, sizeof(int) * n), "status");
    if (n_snap > 0) dfrg::check(dfrg_malloc(&dsnap, sizeof(double) * 2 * n * n_snap), "snap");
    dfrg::check(dfrg_memcpy_h2d(dA, A, sizeof(double) * 4 * n, s), "A");
    dfrg::check(dfrg_memcpy_h2d(dy, y, sizeof(double) * 2 * n, s), "y");
    LinearRHS rhs{(const double *)dA, nullptr, nullptr, 0, c};
    if (nx > 0) {
      rhs.xq = provider().get_device_points<double>(nx);
      rhs.wq = provider().get_device_weights<double>(nx);
      rhs.nx = nx;
    }
    dfrg::tmpl::rk_ensemble<2>(rhs, method, (double *)dy, n, t0, t1, dt, output_dt, (double *)dsnap, n_snap,
                               (int *)dst, (hipStream_t)s);
    dfrg::check(dfrg_memcpy_d2h(y, dy, sizeof(double) * 2 * n, s), "y");
    dfrg::check(dfrg_memcpy_d2h(status, dst, sizeof(int) * n, s), "status");
    if (n_snap > 0) dfrg::check(dfrg_memcpy_d2h(snap, dsnap, sizeof(double) * 2 * n * n_snap, s), "snap");
    dfrg::check(dfrg_stream_sync(s), "sync");
    dfrg_free(dA);
    dfrg_free(dy);
    dfrg_free(dst);
    if (dsnap) dfrg_free(dsnap);
  });
}

int kat_rk_snapshots(double t0, double t1, double dt, double output_dt, int *n)
{
  return guard([&] { *n = dfrg::tmpl::rk_snapshots(t0, t1, dt, output_dt); });
}

// x_extent for f(q^2) = exp(-a q^2) (which = 0), f = 0 (which = 1), or a slowly decaying
// 1/(1 + q^2)^3 (which = 2)
int kat_x_extent(int which, double a, unsigned order, double tol, double *x_extent)
{
  return guard([&] {
    std::function<double(double)> f;
    if (which == 0)
      f = [a](double q2) { return std::exp(-a * q2); };
    else if (which == 1)
      f = [](double) { return 0.0; };
    else
      f = [](double q2) { return 1.0 / powr<3>(1.0 + q2); };
    *x_extent = dfrg::tmpl::optimize_x_extent(provider(), f, order, tol);
  });
}

// NT = dual number: the radial integrator over (f, g) polynomial pairs, get and a batched request;
// out = (get value, get derivative, request value, request derivative)
int kat_radial_dual(int nx, double x_extent, double k, const double *cf, const double *cg, int tbb, double *out)
{
  return guard([&] {
    auto run = [&](auto &&integ) {
      const Dual1 a = integ.get(k, cf[0], cf[1], cf[2], cf[3], cf[4], cf[5], cg[0], cg[1], cg[2], cg[3], cg[4], cg[5]);
      const Dual1 b =
          integ.request(k, cf[0], cf[1], cf[2], cf[3], cf[4], cf[5], cg[0], cg[1], cg[2], cg[3], cg[4], cg[5]).get();
      out[0] = a.v;
      out[1] = a.d;
      out[2] = b.v;
      out[3] = b.d;
    };
    if (tbb)
      run(IntegratorTBB<4, Dual1, PolyRadialDual>(provider(), nx, x_extent));
    else
      run(IntegratorGPU<4, Dual1, PolyRadialDual>(provider(), nx, x_extent));
  });
}

// NT = float: the radial integrator on FP32 nodes (get and a batched request)
int kat_radial_float(int nx, float x_extent, float k, const double *c, float constant, double *out_get,
                     double *out_req)
{
  return guard([&] {
    IntegratorGPU<4, float, PolyRadialF> integ(provider(), nx, x_extent);
    const float cf[6] = {(float)c[0], (float)c[1], (float)c[2], (float)c[3], (float)c[4], (float)c[5]};
    *out_get = integ.get(k, constant, cf[0], cf[1], cf[2], cf[3], cf[4], cf[5]);
    *out_req = integ.request(k, constant, cf[0], cf[1], cf[2], cf[3], cf[4], cf[5]).get();
  });
}

} // extern "C"

// ---------------------------------------------------------------------------------------------
// regulators and thermodynamic helpers (DiFfRG/physics/regulators.hh, thermodynamics.hh),
// evaluated on the device and on the host, and regulators inside a user KERNEL
// ---------------------------------------------------------------------------------------------
namespace
{
  struct BosB1 {
    static constexpr double b = 1.;
  };
  struct PolyOrder1 {
    static constexpr int order = 1;
  };
  struct PolyOrder60 {
    static constexpr int order = 60;
  };
  struct SmoothAlpha1e4 {
    static constexpr double alpha = 1e-4;
  };

  // RB, RBdot, dq2RB, RF, RFdot, dq2RF (NaN where the regulator does not define the function)
  template <class R> __host__ __device__ void reg_eval(double k2, double q2, double *o)
  {
    const double nan = __builtin_nan("");
    o[0] = R::RB(k2, q2);
    o[1] = R::RBdot(k2, q2);
    if constexpr (requires { R::dq2RB(k2, q2) + 0.; })
      o[2] = R::dq2RB(k2, q2);
    else
      o[2] = nan;
    o[3] = R::RF(k2, q2);
    o[4] = R::RFdot(k2, q2);
    if constexpr (requires { R::dq2RF(k2, q2) + 0.; })
      o[5] = R::dq2RF(k2, q2);
    else
      o[5] = nan;
  }
  template <class R> __global__ void reg_kernel(const double *k2, const double *q2, int n, double *out)
  {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) reg_eval<R>(k2[i], q2[i], out + 6 * i);
  }
  template <class R> void reg_run(int n, const double *k2, const double *q2, double *out_dev, double *out_host)
  {
    for (int i = 0; i < n; ++i) reg_eval<R>(k2[i], q2[i], out_host + 6 * i);
    void *d = nullptr, *s = nullptr;
    dfrg::check(dfrg_malloc(&d, sizeof(double) * 8 * n), "reg buffers");
    double *dk = (double *)d, *dq = dk + n, *dout = dq + n;
    dfrg::check(dfrg_thread_stream(&s), "thread stream");
    dfrg::check(dfrg_memcpy_h2d(dk, k2, sizeof(double) * n, s), "k2");
    dfrg::check(dfrg_memcpy_h2d(dq, q2, sizeof(double) * n, s), "q2");
    hipLaunchKernelGGL(reg_kernel<R>, dim3((n + 127) / 128), dim3(128), 0, (hipStream_t)s, dk, dq, n, dout);
    dfrg::check_hip("reg_kernel");
    dfrg::check(dfrg_memcpy_d2h(out_dev, dout, sizeof(double) * 6 * n, s), "out");
    dfrg::check(dfrg_stream_sync(s), "out");
    dfrg_free(d);
  }

  // the LPA threshold integrand of a regulator, RBdot / (q^2 + RB + m2)^2
  template <class R> struct ThresholdOf {
    static __forceinline__ __host__ __device__ double kernel(const double q, const double k, const double m2)
    {
      const double k2 = powr<2>(k), q2 = powr<2>(q);
      return R::RBdot(k2, q2) / powr<2>(q2 + R::RB(k2, q2) + m2);
    }
    static __forceinline__ __host__ __device__ double constant(const double, const double) { return 0.; }
  };
  template <class R> void reg_threshold(int nx, double x_extent, double k, const double *m2, int n, double *gpu,
                                        double *host)
  {
    IntegratorGPU<4, double, ThresholdOf<R>> g(provider(), {{(unsigned)nx}}, x_extent);
    IntegratorTBB<4, double, ThresholdOf<R>> h(provider(), {{(unsigned)nx}}, x_extent);
    std::vector<std::future<double>> f;
    for (int i = 0; i < n; ++i) f.push_back(g.request(k, m2[i]));
    for (int i = 0; i < n; ++i) {
      gpu[i] = f[i].get();
      host[i] = h.get(k, m2[i]);
    }
  }

  template <class F> int reg_dispatch(int kind, F &&f)
  {
    switch (kind) {
    case 0: f.template operator()<LitimRegulator<>>(); return 0;
    case 1: f.template operator()<BosonicRegulator<>>(); return 0;
    case 2: f.template operator()<BosonicRegulator<BosB1>>(); return 0;
    case 3: f.template operator()<ExponentialRegulator<B1>>(); return 0;
    case 4: f.template operator()<ExponentialRegulator<>>(); return 0;
    case 5: f.template operator()<SmoothedLitimRegulator<>>(); return 0;
    case 6: f.template operator()<SmoothedLitimRegulator<SmoothAlpha1e4>>(); return 0;
    case 7: f.template operator()<PolynomialExpRegulator<>>(); return 0;
    case 8: f.template operator()<PolynomialExpRegulator<PolyOrder1>>(); return 0;
    case 9: f.template operator()<PolynomialExpRegulator<PolyOrder60>>(); return 0;
    }
    throw std::invalid_argument("unknown regulator kind");
  }

  constexpr int kThermo = 18;
  __host__ __device__ void thermo_eval(double x, double T, double *o)
  {
    o[0] = CothFiniteT(x, T);
    o[1] = TanhFiniteT(x, T);
    o[2] = SechFiniteT(x, T);
    o[3] = CschFiniteT(x, T);
    o[4] = cothS(x, T);
    o[5] = dcothS(x, T);
    o[6] = ddcothS(x, T);
    o[7] = dddcothS(x, T);
    o[8] = tanhS(x, T);
    o[9] = dtanhS(x, T);
    o[10] = ddtanhS(x, T);
    o[11] = sechS(x, T);
    o[12] = cschS(x, T);
    o[13] = nB(x, T);
    o[14] = dnB(x, T);
    o[15] = ddnB(x, T);
    o[16] = nF(x, T);
    o[17] = dnF(x, T);
  }
  __global__ void thermo_kernel(const double *x, const double *T, int n, double *out)
  {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) thermo_eval(x[i], T[i], out + kThermo * i);
  }

  // the bosonic finite-T loop kernel of the reference's O(N) example in its own form
  // (Examples/ONfiniteT/flows/V/V.kernel:4-15 with a single field): CothFiniteT(E, T)/E RBdot
  template <class R> struct CothKernel {
    static __forceinline__ __host__ __device__ double kernel(const double q, const double k, const double T,
                                                             const double m2)
    {
      const double k2 = powr<2>(k), q2 = powr<2>(q);
      const double E = sqrt(q2 + R::RB(k2, q2) + m2);
      return CothFiniteT(E, T) / E * R::RBdot(k2, q2);
    }
    static __forceinline__ __host__ __device__ double constant(const double, const double, const double)
    {
      return 0.;
    }
  };
} // namespace

extern "C" {

int kat_num_thermo(void) { return kThermo; }

int kat_regulator_eval(int kind, int n, const double *k2, const double *q2, double *out_dev, double *out_host)
{
  return guard([&] {
    reg_dispatch(kind, [&]<class R>() { reg_run<R>(n, k2, q2, out_dev, out_host); });
  });
}

int kat_regulator_threshold(int kind, int nx, double x_extent, double k, const double *m2, int n, double *gpu,
                            double *host)
{
  return guard([&] {
    reg_dispatch(kind, [&]<class R>() { reg_threshold<R>(nx, x_extent, k, m2, n, gpu, host); });
  });
}

int kat_thermo_eval(int n, const double *x, const double *T, double *out_dev, double *out_host)
{
  return guard([&] {
    for (int i = 0; i < n; ++i) thermo_eval(x[i], T[i], out_host + kThermo * i);
    void *d = nullptr, *s = nullptr;
    dfrg::check(dfrg_malloc(&d, sizeof(double) * (2 + kThermo) * n), "thermo buffers");
    double *dx = (double *)d, *dT = dx + n, *dout = dT + n;
    dfrg::check(dfrg_thread_stream(&s), "thread stream");
    dfrg::check(dfrg_memcpy_h2d(dx, x, sizeof(double) * n, s), "x");
    dfrg::check(dfrg_memcpy_h2d(dT, T, sizeof(double) * n, s), "T");
    hipLaunchKernelGGL(thermo_kernel, dim3((n + 127) / 128), dim3(128), 0, (hipStream_t)s, dx, dT, n, dout);
    dfrg::check_hip("thermo_kernel");
    dfrg::check(dfrg_memcpy_d2h(out_dev, dout, sizeof(double) * kThermo * n, s), "out");
    dfrg::check(dfrg_stream_sync(s), "out");
    dfrg_free(d);
  });
}

// sum_j over the radial GL rule of CothFiniteT(E, T)/E RBdot for the exponential regulator (b = 1),
// 3-dimensional (IntegratorGPU<3>), on the GPU and with the host template, for n (T, m2) pairs
int kat_coth_threshold(int nx, double x_extent, double k, const double *T, const double *m2, int n, double *gpu,
                       double *host)
{
  return guard([&] {
    IntegratorGPU<3, double, CothKernel<Reg>> g(provider(), {{(unsigned)nx}}, x_extent);
    IntegratorTBB<3, double, CothKernel<Reg>> h(provider(), {{(unsigned)nx}}, x_extent);
    for (int i = 0; i < n; ++i) {
      gpu[i] = g.get(k, T[i], m2[i]);
      host[i] = h.get(k, T[i], m2[i]);
    }
  });
}

} // extern "C"
