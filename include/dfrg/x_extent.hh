// dfrg/x_extent.hh -- choice of the radial extent x_extent = q_max^2/k^2 of the momentum integrals,
// the reference's FlowEquations::update_x_extent (flow_equations.cc:54-105), over this library's
// Gauss-Legendre rules.
//
// With I(X) the d = 4 radial integral of optimize_x(q^2) over x = q^2/k^2 in [0, X] at k = 1
// (LoopIntegrals::integrate<double, 4>, loop_integrals.hh:35-70), rules of orders f n, 2 f n and
// 10 f n (f = quadrature_factor, n = x_quadrature_order) for I1 = I(X), I2 = I(2X), I3 = I(10X):
// X starts at 1 and grows by 15% until |(I2 - I1)/I1| and |(I2 - I3)/I2| are both <= tolerance.
// An integrand that vanishes at the ten probe points 1 + 1e-3 + (1.5^i - 1) leaves X = 1; three
// consecutive rounds in which both ratios grow throw std::runtime_error.
#pragma once

#include <cmath>
#include <functional>
#include <stdexcept>

#include "dfrg/math.hh"
#include <DiFfRG/common/quadrature/quadrature_provider.hh>

namespace dfrg
{
  namespace tmpl
  {
    inline double radial_integral_4d(DiFfRG::QuadratureProvider &qp, const std::function<double(double)> &f,
                                     const unsigned order, const double extent)
    {
      const auto &x = qp.get_points<double>(order);
      const auto &w = qp.get_weights<double>(order);
      const double prefactor = 2.0 * M_PI * M_PI / std::pow(2.0 * M_PI, 4); // S_4 / (2 pi)^4
      double sum = 0.0;
      for (unsigned i = 0; i < order; ++i) {
        const double xi = x[i] * extent;
        sum += w[i] * extent * prefactor * (0.5 * xi) * f(xi);
      }
      return sum;
    }

    inline double optimize_x_extent(DiFfRG::QuadratureProvider &qp, const std::function<double(double)> &optimize_x,
                                    const unsigned x_quadrature_order, const double x_extent_tolerance,
                                    const unsigned quadrature_factor = 5)
    {
      bool all_zero = true;
      for (int i = 0; i < 10; ++i)
        if (!dfrg::is_close(optimize_x(1.0 + 1e-3 + (std::pow(1.5, (double)i) - 1.0)), 0.0)) all_zero = false;
      if (all_zero) return 1.0;

      const unsigned o1 = quadrature_factor * x_quadrature_order;
      double x_extent = 1.0, eps1 = 1.0, eps2 = 1.0;
      unsigned growing = 0;
      while (eps1 > x_extent_tolerance || eps2 > x_extent_tolerance) {
        const double I1 = radial_integral_4d(qp, optimize_x, o1, x_extent);
        const double I2 = radial_integral_4d(qp, optimize_x, 2 * o1, 2.0 * x_extent);
        const double I3 = radial_integral_4d(qp, optimize_x, 10 * o1, 10.0 * x_extent);
        const double e1 = std::abs((I2 - I1) / I1), e2 = std::abs((I2 - I3) / I2);
        growing = (e1 > eps1 && e2 > eps2) ? growing + 1 : 0;
        if (growing > 2)
          throw std::runtime_error("Cannot reach requested precision for x_extent - increase x_quadrature_order or "
                                   "decrease x_extent_tolerance.");
        eps1 = e1;
        eps2 = e2;
        if (eps1 > x_extent_tolerance || eps2 > x_extent_tolerance) x_extent *= 1.15;
      }
      return x_extent;
    }
  } // namespace tmpl
} // namespace dfrg
