// DiFfRG/common/quadrature/quadrature_provider.hh -- QuadratureProvider with the interface of the
// reference class (quadrature_provider.hh:138-248): cached Gauss-Legendre and Matsubara rules, as
// host vectors and device pointers.  Rules are generated in FP64 on the GPU by libdfrg.so and
// cached per (order) and per (T, E) with the reference's 1e-6 relative key match.
#pragma once

#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <type_traits>
#include <vector>

#include "dfrg/check.hh"
#include "dfrg/math.hh"

namespace DiFfRG
{
  enum class QuadratureType { legendre, chebyshev, laguerre, hermite, jacobi, count };

  class QuadratureProvider
  {
  public:
    // QuadratureProvider() (quadrature_provider.cu:199-204): vacuum 48, min Matsubara 32, add 0
    QuadratureProvider() : QuadratureProvider(48, 32, 0) {}
    // QuadratureProvider(json) (quadrature_provider.cu:206-228)
    template <dfrg::JsonLike JSON>
    QuadratureProvider(const JSON &json)
        : QuadratureProvider((int)json.get_uint("/integration/vacuum_quad_size", 48),
                             (int)json.get_uint("/integration/min_matsubara_size", 16),
                             (int)json.get_uint("/integration/add_matsubara_size", 0))
    {
    }
    QuadratureProvider(int vacuum_quad_size, int min_matsubara_size, int add_matsubara_size)
    {
      dfrg::ensure_device();
      dfrg::check(dfrg_qp_create(&qp, vacuum_quad_size, min_matsubara_size, add_matsubara_size), "QuadratureProvider");
    }
    QuadratureProvider(const QuadratureProvider &) = delete;
    QuadratureProvider &operator=(const QuadratureProvider &) = delete;
    ~QuadratureProvider() { dfrg_qp_destroy(qp); }

    template <typename NT = double>
    const std::vector<NT> &get_points(const size_t order, const QuadratureType type = QuadratureType::legendre)
    {
      return host_gl<NT>(order, type).first;
    }
    template <typename NT = double>
    const std::vector<NT> &get_weights(const size_t order, const QuadratureType type = QuadratureType::legendre)
    {
      return host_gl<NT>(order, type).second;
    }
    template <typename NT = double> const std::vector<NT> &get_matsubara_points(const NT T, const NT typical_E)
    {
      return host_mats<NT>(T, typical_E).first;
    }
    template <typename NT = double> const std::vector<NT> &get_matsubara_weights(const NT T, const NT typical_E)
    {
      return host_mats<NT>(T, typical_E).second;
    }
    template <typename NT = double>
    const NT *get_device_points(const size_t order, const int device = 0,
                                const QuadratureType type = QuadratureType::legendre)
    {
      return device_gl<NT>(order, device, type, 0);
    }
    template <typename NT = double>
    const NT *get_device_weights(const size_t order, const int device = 0,
                                 const QuadratureType type = QuadratureType::legendre)
    {
      return device_gl<NT>(order, device, type, 1);
    }
    template <typename NT = double>
    const NT *get_device_matsubara_points(const NT T, const NT typical_E, const int device = 0)
    {
      return device_mats<NT>(T, typical_E, device, 0);
    }
    template <typename NT = double>
    const NT *get_device_matsubara_weights(const NT T, const NT typical_E, const int device = 0)
    {
      return device_mats<NT>(T, typical_E, device, 1);
    }

    dfrg_qp *handle() const { return qp; }

  private:
    template <typename NT> static constexpr int fp32_flag()
    {
      static_assert(std::is_same_v<NT, double> || std::is_same_v<NT, float>,
                    "QuadratureProvider serves double or float rules");
      return std::is_same_v<NT, float> ? 1 : 0;
    }
    static void check_type(QuadratureType type)
    {
      if (type != QuadratureType::legendre)
        throw std::runtime_error("QuadratureProvider: only Gauss-Legendre rules are provided");
    }
    template <typename NT> const std::pair<std::vector<NT>, std::vector<NT>> &host_gl(size_t order, QuadratureType type)
    {
      check_type(type);
      std::lock_guard<std::mutex> lock(mutex);
      auto &cache = gl_cache<NT>();
      auto it = cache.find(order);
      if (it != cache.end()) return it->second;
      const void *hx, *hw;
      dfrg::check(dfrg_qp_gl(qp, (int)order, fp32_flag<NT>(), nullptr, nullptr, &hx, &hw), "get_points");
      const NT *x = (const NT *)hx, *w = (const NT *)hw;
      return cache.emplace(order, std::make_pair(std::vector<NT>(x, x + order), std::vector<NT>(w, w + order)))
          .first->second;
    }
    template <typename NT> const NT *device_gl(size_t order, int device, QuadratureType type, int which)
    {
      (void)device; // one device per process
      check_type(type);
      const void *dx, *dw;
      dfrg::check(dfrg_qp_gl(qp, (int)order, fp32_flag<NT>(), &dx, &dw, nullptr, nullptr), "get_device_points");
      return (const NT *)(which == 0 ? dx : dw);
    }
    template <typename NT> const std::pair<std::vector<NT>, std::vector<NT>> &host_mats(NT T, NT E)
    {
      std::lock_guard<std::mutex> lock(mutex);
      int n = 0;
      double Tr = 0;
      const void *hx, *hw;
      dfrg::check(dfrg_qp_matsubara(qp, (double)T, (double)E, fp32_flag<NT>(), &n, &Tr, nullptr, nullptr, &hx, &hw),
                 "get_matsubara_points");
      auto &cache = mats_cache<NT>();
      auto it = cache.find(hx);
      if (it != cache.end()) return it->second;
      const NT *x = (const NT *)hx, *w = (const NT *)hw;
      return cache.emplace(hx, std::make_pair(std::vector<NT>(x, x + n), std::vector<NT>(w, w + n))).first->second;
    }
    template <typename NT> const NT *device_mats(NT T, NT E, int device, int which)
    {
      (void)device;
      int n = 0;
      double Tr = 0;
      const void *dx, *dw;
      dfrg::check(dfrg_qp_matsubara(qp, (double)T, (double)E, fp32_flag<NT>(), &n, &Tr, &dx, &dw, nullptr, nullptr),
                 "get_device_matsubara_points");
      return (const NT *)(which == 0 ? dx : dw);
    }
    template <typename NT> std::map<size_t, std::pair<std::vector<NT>, std::vector<NT>>> &gl_cache()
    {
      if constexpr (std::is_same_v<NT, double>)
        return gl_d;
      else
        return gl_f;
    }
    template <typename NT> std::map<const void *, std::pair<std::vector<NT>, std::vector<NT>>> &mats_cache()
    {
      if constexpr (std::is_same_v<NT, double>)
        return mats_d;
      else
        return mats_f;
    }

    dfrg_qp *qp = nullptr;
    std::mutex mutex;
    std::map<size_t, std::pair<std::vector<double>, std::vector<double>>> gl_d;
    std::map<size_t, std::pair<std::vector<float>, std::vector<float>>> gl_f;
    std::map<const void *, std::pair<std::vector<double>, std::vector<double>>> mats_d;
    std::map<const void *, std::pair<std::vector<float>, std::vector<float>>> mats_f;
  };
} // namespace DiFfRG
