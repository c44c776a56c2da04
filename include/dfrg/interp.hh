// dfrg/interp.hh -- shared pieces of the drop-in interpolators: the return type of the FP32
// ("Tex") tables (the role of the reference's interpolation/common.hh:8-25), their derivative
// channel for dual numbers (dfrg/dual.hh), and the host + HBM table storage behind every
// interpolator.
//
// Copies of an interpolator share one storage and are non-owning views: the copy an integrator
// passes by value into a kernel reads the HBM table, and update() on a view throws, as in the
// reference (linear_interpolation_1D.cu:41).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>

#include "dfrg/dual.hh"
#include "dfrg/check.hh"

namespace dfrg
{
  namespace interp
  {
    // value type of the FP32 tables: float for float and double data, the dual type itself for
    // dual data (reference interpolation/common.hh:10-24)
    template <typename NT> struct TexReturn {
      static_assert(DualTraits<NT>::is_dual, "FP32 table interpolators hold float, double or dual-number data");
      using type = NT;
    };
    template <> struct TexReturn<float> {
      using type = float;
    };
    template <> struct TexReturn<double> {
      using type = float;
    };

    // n values of type V on the host and in HBM
    template <typename V> class TableStorage
    {
    public:
      explicit TableStorage(unsigned n) : n(n)
      {
        if (n == 0) throw std::runtime_error("interpolator: empty table");
        ensure_device();
        host = std::shared_ptr<V[]>(new V[n]());
        void *p = nullptr;
        check(dfrg_malloc(&p, sizeof(V) * n), "interpolator table");
        dev = std::shared_ptr<void>(p, [](void *q) { dfrg_free(q); });
      }
      V *host_ptr() const { return host.get(); }
      const V *device_ptr() const { return (const V *)dev.get(); }
      void upload() const
      {
        check(dfrg_memcpy_h2d(dev.get(), host.get(), sizeof(V) * n, nullptr), "interpolator upload");
        check(dfrg_stream_sync(nullptr), "interpolator upload");
      }
      unsigned size() const { return n; }

    private:
      unsigned n;
      std::shared_ptr<V[]> host;
      std::shared_ptr<void> dev;
    };

    // Neighbours and weights of a clamped index on an axis of n >= 2 points: i1 = min(ceil(idx +
    // 1e-16), n - 1), i0 = i1 - 1, w0 = i1 - idx, w1 = idx - i0 (the corner rule of the reference's
    // multilinear interpolators, linear_interpolation_2D.hh:93-106).  The clamp is fmax(0, fmin(idx,
    // n - 1)), the reference's device semantics: a NaN index (a point the coordinates cannot map,
    // such as log1p below -1) clamps to the last point, on the host as on the device.
    template <typename C> struct Axis {
      unsigned i0, i1;
      C w0, w1;
      __forceinline__ __host__ __device__ Axis(C idx, unsigned n)
      {
        idx = fmax(C(0), fmin(idx, C(n - 1)));
        const C top = fmin(ceil(idx + C(1e-16)), C(n - 1));
        i1 = top < C(1) ? 1u : unsigned(top);
        i0 = i1 - 1;
        w0 = C(i1) - idx;
        w1 = idx - C(i0);
      }
    };

    // Clamped 1D lerp between floor and ceil (reference linear_interpolation_1D.hh:75-92), same clamp.
    template <typename V, typename C> __forceinline__ __host__ __device__ V lerp1(const V *d, C idx, unsigned n)
    {
      idx = fmax(C(0), fmin(idx, C(n - 1)));
      const C fl = floor(idx);
      return tld(d + unsigned(fl)) * (C(1) - idx + fl) + tld(d + unsigned(ceil(idx))) * (idx - fl);
    }

    template <typename V, typename C>
    __forceinline__ __host__ __device__ V lerp2(const V *d, C ix, C iy, unsigned nx, unsigned ny)
    {
      const Axis<C> x(ix, nx), y(iy, ny);
      const V *r0 = d + (size_t)x.i0 * ny, *r1 = d + (size_t)x.i1 * ny;
      return tld(r0 + y.i0) * x.w0 * y.w0 + tld(r0 + y.i1) * x.w0 * y.w1 + tld(r1 + y.i0) * x.w1 * y.w0 +
             tld(r1 + y.i1) * x.w1 * y.w1;
    }

    template <typename V, typename C>
    __forceinline__ __host__ __device__ V lerp3(const V *d, C ix, C iy, C iz, unsigned nx, unsigned ny, unsigned nz)
    {
      const Axis<C> x(ix, nx), y(iy, ny), z(iz, nz);
      auto at = [&](unsigned a, unsigned b, unsigned c) { return tld(d + ((size_t)a * ny + b) * nz + c); };
      return at(x.i0, y.i0, z.i0) * x.w0 * y.w0 * z.w0 + at(x.i0, y.i0, z.i1) * x.w0 * y.w0 * z.w1 +
             at(x.i0, y.i1, z.i0) * x.w0 * y.w1 * z.w0 + at(x.i0, y.i1, z.i1) * x.w0 * y.w1 * z.w1 +
             at(x.i1, y.i0, z.i0) * x.w1 * y.w0 * z.w0 + at(x.i1, y.i0, z.i1) * x.w1 * y.w0 * z.w1 +
             at(x.i1, y.i1, z.i0) * x.w1 * y.w1 * z.w0 + at(x.i1, y.i1, z.i1) * x.w1 * y.w1 * z.w1;
    }

    // The FP32 tables of a Tex interpolator over n points: values, and derivatives when NT is a
    // dual number (zero until data with derivatives is given, as in the reference).
    template <typename NT> class TexTables
    {
    public:
      static constexpr bool dual = DualTraits<NT>::is_dual;
      explicit TexTables(unsigned n)
          : n(n), val(std::make_shared<TableStorage<float>>(n)),
            der(dual ? std::make_shared<TableStorage<float>>(n) : nullptr)
      {
        upload();
      }
      // values (and derivatives, when both NT and NT2 carry them) from n entries of data
      template <typename NT2> void fill(const NT2 *data) const
      {
        float *v = val->host_ptr(), *dv = der ? der->host_ptr() : nullptr;
        for (unsigned i = 0; i < n; ++i) {
          if constexpr (DualTraits<NT2>::is_dual) {
            v[i] = static_cast<float>(DualTraits<NT2>::value(data[i]));
            if (dv) dv[i] = static_cast<float>(DualTraits<NT2>::derivative(data[i]));
          } else
            v[i] = static_cast<float>(data[i]);
        }
      }
      void upload() const
      {
        val->upload();
        if (der) der->upload();
      }
      float *host() const { return val->host_ptr(); }
      float *host_AD() const { return der ? der->host_ptr() : nullptr; }
      const float *device() const { return val->device_ptr(); }
      const float *device_AD() const { return der ? der->device_ptr() : nullptr; }

    private:
      unsigned n;
      std::shared_ptr<TableStorage<float>> val, der;
    };

    // One interpolated value of a Tex interpolator: lerp(values), or for dual data
    // make(lerp(values), lerp(derivatives)).  lerp maps a table pointer to the FP32 result.
    template <typename NT, typename L>
    __forceinline__ __host__ __device__ typename TexReturn<NT>::type tex_value(const float *v, const float *dv,
                                                                                       const L &lerp)
    {
      if constexpr (DualTraits<NT>::is_dual)
        return DualTraits<NT>::make(double(lerp(v)), double(lerp(dv)));
      else
        return lerp(v);
    }

    inline void require_axes(const unsigned *n, int dim, const char *who)
    {
      for (int i = 0; i < dim; ++i)
        if (n[i] < 2) throw std::runtime_error(std::string(who) + ": every axis needs at least 2 points");
    }
  } // namespace interp
} // namespace dfrg

