// DiFfRG/physics/interpolation/linear_interpolation_1D.hh -- tabulated dressing function
// evaluated inside kernels (reference linear_interpolation_1D.hh:22-113, .cu:15-98).
//
// The table lives in host memory and in HBM.  Copies are non-owning views (the copy the
// integrator passes by value into a kernel); update() on a view throws, as in the reference.
// operator() is the FP64-capable software lerp on the coordinate's index, clamped to [0, n-1];
// on the device it reads the HBM copy.  It replaces both LinearInterpolator1D and the FP32
// texture interpolator TexLinearInterpolator1D of the reference.
#pragma once

#include <cmath>
#include <memory>
#include <stdexcept>
#include <vector>

#include "dfrg/check.hh"
#include "dfrg/math.hh"
#include <DiFfRG/discretization/grid/coordinates.hh>

namespace DiFfRG
{
  template <typename NT, typename Coordinates> class LinearInterpolator1D
  {
    static_assert(Coordinates::dim == 1, "LinearInterpolator1D requires 1D coordinates");

  public:
    LinearInterpolator1D(const std::vector<NT> &data, const Coordinates &coordinates)
        : LinearInterpolator1D(data.data(), coordinates)
    {
      if (data.size() != coordinates.size()) throw std::runtime_error("Data and coordinates must have the same size");
    }
    LinearInterpolator1D(const Coordinates &coordinates)
        : LinearInterpolator1D(std::vector<NT>(coordinates.size(), NT(0)), coordinates)
    {
    }
    LinearInterpolator1D(const NT *in_data, const Coordinates &coordinates)
        : size(coordinates.size()), coordinates(coordinates), owner(true)
    {
      dfrg::ensure_device();
      m_data = std::shared_ptr<NT[]>(new NT[size]);
      for (unsigned i = 0; i < size; ++i) m_data[i] = NT(in_data[i]);
      void *p = nullptr;
      dfrg::check(dfrg_malloc(&p, sizeof(NT) * size), "LinearInterpolator1D");
      device_data = std::shared_ptr<void>(p, [](void *q) { dfrg_free(q); });
      device_data_ptr = (const NT *)p;
      update();
    }
    LinearInterpolator1D(const LinearInterpolator1D &other)
        : size(other.size), coordinates(other.coordinates), m_data(other.m_data), device_data(other.device_data),
          device_data_ptr(other.device_data_ptr), owner(false)
    {
    }

    template <typename NT2> void update(const NT2 *data)
    {
      if (!owner) throw std::runtime_error("Cannot update data of non-owner interpolator");
      for (unsigned i = 0; i < size; ++i) m_data[i] = NT(data[i]);
      update();
    }
    void update()
    {
      if (!owner) throw std::runtime_error("Cannot update data of non-owner interpolator");
      dfrg::check(dfrg_memcpy_h2d((void *)device_data_ptr, m_data.get(), sizeof(NT) * size, nullptr),
                 "LinearInterpolator1D::update");
      dfrg::check(dfrg_stream_sync(nullptr), "LinearInterpolator1D::update");
    }

    NT *data() const { return m_data.get(); }
    NT &operator[](const unsigned i) { return m_data[i]; }
    const NT &operator[](const unsigned i) const { return m_data[i]; }
    const Coordinates &get_coordinates() const { return coordinates; }

    __host__ __device__ NT operator()(const typename Coordinates::ctype x) const
    {
      using ct = typename Coordinates::ctype;
      ct idx = coordinates.backward(x);
      idx = fmax(ct(0), fmin(idx, ct(size - 1)));
      const ct fl = floor(idx);
#ifdef __HIP_DEVICE_COMPILE__
      const NT *d = device_data_ptr;
#else
      const NT *d = m_data.get();
#endif
      return d[unsigned(fl)] * (ct(1) - idx + fl) + d[unsigned(ceil(idx))] * (idx - fl);
    }

  private:
    unsigned size;
    Coordinates coordinates;
    std::shared_ptr<NT[]> m_data;
    std::shared_ptr<void> device_data;
    const NT *device_data_ptr = nullptr;
    bool owner;
  };
} // namespace DiFfRG
