// DiFfRG/physics/interpolation/linear_interpolation_2D.hh -- LinearInterpolator2D<NT, Coordinates>
// (reference linear_interpolation_2D.hh:16-130): bilinear interpolation of a row-major table over
// 2D coordinates (a CoordinatePackND of two 1D systems), clamped to the table, on the host and
// inside kernels (HBM table, full NT precision).
#pragma once

#include <array>
#include <vector>

#include "dfrg/math.hh"
#include <DiFfRG/discretization/grid/coordinates.hh>
#include "dfrg/interp.hh"

namespace DiFfRG
{
  template <typename NT, typename Coordinates> class LinearInterpolator2D
  {
    static_assert(Coordinates::dim == 2, "LinearInterpolator2D requires 2D coordinates");
    using C = typename Coordinates::ctype;

  public:
    LinearInterpolator2D(const std::vector<NT> &data, const Coordinates &coordinates)
        : LinearInterpolator2D(data.data(), coordinates)
    {
      if (data.size() != coordinates.size()) throw std::runtime_error("Data and coordinates must have the same size");
    }
    LinearInterpolator2D(const NT *data, const Coordinates &coordinates) : LinearInterpolator2D(coordinates)
    {
      update(data);
    }
    LinearInterpolator2D(const Coordinates &coordinates)
        : size(coordinates.size()), shape(coordinates.sizes()), coordinates(coordinates),
          table(std::make_shared<dfrg::interp::TableStorage<NT>>(coordinates.size())), owner(true)
    {
      dfrg::interp::require_axes(shape.data(), 2, "LinearInterpolator2D");
      d_table = table->device_ptr();
      h_table = table->host_ptr();
      table->upload();
    }
    LinearInterpolator2D(const LinearInterpolator2D &o)
        : size(o.size), shape(o.shape), coordinates(o.coordinates), table(o.table), d_table(o.d_table),
          h_table(o.h_table), owner(false)
    {
    }

    template <typename NT2> void update(const NT2 *data)
    {
      if (!owner) throw std::runtime_error("Cannot update data of non-owner interpolator");
      for (unsigned i = 0; i < size; ++i) h_table[i] = NT(data[i]);
      update();
    }
    void update()
    {
      if (!owner) throw std::runtime_error("Cannot update data of non-owner interpolator");
      table->upload();
    }
    NT *data() const { return h_table; }
    NT &operator[](const unsigned i) { return h_table[i]; }
    const NT &operator[](const unsigned i) const { return h_table[i]; }
    const Coordinates &get_coordinates() const { return coordinates; }

    __host__ __device__ NT operator()(const C x, const C y) const
    {
      const auto idx = coordinates.backward(x, y);
#ifdef __HIP_DEVICE_COMPILE__
      const NT *d = d_table;
#else
      const NT *d = h_table;
#endif
      return dfrg::interp::lerp2(d, C(idx[0]), C(idx[1]), shape[0], shape[1]);
    }

  private:
    unsigned size;
    std::array<unsigned, 2> shape;
    Coordinates coordinates;
    std::shared_ptr<dfrg::interp::TableStorage<NT>> table;
    const NT *d_table = nullptr;
    NT *h_table = nullptr;
    bool owner;
  };
} // namespace DiFfRG
