// DiFfRG/physics/interpolation/tex_linear_interpolation_2D.hh -- TexLinearInterpolator2D<NT,
// Coordinates> (reference tex_linear_interpolation_2D.hh:20-150): FP32 table over 2D coordinates,
// bilinear with the reference's clamp and corner rule, computed in FP32 on the device (HBM table)
// and on the host.  Returns float.
#pragma once

#include <array>
#include <vector>

#include <DiFfRG/common/math.hh>
#include <DiFfRG/discretization/grid/coordinates.hh>
#include <DiFfRG/physics/interpolation/common.hh>

namespace DiFfRG
{
  template <typename NT, typename Coordinates> class TexLinearInterpolator2D
  {
    static_assert(Coordinates::dim == 2, "TexLinearInterpolator2D requires 2D coordinates");

  public:
    using ReturnType = typename internal::__TLITypes<NT>::ReturnType;

    TexLinearInterpolator2D(const std::vector<NT> &data, const Coordinates &coordinates)
        : TexLinearInterpolator2D(data.data(), coordinates)
    {
      if (data.size() != coordinates.size()) throw std::runtime_error("Data and coordinates must have the same size");
    }
    TexLinearInterpolator2D(const NT *data, const Coordinates &coordinates) : TexLinearInterpolator2D(coordinates)
    {
      update(data);
    }
    TexLinearInterpolator2D(const Coordinates &coordinates)
        : size(coordinates.size()), shape(coordinates.sizes()), coordinates(coordinates),
          table(std::make_shared<internal::TableStorage<float>>(coordinates.size())), owner(true)
    {
      internal::require_axes(shape.data(), 2, "TexLinearInterpolator2D");
      d_table = table->device_ptr();
      h_table = table->host_ptr();
      table->upload();
    }
    TexLinearInterpolator2D(const TexLinearInterpolator2D &o)
        : size(o.size), shape(o.shape), coordinates(o.coordinates), table(o.table), d_table(o.d_table),
          h_table(o.h_table), owner(false)
    {
    }

    template <typename NT2> void update(const NT2 *data)
    {
      if (!owner) throw std::runtime_error("Cannot update data of non-owner interpolator");
      for (unsigned i = 0; i < size; ++i) h_table[i] = static_cast<float>(data[i]);
      update();
    }
    void update()
    {
      if (!owner) throw std::runtime_error("Cannot update data of non-owner interpolator");
      table->upload();
    }
    float *data() const { return h_table; }
    ReturnType &operator[](const unsigned i) { return h_table[i]; }
    const ReturnType &operator[](const unsigned i) const { return h_table[i]; }
    const Coordinates &get_coordinates() const { return coordinates; }

    __host__ __device__ ReturnType operator()(const float x, const float y) const
    {
      const auto idx = coordinates.backward(x, y);
#ifdef __HIP_DEVICE_COMPILE__
      const float *d = d_table;
#else
      const float *d = h_table;
#endif
      return internal::lerp2(d, float(idx[0]), float(idx[1]), shape[0], shape[1]);
    }

  private:
    unsigned size;
    std::array<unsigned, 2> shape;
    Coordinates coordinates;
    std::shared_ptr<internal::TableStorage<float>> table;
    const float *d_table = nullptr;
    float *h_table = nullptr;
    bool owner;
  };
} // namespace DiFfRG
