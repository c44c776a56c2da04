// DiFfRG/physics/interpolation/tex_linear_interpolation_3D.hh -- TexLinearInterpolator3D<NT,
// Coordinates> (reference tex_linear_interpolation_3D.hh:20-170): FP32 table over 3D coordinates,
// trilinear with the reference's clamp and corner rule, computed in FP32 on the device (HBM table)
// and on the host.  Returns float.
#pragma once

#include <array>
#include <vector>

#include <DiFfRG/common/math.hh>
#include <DiFfRG/discretization/grid/coordinates.hh>
#include <DiFfRG/physics/interpolation/common.hh>

namespace DiFfRG
{
  template <typename NT, typename Coordinates> class TexLinearInterpolator3D
  {
    static_assert(Coordinates::dim == 3, "TexLinearInterpolator3D requires 3D coordinates");

  public:
    using ReturnType = typename internal::__TLITypes<NT>::ReturnType;

    TexLinearInterpolator3D(const std::vector<NT> &data, const Coordinates &coordinates)
        : TexLinearInterpolator3D(data.data(), coordinates)
    {
      if (data.size() != coordinates.size()) throw std::runtime_error("Data and coordinates must have the same size");
    }
    TexLinearInterpolator3D(const NT *data, const Coordinates &coordinates) : TexLinearInterpolator3D(coordinates)
    {
      update(data);
    }
    TexLinearInterpolator3D(const Coordinates &coordinates)
        : size(coordinates.size()), shape(coordinates.sizes()), coordinates(coordinates),
          table(std::make_shared<internal::TableStorage<float>>(coordinates.size())), owner(true)
    {
      internal::require_axes(shape.data(), 3, "TexLinearInterpolator3D");
      d_table = table->device_ptr();
      h_table = table->host_ptr();
      table->upload();
    }
    TexLinearInterpolator3D(const TexLinearInterpolator3D &o)
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

    __host__ __device__ ReturnType operator()(const float x, const float y, const float z) const
    {
      const auto idx = coordinates.backward(x, y, z);
#ifdef __HIP_DEVICE_COMPILE__
      const float *d = d_table;
#else
      const float *d = h_table;
#endif
      return internal::lerp3(d, float(idx[0]), float(idx[1]), float(idx[2]), shape[0], shape[1], shape[2]);
    }

  private:
    unsigned size;
    std::array<unsigned, 3> shape;
    Coordinates coordinates;
    std::shared_ptr<internal::TableStorage<float>> table;
    const float *d_table = nullptr;
    float *h_table = nullptr;
    bool owner;
  };
} // namespace DiFfRG
