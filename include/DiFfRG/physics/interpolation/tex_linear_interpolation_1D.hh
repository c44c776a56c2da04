// DiFfRG/physics/interpolation/tex_linear_interpolation_1D.hh -- TexLinearInterpolator1D<NT,
// Coordinates> (reference tex_linear_interpolation_1D.hh:20-140): an FP32 table with clamped
// linear interpolation, the reference's texture-memory interpolator.  The reference filters with
// NVIDIA texture hardware (weights quantised to 1/256); here the table is FP32 in HBM and the lerp
// is computed in FP32 on the device and on the host alike, which is the reference's host formula.
// Values are returned as float (internal::__TLITypes).
#pragma once

#include <vector>

#include <DiFfRG/common/math.hh>
#include <DiFfRG/discretization/grid/coordinates.hh>
#include <DiFfRG/physics/interpolation/common.hh>

namespace DiFfRG
{
  template <typename NT, typename Coordinates> class TexLinearInterpolator1D
  {
    static_assert(Coordinates::dim == 1, "TexLinearInterpolator1D requires 1D coordinates");

  public:
    using ReturnType = typename internal::__TLITypes<NT>::ReturnType;

    TexLinearInterpolator1D(const std::vector<NT> &data, const Coordinates &coordinates)
        : TexLinearInterpolator1D(data.data(), coordinates)
    {
      if (data.size() != coordinates.size()) throw std::runtime_error("Data and coordinates must have the same size");
    }
    TexLinearInterpolator1D(const NT *data, const Coordinates &coordinates) : TexLinearInterpolator1D(coordinates)
    {
      update(data);
    }
    TexLinearInterpolator1D(const Coordinates &coordinates)
        : size(coordinates.size()), coordinates(coordinates),
          table(std::make_shared<internal::TableStorage<float>>(coordinates.size())), owner(true)
    {
      if (size < 2) throw std::runtime_error("TexLinearInterpolator1D: the table needs at least 2 points");
      d_table = table->device_ptr();
      h_table = table->host_ptr();
      table->upload();
    }
    TexLinearInterpolator1D(const TexLinearInterpolator1D &o)
        : size(o.size), coordinates(o.coordinates), table(o.table), d_table(o.d_table), h_table(o.h_table),
          owner(false)
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

    __host__ __device__ ReturnType operator()(const float x) const
    {
      const float idx = float(coordinates.backward(x));
#ifdef __HIP_DEVICE_COMPILE__
      return internal::lerp1(d_table, idx, size);
#else
      return internal::lerp1((const float *)h_table, idx, size);
#endif
    }

  private:
    unsigned size;
    Coordinates coordinates;
    std::shared_ptr<internal::TableStorage<float>> table;
    const float *d_table = nullptr;
    float *h_table = nullptr;
    bool owner;
  };
} // namespace DiFfRG
