// DiFfRG/physics/interpolation/tex_linear_interpolation_2D.hh -- TexLinearInterpolator2D<NT,
// Coordinates> (reference tex_linear_interpolation_2D.hh:20-150): FP32 table over 2D coordinates,
// bilinear with the reference's clamp and corner rule, computed in FP32 on the device (HBM table)
// and on the host.  Returns float, or both channels of dual-number data.
#pragma once

#include <array>
#include <vector>

#include "dfrg/math.hh"
#include <DiFfRG/discretization/grid/coordinates.hh>
#include "dfrg/interp.hh"

namespace DiFfRG
{
  template <typename NT, typename Coordinates> class TexLinearInterpolator2D
  {
    static_assert(Coordinates::dim == 2, "TexLinearInterpolator2D requires 2D coordinates");

  public:
    using ReturnType = typename dfrg::interp::TexReturn<NT>::type;

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
        : size(coordinates.size()), shape(coordinates.sizes()), coordinates(coordinates), owner(true)
    {
      dfrg::interp::require_axes(shape.data(), 2, "TexLinearInterpolator2D");
      tables = std::make_shared<dfrg::interp::TexTables<NT>>(size);
      d_table = tables->device();
      d_table_AD = tables->device_AD();
      h_table = tables->host();
      h_table_AD = tables->host_AD();
    }
    TexLinearInterpolator2D(const TexLinearInterpolator2D &o)
        : size(o.size), shape(o.shape), coordinates(o.coordinates), tables(o.tables), d_table(o.d_table),
          d_table_AD(o.d_table_AD), h_table(o.h_table), h_table_AD(o.h_table_AD), owner(false)
    {
    }

    template <typename NT2> void update(const NT2 *data)
    {
      if (!owner) throw std::runtime_error("Cannot update data of non-owner interpolator");
      tables->fill(data);
      update();
    }
    void update()
    {
      if (!owner) throw std::runtime_error("Cannot update data of non-owner interpolator");
      tables->upload();
    }
    float *data() const { return h_table; }
    float *data_AD() const { return h_table_AD; }
    ReturnType &operator[](const unsigned i)
    {
      if constexpr (dfrg::DualTraits<NT>::is_dual)
        throw std::runtime_error("Cannot access dual-number data directly");
      else
        return h_table[i];
    }
    const ReturnType &operator[](const unsigned i) const
    {
      if constexpr (dfrg::DualTraits<NT>::is_dual)
        throw std::runtime_error("Cannot access dual-number data directly");
      else
        return h_table[i];
    }
    const Coordinates &get_coordinates() const { return coordinates; }

    __host__ __device__ ReturnType operator()(const float x, const float y) const
    {
      const auto idx = coordinates.backward(x, y);
      const float ix = float(idx[0]), iy = float(idx[1]);
      const unsigned nx = shape[0], ny = shape[1];
      const auto lerp = [=](const float *t) { return dfrg::interp::lerp2(t, ix, iy, nx, ny); };
#ifdef __HIP_DEVICE_COMPILE__
      return dfrg::interp::tex_value<NT>(d_table, d_table_AD, lerp);
#else
      return dfrg::interp::tex_value<NT>(h_table, h_table_AD, lerp);
#endif
    }

  private:
    unsigned size;
    std::array<unsigned, 2> shape;
    Coordinates coordinates;
    std::shared_ptr<dfrg::interp::TexTables<NT>> tables;
    const float *d_table = nullptr, *d_table_AD = nullptr;
    float *h_table = nullptr, *h_table_AD = nullptr;
    bool owner;
  };
} // namespace DiFfRG
