// DiFfRG/physics/interpolation/tex_linear_interpolation_1D.hh -- TexLinearInterpolator1D<NT,
// Coordinates> (reference tex_linear_interpolation_1D.hh:20-140): an FP32 table with clamped
// linear interpolation, the reference's texture-memory interpolator.  The reference filters with
// NVIDIA texture hardware (weights quantised to 1/256); here the table is FP32 in HBM and the lerp
// is computed in FP32 on the device and on the host alike, which is the reference's host formula.
// Values are returned as float (internal::__TLITypes); dual-number data (autodiff::real) keeps a
// second FP32 table for the derivative and returns both channels.
#pragma once

#include <vector>

#include "dfrg/math.hh"
#include <DiFfRG/discretization/grid/coordinates.hh>
#include "dfrg/interp.hh"

namespace DiFfRG
{
  template <typename NT, typename Coordinates> class TexLinearInterpolator1D
  {
    static_assert(Coordinates::dim == 1, "TexLinearInterpolator1D requires 1D coordinates");

  public:
    using ReturnType = typename dfrg::interp::TexReturn<NT>::type;

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
        : size(coordinates.size()), coordinates(coordinates), owner(true)
    {
      if (size < 2) throw std::runtime_error("TexLinearInterpolator1D: the table needs at least 2 points");
      tables = std::make_shared<dfrg::interp::TexTables<NT>>(size);
      d_table = tables->device();
      d_table_AD = tables->device_AD();
      h_table = tables->host();
      h_table_AD = tables->host_AD();
    }
    TexLinearInterpolator1D(const TexLinearInterpolator1D &o)
        : size(o.size), coordinates(o.coordinates), tables(o.tables), d_table(o.d_table), d_table_AD(o.d_table_AD),
          h_table(o.h_table), h_table_AD(o.h_table_AD), owner(false)
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

    __host__ __device__ ReturnType operator()(const float x) const
    {
      const float idx = float(coordinates.backward(x));
      const unsigned n = size;
      const auto lerp = [idx, n](const float *t) { return dfrg::interp::lerp1(t, idx, n); };
#ifdef __HIP_DEVICE_COMPILE__
      return dfrg::interp::tex_value<NT>(d_table, d_table_AD, lerp);
#else
      return dfrg::interp::tex_value<NT>(h_table, h_table_AD, lerp);
#endif
    }

  private:
    unsigned size;
    Coordinates coordinates;
    std::shared_ptr<dfrg::interp::TexTables<NT>> tables;
    const float *d_table = nullptr, *d_table_AD = nullptr;
    float *h_table = nullptr, *h_table_AD = nullptr;
    bool owner;
  };
} // namespace DiFfRG
