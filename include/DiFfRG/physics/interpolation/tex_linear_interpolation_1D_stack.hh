// DiFfRG/physics/interpolation/tex_linear_interpolation_1D_stack.hh --
// TexLinearInterpolator1DStack<NT, Coordinates, max_stack_size> (reference
// tex_linear_interpolation_1D_stack.hh:20-150): a stack of FP32 1D tables, one per Matsubara
// frequency (or index) of 2D stack coordinates such as BosonicCoordinates1DFiniteT.  operator()(m,
// p) takes the stack entry the coordinates select for m and lerps the momentum index, clamped.
// The stack is one row-major FP32 array in HBM (row s at s * p_size), with a second one for the
// derivative of dual-number data; FP32 lerp on device and host.
#pragma once

#include <vector>

#include "dfrg/math.hh"
#include <DiFfRG/discretization/grid/combined_coordinates.hh>
#include "dfrg/interp.hh"

namespace DiFfRG
{
  template <typename NT, typename Coordinates, unsigned max_stack_size = 64> class TexLinearInterpolator1DStack
  {
    static_assert(Coordinates::dim == 2, "TexLinearInterpolator1DStack requires 2D coordinates");

  public:
    using ReturnType = typename dfrg::interp::TexReturn<NT>::type;

    TexLinearInterpolator1DStack(const std::vector<NT> &data, const Coordinates &coordinates)
        : TexLinearInterpolator1DStack(data.data(), coordinates)
    {
      if (data.size() != coordinates.size()) throw std::runtime_error("Data and coordinates must have the same size");
    }
    TexLinearInterpolator1DStack(const NT *data, const Coordinates &coordinates)
        : TexLinearInterpolator1DStack(coordinates)
    {
      update(data);
    }
    TexLinearInterpolator1DStack(const Coordinates &coordinates)
        : stack_size(coordinates.sizes()[0]), p_size(coordinates.sizes()[1]), coordinates(coordinates), owner(true)
    {
      if (stack_size > max_stack_size)
        throw std::runtime_error("TexLinearInterpolator1DStack: stack size exceeds max_stack_size");
      if (p_size < 2) throw std::runtime_error("TexLinearInterpolator1DStack: every table needs at least 2 points");
      tables = std::make_shared<dfrg::interp::TexTables<NT>>(stack_size * p_size);
      d_table = tables->device();
      d_table_AD = tables->device_AD();
      h_table = tables->host();
      h_table_AD = tables->host_AD();
    }
    template <typename Coordinates2>
    TexLinearInterpolator1DStack(const Coordinates2 &coordinates, const unsigned m_start, const unsigned m_end,
                                 const double T)
        : TexLinearInterpolator1DStack(Coordinates(coordinates, m_start, m_end, T))
    {
    }
    TexLinearInterpolator1DStack(const TexLinearInterpolator1DStack &o)
        : stack_size(o.stack_size), p_size(o.p_size), coordinates(o.coordinates), tables(o.tables),
          d_table(o.d_table), d_table_AD(o.d_table_AD), h_table(o.h_table), h_table_AD(o.h_table_AD), owner(false)
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
    float *data(const unsigned s) const { return h_table + (size_t)s * p_size; }
    float *data_AD(const unsigned s) const { return h_table_AD ? h_table_AD + (size_t)s * p_size : nullptr; }
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

    __host__ __device__ ReturnType operator()(const float m, const float p) const
    {
      const auto mp = coordinates.backward(m, p);
      int s = int(std::get<0>(mp));
      s = s < 0 ? 0 : (s >= int(stack_size) ? int(stack_size) - 1 : s);
      const float ip = float(std::get<1>(mp));
      const unsigned n = p_size;
      const size_t off = (size_t)s * p_size;
      const auto lerp = [=](const float *t) { return dfrg::interp::lerp1(t + off, ip, n); };
#ifdef __HIP_DEVICE_COMPILE__
      return dfrg::interp::tex_value<NT>(d_table, d_table_AD, lerp);
#else
      return dfrg::interp::tex_value<NT>(h_table, h_table_AD, lerp);
#endif
    }

  private:
    unsigned stack_size, p_size;
    Coordinates coordinates;
    std::shared_ptr<dfrg::interp::TexTables<NT>> tables;
    const float *d_table = nullptr, *d_table_AD = nullptr;
    float *h_table = nullptr, *h_table_AD = nullptr;
    bool owner;
  };
} // namespace DiFfRG
