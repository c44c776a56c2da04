// DiFfRG/physics/interpolation/linear_interpolation_1D.hh -- tabulated dressing function
// evaluated inside kernels (reference linear_interpolation_1D.hh:22-113, .cu:15-98).
//
// The table lives in host memory and in HBM.  Copies are non-owning views (the copy the
// integrator passes by value into a kernel); update() on a view throws, as in the reference.
// operator() is the FP64-capable software lerp on the coordinate's index, clamped to [0, n-1];
// on the device it reads the HBM copy.  It replaces both LinearInterpolator1D and the FP32
// texture interpolator TexLinearInterpolator1D of the reference.
//
// FP64 logarithmic coordinates on the device: the reference's index log1p((x - c)/b) (N-1)/a costs
// an FP64 division and log1p per call.  Here the cell comes from an FP32 log estimate and the
// position in the cell from the tabulated cell start u_i = expm1(a i/(N-1)):
//   idx - i = (N-1)/a log((1 + u)/(1 + u_i)) = (N-1)/a 2 atanh(z),  z = delta/(2 + delta),
//   delta = (u - u_i)/(1 + u_i),
// with the atanh series to z^17 (|z| <= 0.125 when a/(N-1) <= 1/4: truncation below 1e-17), and
// the estimate's cell moved by one where the position falls outside [0, 1).  The index agrees with
// the reference's to its own FP64 rounding (~1e-14 of a cell); the clamp to [0, N-1] and its NaN
// rule (an index that is NaN takes the last point) are the reference's: points with (x - c)/b <= -1/2
// (negative index, or NaN below -1) and NaN take the clamp's endpoint.  Coarser grids take the
// reference's formula.
#pragma once

#include <cmath>
#include <memory>
#include <stdexcept>
#include <type_traits>
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
      if constexpr (kLogFP64) make_cells();
    }
    LinearInterpolator1D(const LinearInterpolator1D &other)
        : size(other.size), coordinates(other.coordinates), m_data(other.m_data), device_data(other.device_data),
          device_data_ptr(other.device_data_ptr), owner(false), cells(other.cells), cell_ptr(other.cell_ptr),
          lc_c(other.lc_c), lc_inv_b(other.lc_inv_b), lc_2s(other.lc_2s), lc_s_ln2(other.lc_s_ln2)
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
#ifdef __HIP_DEVICE_COMPILE__
      if constexpr (kLogFP64) {
        if (cell_ptr) return log_cell_lerp((x - lc_c) * lc_inv_b);
      }
#endif
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
    static constexpr bool kLogFP64 = std::is_same_v<Coordinates, LogarithmicCoordinates1D<double>>;

    // cell starts {u_i, 1/(1 + u_i)} of FP64 log coordinates, when a/(N-1) <= 1/4 and N <= 65536
    // (the FP32 estimate then lands within one cell)
    void make_cells()
    {
      const double a = coordinates.bias, gem1inv = 1. / (size - 1.);
      if (size < 2 || size > 65536 || !(a * gem1inv <= 0.25) || !(a > 0)) return;
      std::vector<double> h(2 * size);
      for (unsigned i = 0; i < size; ++i) {
        h[2 * i] = std::expm1(a * i * gem1inv);
        h[2 * i + 1] = 1. / (1. + h[2 * i]);
      }
      void *p = nullptr;
      dfrg::check(dfrg_malloc(&p, sizeof(double) * h.size()), "LinearInterpolator1D cells");
      cells = std::shared_ptr<void>(p, [](void *q) { dfrg_free(q); });
      dfrg::check(dfrg_memcpy_h2d(p, h.data(), sizeof(double) * h.size(), nullptr), "LinearInterpolator1D cells");
      dfrg::check(dfrg_stream_sync(nullptr), "LinearInterpolator1D cells");
      cell_ptr = (const double *)p;
      lc_c = coordinates.start;
      lc_inv_b = 1. / ((coordinates.stop - coordinates.start) / std::expm1(a));
      lc_2s = 2. * (size - 1.) / a;
      lc_s_ln2 = float((size - 1.) / a * 0.69314718055994530942);
    }

    // the lerp at the index of u = (x - c)/b (device).  For u <= -1/2 the reference's index is
    // negative (-1 <= u: the first point) or NaN (u < -1, NaN: the last point, its NaN clamp); those
    // take the clamp's endpoint by selects, so the loop that calls this has no branch
    __device__ __forceinline__ NT log_cell_lerp(const double u_in) const
    {
      const int last = (int)size - 2;
      const bool inside = u_in > -0.5;
      const double u = inside ? u_in : 0.;
      const float est = __builtin_amdgcn_logf(1.f + (float)u) * lc_s_ln2; // v_log_f32 is log2
      int i = (int)floorf(fmaxf(0.f, fminf(est, (float)last)));
      const double *cl = cell_ptr + 2 * i;
      double dl = (u - cl[0]) * cl[1];
      dl = fmax(-0.5, fmin(dl, 1.0)); // beyond the cell by more than one: the clamp decides
      // z = dl/(2 + dl): FP32 seed, two FP64 Newton steps
      const double den = 2. + dl;
      double r = (double)__builtin_amdgcn_rcpf((float)den);
      r = fma(r, fma(-den, r, 1.), r);
      r = fma(r, fma(-den, r, 1.), r);
      const double z = dl * r, w = z * z;
      double pz = fma(w, 1. / 17, 1. / 15);
      pz = fma(pz, w, 1. / 13);
      pz = fma(pz, w, 1. / 11);
      pz = fma(pz, w, 1. / 9);
      pz = fma(pz, w, 1. / 7);
      pz = fma(pz, w, 1. / 5);
      pz = fma(pz, w, 1. / 3);
      pz = fma(pz, w, 1.);
      double f = lc_2s * (z * pz);
      // the FP32 estimate can be one cell off where the index is within ~1e-5 of an integer
      if (f < 0. && i > 0) {
        --i;
        f += 1.;
      } else if (f >= 1. && i < last) {
        ++i;
        f -= 1.;
      }
      f = fmax(0., fmin(f, 1.));
      if (!inside) {
        const bool first = u_in >= -1.; // false for NaN
        i = first ? 0 : last;
        f = first ? 0. : 1.;
      }
      const NT *d = device_data_ptr + i;
      return d[0] * (1. - f) + d[1] * f;
    }

    unsigned size;
    Coordinates coordinates;
    std::shared_ptr<NT[]> m_data;
    std::shared_ptr<void> device_data;
    const NT *device_data_ptr = nullptr;
    bool owner;
    std::shared_ptr<void> cells;
    const double *cell_ptr = nullptr;
    double lc_c = 0, lc_inv_b = 0, lc_2s = 0;
    float lc_s_ln2 = 0;
  };
} // namespace DiFfRG
