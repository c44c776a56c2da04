// DiFfRG/physics/interpolation/linear_interpolation_1D.hh -- tabulated dressing function
// evaluated inside kernels (reference linear_interpolation_1D.hh:22-113, .cu:15-98).
//
// The table lives in host memory and in HBM.  Copies are non-owning views (the copy the
// integrator passes by value into a kernel); update() on a view throws, as in the reference.
// operator() is the FP64-capable software lerp on the coordinate's index, clamped to [0, n-1];
// on the device it reads the HBM copy.  It replaces both LinearInterpolator1D and the FP32
// texture interpolator TexLinearInterpolator1D of the reference.
//
// FP64 logarithmic coordinates on the device: the reference's index idx = log1p(u) (N-1)/a,
// u = (x - c)/b, costs an FP64 division and log1p per call.  Here it comes from a table of sub-cell
// starts: K sub-cells per cell (K the smallest power of two with |a|/(K(N-1)) <= 1/4), their starts
// u_m = expm1(a m/(K(N-1))) and 1/(1 + u_m).  An FP32 log2 of the mantissa of 1 + u plus its exponent
// locates the sub-cell m, and the position in it is
//   K(N-1)/a log((1 + u)/(1 + u_m)) = K(N-1)/a 2 atanh(z),  z = delta/(2 + delta),
//   delta = (u - u_m)/(1 + u_m),
// with the atanh series to z^17 (|z| <= 0.125: truncation below 1e-17); an estimate one sub-cell off
// (within ~1e-6 of a boundary) is moved by one.  The index agrees with the reference's to its own FP64
// rounding (~1e-14 of a cell); the clamp to [0, N-1] and the reference's NaN rule (fmax(0, fmin(NaN,
// N-1)): a NaN index takes the last point) give the ends: u <= -1 takes the first point when
// -1 <= u and a > 0 (index -inf or negative), the last point otherwise (NaN, or +inf for a < 0);
// N = 1, bias 0 and NaN bias have a NaN or zero-width index: the last point.  The device code has no
// branch on the grid, so a user KERNEL's loop-invariant lookups (at q alone) stay hoistable.  Biases
// beyond +-700 (exp overflows) and grids of more than 2^21 sub-cells per unit of log2 are refused.
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
      if constexpr (kLogFP64) make_cells(); // may refuse the grid: before the table's allocation
      m_data = std::shared_ptr<NT[]>(new NT[size]);
      for (unsigned i = 0; i < size; ++i) m_data[i] = NT(in_data[i]);
      void *p = nullptr;
      // at least two points in HBM: a one-point table's lerp reads its point twice
      dfrg::check(dfrg_malloc(&p, sizeof(NT) * (size < 2 ? 2 : size)), "LinearInterpolator1D");
      device_data = std::shared_ptr<void>(p, [](void *q) { dfrg_free(q); });
      device_data_ptr = (const NT *)p;
      update();
    }
    LinearInterpolator1D(const LinearInterpolator1D &other)
        : size(other.size), coordinates(other.coordinates), m_data(other.m_data), device_data(other.device_data),
          device_data_ptr(other.device_data_ptr), owner(false), cells(other.cells), lc(other.lc)
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
      if (size == 1)
        dfrg::check(dfrg_memcpy_h2d((void *)(device_data_ptr + 1), m_data.get(), sizeof(NT), nullptr),
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
      if constexpr (kLogFP64) return log_cell_lerp((x - lc.c) * lc.inv_b);
#endif
      ct idx = coordinates.backward(x);
      idx = fmax(ct(0), fmin(idx, ct(size - 1)));
      const ct fl = floor(idx);
#ifdef __HIP_DEVICE_COMPILE__
      const NT *d = device_data_ptr;
#else
      const NT *d = m_data.get();
#endif
      return dfrg::tld(d + unsigned(fl)) * (ct(1) - idx + fl) + dfrg::tld(d + unsigned(ceil(idx))) * (idx - fl);
    }

  private:
    static constexpr bool kLogFP64 = std::is_same_v<Coordinates, LogarithmicCoordinates1D<double>>;

    // the device index of FP64 log coordinates: sub-cell table and constants
    struct LogCells {
      const double2 *cell = nullptr; // {u_m, 1/(1 + u_m)}, m = 0 .. K(N-1)
      double c = 0, inv_b = 0;       // u = (x - c) inv_b
      double two_s = 0, s_ln2 = 0;   // 2 K(N-1)/a, K(N-1)/a ln 2
      double inv_k = 1;              // 1/K
      int m_last = 0, k_shift = 0;   // last sub-cell (K(N-1) - 1), log2 K
      int last = 0;                  // last cell (N - 2; 0 for N = 1)
      int const_last = 0;            // 1: every point takes the last point (N = 1, bias 0 or NaN)
      int a_pos = 1;                 // bias > 0
    };

    void make_cells()
    {
      const double a = coordinates.bias;
      const unsigned N = size;
      LogCells L;
      L.last = N >= 2 ? int(N) - 2 : 0;
      std::vector<double2> h;
      if (N >= 2 && a == a && a != 0) {
        if (!(std::fabs(a) <= 700.))
          throw std::invalid_argument("LinearInterpolator1D: the bias of LogarithmicCoordinates1D must be within "
                                      "[-700, 700] (exp(bias) overflows beyond)");
        int k_shift = 0;
        while (std::fabs(a) / ((double)(1u << k_shift) * (N - 1.)) > 0.25) ++k_shift;
        const double K = double(1u << k_shift), subs = K * (N - 1.);
        if (subs / std::fabs(a) * 0.69314718055994530942 > double(1 << 21) || subs > double(1 << 24))
          throw std::invalid_argument("LinearInterpolator1D: too many points for this bias of "
                                      "LogarithmicCoordinates1D (more than 2^21 sub-cells per unit of log2)");
        h.resize((size_t)subs + 1);
        for (size_t m = 0; m < h.size(); ++m) {
          const double t = a * (double)m / subs;
          h[m] = make_double2(std::expm1(t), std::exp(-t));
        }
        L.c = coordinates.start;
        L.inv_b = 1. / ((coordinates.stop - coordinates.start) / std::expm1(a));
        L.two_s = 2. * subs / a;
        L.s_ln2 = subs / a * 0.69314718055994530942;
        L.inv_k = 1. / K;
        L.m_last = int(subs) - 1;
        L.k_shift = k_shift;
        L.a_pos = a > 0;
      } else {
        // one point, bias 0 or NaN: the reference's index is 0 x (...)/a with a NaN or infinite
        // factor, or the one point; all points take the last point.  A valid dummy table keeps the
        // device code branch-free.
        h = {make_double2(0., 1.), make_double2(0., 1.)};
        L.const_last = 1;
      }
      void *p = nullptr;
      dfrg::check(dfrg_malloc(&p, sizeof(double2) * h.size()), "LinearInterpolator1D cells");
      cells = std::shared_ptr<void>(p, [](void *q) { dfrg_free(q); });
      dfrg::check(dfrg_memcpy_h2d(p, h.data(), sizeof(double2) * h.size(), nullptr), "LinearInterpolator1D cells");
      dfrg::check(dfrg_stream_sync(nullptr), "LinearInterpolator1D cells");
      L.cell = (const double2 *)p;
      lc = L;
    }

    // the lerp at the reference's index of u = (x - c)/b (device, no branch on the grid)
    __device__ __forceinline__ NT log_cell_lerp(const double u_in) const
    {
      const bool inside = u_in > -1.; // false for NaN
      const double u = inside ? u_in : 0.;
      // sub-cell estimate: log2(1 + u) = e + log2(mant), mant in [1/2, 1), FP32 log of the mantissa
      int e = 0;
      const double mant = frexp(1. + u, &e);
      const double est = ((double)e + (double)__builtin_amdgcn_logf((float)mant)) * lc.s_ln2;
      int m = (int)floor(fmax(0., fmin(est, (double)lc.m_last)));
      const double2 cl = dfrg::gld(lc.cell + m); // {u_m, 1/(1 + u_m)}
      double dl = (u - cl.x) * cl.y;
      dl = fmax(-0.5, fmin(dl, 1.0)); // beyond the sub-cell by more than one: the clamp decides
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
      double fs = lc.two_s * (z * pz); // position in the sub-cell
      // the estimate can be one sub-cell off where the index is within ~1e-6 of a boundary
      const bool down = fs < 0. && m > 0, up = !down && fs >= 1. && m < lc.m_last;
      m += up ? 1 : down ? -1 : 0;
      fs += down ? 1. : up ? -1. : 0.;
      fs = fmax(0., fmin(fs, 1.));
      // sub-cell -> cell (K a power of two: exact)
      int i = m >> lc.k_shift;
      double f = ((double)(m & ((1 << lc.k_shift) - 1)) + fs) * lc.inv_k;
      // the clamp's ends: u <= -1 or NaN, and the degenerate grids
      const bool first = !lc.const_last && !inside && lc.a_pos && u_in >= -1.;
      const bool to_last = lc.const_last || (!inside && !first);
      i = first ? 0 : to_last ? lc.last : i;
      f = first ? 0. : to_last ? 1. : f;
      const NT *d = device_data_ptr + i;
      return dfrg::gld(d) * (1. - f) + dfrg::gld(d + 1) * f;
    }

    unsigned size;
    Coordinates coordinates;
    std::shared_ptr<NT[]> m_data;
    std::shared_ptr<void> device_data;
    const NT *device_data_ptr = nullptr;
    bool owner;
    std::shared_ptr<void> cells;
    LogCells lc;
  };
} // namespace DiFfRG
