// DiFfRG/discretization/grid/coordinates.hh -- 1D index <-> momentum maps used by the
// interpolators and for external-momentum grids (reference coordinates.hh:97-252).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <stdexcept>
#include <vector>

// coordinate systems: a ctype, a dimension and a size (reference coordinates.hh:15-19)
template <typename T>
concept IsCoordinate = requires(T t) {
  typename T::ctype;
  T::dim;
  t.size();
};

namespace DiFfRG
{
  template <typename NT> class LinearCoordinates1D
  {
  public:
    using ctype = NT;
    static constexpr unsigned dim = 1;

    LinearCoordinates1D(unsigned grid_extent, double start, double stop)
        : start(start), stop(stop), grid_extent(grid_extent)
    {
      if (grid_extent == 0) throw std::runtime_error("LinearCoordinates1D: grid_extent must be > 0");
      a = (stop - start) / (grid_extent - 1.);
    }
    template <typename NT2>
    LinearCoordinates1D(const LinearCoordinates1D<NT2> &other) : LinearCoordinates1D(other.size(), other.start, other.stop)
    {
    }
    __forceinline__ __host__ __device__ NT forward(const unsigned &x) const { return start + a * x; }
    __forceinline__ __host__ __device__ NT backward(const NT &y) const { return (y - start) / a; }
    unsigned size() const { return grid_extent; }

    NT start, stop;

  private:
    unsigned grid_extent;
    NT a;
  };

  // y(i) = b expm1(a i/(n-1)) + c with a = bias, b = (stop - start)/expm1(bias), c = start
  template <typename NT> class LogarithmicCoordinates1D
  {
  public:
    using ctype = NT;
    static constexpr unsigned dim = 1;

    LogarithmicCoordinates1D(unsigned grid_extent, NT start, NT stop, NT bias)
        : start(start), stop(stop), bias(bias), grid_extent(grid_extent), gem1(grid_extent - 1.),
          gem1inv(1. / (grid_extent - 1.))
    {
      if (grid_extent == 0) throw std::runtime_error("LogarithmicCoordinates1D: grid_extent must be > 0");
      a = bias;
      b = (stop - start) / std::expm1(a);
      c = start;
    }
    template <typename NT2>
    LogarithmicCoordinates1D(const LogarithmicCoordinates1D<NT2> &other)
        : LogarithmicCoordinates1D(other.size(), other.start, other.stop, other.bias)
    {
    }
    __forceinline__ __host__ __device__ NT forward(const unsigned &x) const { return b * expm1(a * x * gem1inv) + c; }
    __forceinline__ __host__ __device__ NT backward(const NT &y) const { return log1p((y - c) / b) * gem1 / a; }
    __forceinline__ __host__ __device__ NT backward_derivative(const NT &y) const { return 1. / (y - c) * gem1 / a; }
    unsigned size() const { return grid_extent; }

    NT start, stop, bias;

  private:
    unsigned grid_extent;
    NT gem1, gem1inv;
    NT a, b, c;
  };

  template <typename Coordinates> std::vector<typename Coordinates::ctype> make_grid(const Coordinates &coordinates)
  {
    static_assert(Coordinates::dim == 1, "make_grid: 1D coordinates");
    std::vector<typename Coordinates::ctype> grid(coordinates.size());
    for (unsigned i = 0; i < coordinates.size(); ++i) grid[i] = coordinates.forward(i);
    return grid;
  }
} // namespace DiFfRG
