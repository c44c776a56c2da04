// DiFfRG/discretization/grid/coordinates.hh -- index <-> momentum maps used by the interpolators
// and for external-momentum grids (reference coordinates.hh:15-252): linear and logarithmic 1D
// coordinates and their N-dimensional products.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cmath>
#include <stdexcept>
#include <tuple>
#include <utility>
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

  // Product of 1D coordinate systems (reference coordinates.hh:24-95): forward/backward act axis by
  // axis, sizes() lists the axis lengths, size() is their product.  Row-major order (last axis
  // fastest) everywhere a pack indexes a table.
  template <typename... Coordinates> class CoordinatePackND
  {
    static_assert(sizeof...(Coordinates) > 0, "CoordinatePackND requires at least one coordinate system");
    static_assert(((Coordinates::dim == 1) && ...), "CoordinatePackND requires all coordinates to have dim 1");
    using First = std::tuple_element_t<0, std::tuple<Coordinates...>>;
    static_assert((std::is_same_v<typename Coordinates::ctype, typename First::ctype> && ...),
                  "CoordinatePackND requires one ctype");

  public:
    using ctype = typename First::ctype;
    static constexpr unsigned dim = sizeof...(Coordinates);

    CoordinatePackND(Coordinates... c) : coordinates(c...) {}

    template <typename... I> __forceinline__ __host__ __device__ std::array<ctype, dim> forward(I &&...i) const
    {
      static_assert(sizeof...(I) == dim);
      return map(std::make_index_sequence<dim>(), [](const auto &c, const auto &v) { return c.forward(v); }, i...);
    }
    template <typename... I> __forceinline__ __host__ __device__ std::array<ctype, dim> backward(I &&...i) const
    {
      static_assert(sizeof...(I) == dim);
      return map(std::make_index_sequence<dim>(), [](const auto &c, const auto &v) { return c.backward(v); }, i...);
    }
    template <unsigned i> const auto &get_coordinates() const { return std::get<i>(coordinates); }
    unsigned size() const
    {
      unsigned n = 1;
      for (unsigned s : sizes()) n *= s;
      return n;
    }
    std::array<unsigned, dim> sizes() const
    {
      return std::apply([](const auto &...c) { return std::array<unsigned, dim>{{c.size()...}}; }, coordinates);
    }

  private:
    template <size_t... Is, typename F, typename... I>
    __forceinline__ __host__ __device__ std::array<ctype, dim> map(std::index_sequence<Is...>, F f, const I &...i) const
    {
      const auto args = std::forward_as_tuple(i...);
      return {{ctype(f(std::get<Is>(coordinates), std::get<Is>(args)))...}};
    }
    std::tuple<Coordinates...> coordinates;
  };

  // physical points of a coordinate system in table order (reference coordinates.hh:199-230):
  // values for 1D, std::array points for 2D and 3D
  template <typename Coordinates> auto make_grid(const Coordinates &coordinates)
  {
    using ctype = typename Coordinates::ctype;
    if constexpr (Coordinates::dim == 1) {
      std::vector<ctype> grid(coordinates.size());
      for (unsigned i = 0; i < coordinates.size(); ++i) grid[i] = coordinates.forward(i);
      return grid;
    } else if constexpr (Coordinates::dim == 2) {
      const auto n = coordinates.sizes();
      std::vector<std::array<ctype, 2>> grid(coordinates.size());
      for (unsigned i = 0; i < n[0]; ++i)
        for (unsigned j = 0; j < n[1]; ++j) {
          const auto v = coordinates.forward(i, j);
          grid[i * n[1] + j] = {{ctype(v[0]), ctype(v[1])}};
        }
      return grid;
    } else {
      static_assert(Coordinates::dim == 3, "make_grid: 1D, 2D and 3D coordinates");
      const auto n = coordinates.sizes();
      std::vector<std::array<ctype, 3>> grid(coordinates.size());
      for (unsigned i = 0; i < n[0]; ++i)
        for (unsigned j = 0; j < n[1]; ++j)
          for (unsigned k = 0; k < n[2]; ++k) {
            const auto v = coordinates.forward(i, j, k);
            grid[(i * n[1] + j) * n[2] + k] = {{ctype(v[0]), ctype(v[1]), ctype(v[2])}};
          }
      return grid;
    }
  }

  // flat table indices in the order of make_grid (reference coordinates.hh:232-252)
  template <typename Coordinates> std::vector<double> make_idx_grid(const Coordinates &coordinates)
  {
    static_assert(Coordinates::dim >= 1 && Coordinates::dim <= 3, "make_idx_grid: 1D, 2D and 3D coordinates");
    std::vector<double> grid(coordinates.size());
    for (unsigned i = 0; i < grid.size(); ++i) grid[i] = i;
    return grid;
  }
} // namespace DiFfRG
