// DiFfRG/physics/utils.hh -- request_data / update_data / update_interpolator / get_data over a grid
// of external points (reference physics/utils.hh:31-220).
//
// fun.request<NT>(k, point..., args...) is called for every grid point, as in the reference.  The
// GPU integrators of this library collect those requests into one batch per integrator
// (dfrg::tmpl::RequestBatcher), and request_data launches the batches of its loop as it returns,
// asynchronously, so host work after it overlaps the GPU: n points cost one launch and one
// synchronisation instead of n.  get_data, which the reference runs as a TBB parallel loop of
// get() calls, is built on the same requests: one batched launch, then the results in order.
#pragma once

#include <future>
#include <memory>
#include <tuple>
#include <type_traits>
#include <vector>

#include <DiFfRG/common/types.hh>
#include <DiFfRG/discretization/grid/coordinates.hh>

#include "dfrg/integrate.hh"

namespace DiFfRG
{
  namespace internal
  {
    template <typename P, typename F> decltype(auto) call_at(const P &p, F &&f)
    {
      if constexpr (std::is_arithmetic_v<std::decay_t<P>>)
        return f(p);
      else
        return std::apply([&](const auto &...q) { return f(q...); }, p);
    }
  } // namespace internal

  // fun.request<NT>(k, grid[i]..., args...) for every grid point (utils.hh:113-147)
  template <typename NT, typename FUN, typename GRID, typename... T>
    requires IsContainer<GRID>
  std::shared_ptr<std::vector<std::future<NT>>> request_data(FUN &fun, const GRID &grid, const double k, T... args)
  {
    auto futures = std::make_shared<std::vector<std::future<NT>>>();
    futures->reserve(grid.size());
    for (unsigned i = 0; i < grid.size(); ++i)
      futures->emplace_back(
          internal::call_at(grid[i], [&](const auto &...p) { return fun.template request<NT>(k, p..., args...); }));
    dfrg::tmpl::flush_thread(); // the loop's requests are one batch per integrator: launch them now
    return futures;
  }

  template <typename NT, typename FUN, typename GRID, typename... T>
    requires IsContainer<GRID>
  std::shared_ptr<std::vector<std::future<NT>>> request_data(FUN &fun, const GRID &grid, const double k,
                                                             std::tuple<T...> args)
  {
    return std::apply([&](auto... a) { return request_data<NT>(fun, grid, k, a...); }, args);
  }

  // the same on the grid of a coordinate system (utils.hh:31-74)
  template <typename NT, typename FUN, typename Coordinates, typename... T>
    requires IsCoordinate<Coordinates>
  std::shared_ptr<std::vector<std::future<NT>>> request_data(FUN &fun, Coordinates &coordinates, const double k,
                                                             T... args)
  {
    return request_data<NT>(fun, make_grid(coordinates), k, args...);
  }

  template <typename NT, typename FUN, typename Coordinates, typename... T>
    requires IsCoordinate<Coordinates>
  std::shared_ptr<std::vector<std::future<NT>>> request_data(FUN &fun, Coordinates &coordinates, const double k,
                                                             std::tuple<T...> args)
  {
    return std::apply([&](auto... a) { return request_data<NT>(fun, coordinates, k, a...); }, args);
  }

  // results of request_data in order into an array (utils.hh:149-163)
  template <typename NT1, typename NT2, typename... T>
  void update_data(std::shared_ptr<std::vector<std::future<NT1>>> futures, NT2 *destination)
  {
    for (unsigned i = 0; i < futures->size(); ++i) destination[i] = static_cast<NT2>((*futures)[i].get());
  }

  // results of request_data in order into an interpolator, then update() (utils.hh:165-180)
  template <typename NT, typename INT, typename... T>
  void update_interpolator(std::shared_ptr<std::vector<std::future<NT>>> futures, INT &destination)
  {
    for (unsigned i = 0; i < futures->size(); ++i) destination[i] = (*futures)[i].get();
    destination.update();
  }

  // dest[i] = fun.get<NT>(k, grid[i]..., args...) (utils.hh:182-220), as one batch of requests
  template <typename NT, typename FUN, typename GRID, typename... T>
  void get_data(NT *dest, FUN &fun, const GRID &grid, const double k, T... args)
  {
    update_data(request_data<NT>(fun, grid, k, args...), dest);
  }

  template <typename NT, typename FUN, typename GRID, typename... T>
  void get_data(NT *dest, FUN &fun, const GRID &grid, const double k, std::tuple<T...> args)
  {
    std::apply([&](auto... a) { get_data<NT>(dest, fun, grid, k, a...); }, args);
  }
} // namespace DiFfRG
