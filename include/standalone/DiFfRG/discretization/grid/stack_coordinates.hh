// DiFfRG/discretization/grid/stack_coordinates.hh -- integer-indexed stacks and Matsubara
// frequency grids (reference stack_coordinates.hh:10-140).
//   IndexStack:               i <-> i + start
//   BosonicMatsubaraValues:   n <-> 2 pi T (n + start)
//   FermionicMatsubaraValues: n <-> 2 pi T (n + start + 1/2)
// backward() rounds to the nearest frequency.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <stdexcept>
#include <string>

#include <DiFfRG/discretization/grid/coordinates.hh>

namespace DiFfRG
{
  namespace internal
  {
    template <typename Idx> void check_stack_range(Idx start, Idx stop)
    {
      if (start > stop)
        throw std::invalid_argument("start must be smaller than stop; start = " + std::to_string(start) +
                                    ", stop = " + std::to_string(stop));
    }
  } // namespace internal

  template <typename Idx> class IndexStack
  {
  public:
    using ctype = Idx;
    static constexpr unsigned dim = 1;

    IndexStack(Idx start, Idx stop) : start(start), stop(stop), m_size(stop - start)
    {
      internal::check_stack_range(start, stop);
    }
    template <typename Idx2> IndexStack(const LinearCoordinates1D<Idx2> &other) : IndexStack(other.start, other.stop) {}

    __forceinline__ __host__ __device__ Idx forward(const Idx x) const { return x + start; }
    __forceinline__ __host__ __device__ Idx backward(const Idx y) const { return y - start; }
    unsigned size() const { return m_size; }

    Idx start, stop;

  private:
    unsigned m_size;
  };

  template <typename Idx, typename NT> class BosonicMatsubaraValues
  {
  public:
    using ctype = NT;
    static constexpr unsigned dim = 1;

    BosonicMatsubaraValues(Idx start, Idx stop, NT T) : start(start), stop(stop), T(T), m_size(stop - start)
    {
      internal::check_stack_range(start, stop);
    }
    template <typename Idx2, typename NT2>
    BosonicMatsubaraValues(const BosonicMatsubaraValues<Idx2, NT2> &other)
        : BosonicMatsubaraValues(other.start, other.stop, other.T)
    {
    }

    __forceinline__ __host__ __device__ NT forward(const Idx &x) const { return NT(x + start) * 2. * M_PI * T; }
    __forceinline__ __host__ __device__ Idx backward(const NT &y) const
    {
      return Idx(std::round(y / (2. * M_PI * T))) - start;
    }
    unsigned size() const { return m_size; }

    Idx start, stop;
    NT T;

  private:
    unsigned m_size;
  };

  template <typename Idx, typename NT> class FermionicMatsubaraValues
  {
  public:
    using ctype = NT;
    static constexpr unsigned dim = 1;

    FermionicMatsubaraValues(Idx start, Idx stop, NT T) : start(start), stop(stop), T(T), m_size(stop - start)
    {
      internal::check_stack_range(start, stop);
    }
    template <typename Idx2, typename NT2>
    FermionicMatsubaraValues(const FermionicMatsubaraValues<Idx2, NT2> &other)
        : FermionicMatsubaraValues(other.start, other.stop, other.T)
    {
    }

    __forceinline__ __host__ __device__ NT forward(const Idx &x) const
    {
      return (NT(x + start) + 0.5) * 2. * M_PI * T;
    }
    __forceinline__ __host__ __device__ Idx backward(const NT &y) const
    {
      return Idx(std::round((y - M_PI * T) / (2. * M_PI * T))) - start;
    }
    unsigned size() const { return m_size; }

    Idx start, stop;
    NT T;

  private:
    unsigned m_size;
  };
} // namespace DiFfRG
