// DiFfRG/discretization/grid/combined_coordinates.hh -- (Matsubara frequency, |p|) coordinates of
// finite-temperature tables (reference combined_coordinates.hh:12-160).  The frequency axis is a
// bosonic or fermionic Matsubara stack [m_start, m_stop), the momentum axis logarithmic.
// backward(p0, |p|) returns the nearest stack index and the momentum index; a frequency beyond
// the stack is mapped to the nearest end, with the excess added to |p| in quadrature (O(4)
// continuation), as in the reference.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cmath>
#include <tuple>

#include <DiFfRG/common/math.hh>
#include <DiFfRG/discretization/grid/coordinates.hh>
#include <DiFfRG/discretization/grid/stack_coordinates.hh>

namespace DiFfRG
{
  namespace internal
  {
    template <typename Idx, typename NT, template <typename, typename> class Values> class MatsubaraLogCoordinates
    {
    public:
      using ctype = NT;
      static constexpr unsigned dim = 2;

      MatsubaraLogCoordinates(Idx m_start, Idx m_stop, NT m_T, unsigned grid_extent, NT p_start, NT p_stop, NT p_bias)
          : m_start(m_start), m_stop(m_stop), m_T(m_T), p_start(p_start), p_stop(p_stop), p_bias(p_bias),
            grid_extent(grid_extent), m_size(m_stop - m_start), matsubara_values(m_start, m_stop, m_T),
            logarithmic_coordinates(grid_extent, p_start, p_stop, p_bias)
      {
      }
      template <typename NT2>
      MatsubaraLogCoordinates(const LogarithmicCoordinates1D<NT2> &lc, Idx m_start, Idx m_stop, NT m_T)
          : MatsubaraLogCoordinates(m_start, m_stop, m_T, lc.size(), lc.start, lc.stop, lc.bias)
      {
      }

      __forceinline__ __host__ __device__ std::array<NT, 2> forward(const unsigned m, const unsigned p) const
      {
        return {{matsubara_values.forward(Idx(m)), logarithmic_coordinates.forward(p)}};
      }
      __forceinline__ __host__ __device__ std::tuple<Idx, NT> backward(const NT m, const NT p) const
      {
        Idx m_idx = matsubara_values.backward(m);
        const Idx last = Idx(m_size) - 1;
        if (m_idx > last || m_idx < 0) {
          m_idx = m_idx > last ? last : Idx(0);
          const NT dm = m - matsubara_values.forward(m_idx);
          return {m_idx, logarithmic_coordinates.backward(std::sqrt(p * p + dm * dm))};
        }
        return {m_idx, logarithmic_coordinates.backward(p)};
      }
      unsigned size() const { return m_size * grid_extent; }
      std::array<unsigned, 2> sizes() const { return {{m_size, grid_extent}}; }

      NT m_start, m_stop, m_T, p_start, p_stop, p_bias;

    private:
      unsigned grid_extent, m_size;
      Values<Idx, NT> matsubara_values;
      LogarithmicCoordinates1D<NT> logarithmic_coordinates;
    };
  } // namespace internal

  template <typename Idx, typename NT>
  class BosonicCoordinates1DFiniteT : public internal::MatsubaraLogCoordinates<Idx, NT, BosonicMatsubaraValues>
  {
    using Base = internal::MatsubaraLogCoordinates<Idx, NT, BosonicMatsubaraValues>;

  public:
    using Base::Base;
    template <typename Idx2, typename NT2>
    BosonicCoordinates1DFiniteT(const BosonicCoordinates1DFiniteT<Idx2, NT2> &o)
        : Base(Idx(o.m_start), Idx(o.m_stop), NT(o.m_T), o.sizes()[1], NT(o.p_start), NT(o.p_stop), NT(o.p_bias))
    {
    }
  };

  template <typename Idx, typename NT>
  class FermionicCoordinates1DFiniteT : public internal::MatsubaraLogCoordinates<Idx, NT, FermionicMatsubaraValues>
  {
    using Base = internal::MatsubaraLogCoordinates<Idx, NT, FermionicMatsubaraValues>;

  public:
    using Base::Base;
    template <typename Idx2, typename NT2>
    FermionicCoordinates1DFiniteT(const FermionicCoordinates1DFiniteT<Idx2, NT2> &o)
        : Base(Idx(o.m_start), Idx(o.m_stop), NT(o.m_T), o.sizes()[1], NT(o.p_start), NT(o.p_stop), NT(o.p_bias))
    {
    }
  };
} // namespace DiFfRG
