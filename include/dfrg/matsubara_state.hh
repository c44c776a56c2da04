// dfrg/matsubara_state.hh -- temperature, typical energy scale and Matsubara rule of a finite-T
// integrator (reference integrator_finiteTq0_gpu.hh:105-126, 276-281; the CPU classes alike).
//
// E defaults to 10 T.  Unless it was set explicitly (set_T(T, E != 0) or set_E), it follows the
// RG scale k with a 2.5% hysteresis, and each change fetches the provider's rule for (T, E).
#pragma once

#include <algorithm>
#include <cmath>
#include <limits>

#include "dfrg/math.hh"
#include <DiFfRG/common/quadrature/quadrature_provider.hh>

namespace dfrg
{
  namespace tmpl
  {
    template <class ctype, bool device> struct MatsubaraState {
      DiFfRG::QuadratureProvider *provider = nullptr;
      ctype T = std::numeric_limits<ctype>::quiet_NaN(), E = 0;
      bool manual_E = false;
      unsigned size = 0;
      const ctype *x = nullptr, *w = nullptr; // device or host rule, per `device`

      void set_T(const ctype T_new, const ctype E_new = 0)
      {
        const bool keep = dfrg::is_close(T_new, T) && (!device || E_new != 0) &&
                          (std::abs(E_new - E) / std::max(E_new, E) < 2.5e-2);
        if (keep) return;
        T = T_new;
        E = dfrg::is_close(E_new, (ctype)0) ? 10 * T : E_new;
        manual_E = !dfrg::is_close(E_new, (ctype)0);
        size = (unsigned)provider->template get_matsubara_points<ctype>(T, E).size();
        if constexpr (device) {
          x = provider->template get_device_matsubara_points<ctype>(T, E);
          w = provider->template get_device_matsubara_weights<ctype>(T, E);
        } else {
          x = provider->template get_matsubara_points<ctype>(T, E).data();
          w = provider->template get_matsubara_weights<ctype>(T, E).data();
        }
      }
      void set_E(const ctype E_new) { set_T(T, E_new); }
      void track(const ctype k)
      {
        if (!manual_E && (std::abs(k - E) / std::max(k, E) > 2.5e-2)) {
          set_T(T, k);
          manual_E = false;
        }
      }
    };
  } // namespace tmpl
} // namespace dfrg
