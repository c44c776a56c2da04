// dfrg/integrate.hh -- the engine behind the drop-in integrator templates.
//
// A user KERNEL (static __host__ __device__ kernel()/constant()) is device code of the caller's
// translation unit, so these kernels are templates compiled there by hipcc.  One integral is the
// flat node space (radial node, inner nodes...) of its integrator, cut into units that depend on its
// node counts only (Plan).  A block takes a unit: for rows of many inner nodes the unit is one
// radial node, and its lanes walk the inner nodes with q fixed, so the KERNEL's q-only
// subexpressions are loop-invariant and computed once per lane and row; smaller rows are grouped,
// their radial factors staged in LDS, and walked flat.  Lanes walk nodes in mixed radix (no
// division per node), read the inner nodes' coordinates and weight factors from a per-integrator
// table, and add KERNEL::kernel(...) x (radial factor x inner factors) to an FP64 accumulator.
// Blocks reduce with a wave64 DPP tree and LDS; the units of one integral meet in per-unit partial
// slots whose last-arriving block adds them with a fixed tree over the units (wave64 DPP, 64
// units per round, rounds in unit order).  Units, their lane assignment and the
// reduction order depend on the node counts only: results are bitwise reproducible, and a request
// in any batch gives bitwise the value of get().  Launches keep as many blocks as fit on the device
// resident and let them stride over the units.
//
// Accumulation is in FP64 for every number type: FP32 number types (float, complex<float>) evaluate
// the KERNEL in FP32 on FP32 nodes and accumulate in FP64, the mixed form of the template path.
//
// get() is a batch of one, launched at once and waited for.  request() adds to the integrator's
// open batch; a batch is launched (asynchronously: one kernel that reads the tuples from pinned
// host memory and whose finishing blocks store the results there, an event) when any future of
// this thread is read,
// when request_data() finishes its loop, when the batch is full, or when a request needs another
// launch configuration.  Reading a future launches every open batch of the calling thread first,
// so batches of different integrators run back to back before the first wait.
//
// Scratch: every host thread owns an arena in HBM on its stream (ticket counters, partial slots,
// argument tuples).  Counters are zeroed when a thread takes an arena (new, grown, or left
// by a thread that ended) and every launch's finishing blocks reset the counters they used, so
// consecutive launches on the stream share it without clearing.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <functional>
#include <future>
#include <initializer_list>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

#include "dfrg/check.hh"
#include "dfrg/device.hh"
#include "dfrg/dual.hh"
#include "dfrg/math.hh"

namespace dfrg
{
  namespace tmpl
  {
    constexpr int kMaxQ = 256; // radial nodes a flat unit may span (LDS staging)

    // ------------------------------------------------------------------------------------------
    // number types: scalar (float, double) or two-component (complex: real/imag; dual numbers:
    // value/derivative), accumulated component-wise in FP64
    // ------------------------------------------------------------------------------------------
    template <class NT> constexpr bool is_scalar_nt = std::is_same_v<NT, double> || std::is_same_v<NT, float>;

    template <class NT> struct Parts {
      static constexpr bool dual = DualTraits<NT>::is_dual;
      static __host__ __device__ __forceinline__ auto first(const NT &v)
      {
        if constexpr (dual)
          return DualTraits<NT>::value(v);
        else
          return v.real();
      }
      static __host__ __device__ __forceinline__ auto second(const NT &v)
      {
        if constexpr (dual)
          return DualTraits<NT>::derivative(v);
        else
          return v.imag();
      }
      using P = std::remove_cv_t<std::remove_reference_t<decltype(first(std::declval<NT>()))>>;
      static __host__ __device__ __forceinline__ NT make(double a, double b)
      {
        if constexpr (dual)
          return DualTraits<NT>::make(P(a), P(b));
        else
          return NT(P(a), P(b));
      }
    };

    template <class NT> struct Acc {
      static constexpr int N = is_scalar_nt<NT> ? 1 : 2;
      double v[N];
      __host__ __device__ __forceinline__ Acc()
      {
        for (int i = 0; i < N; ++i) v[i] = 0.0;
      }
      __host__ __device__ __forceinline__ void add(const NT &x)
      {
        if constexpr (N == 1)
          v[0] += double(x);
        else {
          v[0] += double(Parts<NT>::first(x));
          v[1] += double(Parts<NT>::second(x));
        }
      }
      static __host__ __device__ __forceinline__ NT value(const double (&s)[N])
      {
        if constexpr (N == 1)
          return NT(s[0]);
        else
          return Parts<NT>::make(s[0], s[1]);
      }
    };

    // ------------------------------------------------------------------------------------------
    // node sets.  Radial nodes x_j of a Gauss-Legendre rule on [0,1] give q_j = k sqrt(x_j x_ext)
    // with factor element(q_j) w_j x_ext; angular nodes come from a per-integrator table of
    // (coordinate, weight factor); Matsubara nodes from the provider's rule.
    // ------------------------------------------------------------------------------------------
    template <class ctype> struct AngNode {
      ctype c, a;
    };

    template <class ctype> struct Nodes {
      const ctype *xq = nullptr, *wq = nullptr;    // radial rule on [0,1]
      const AngNode<ctype> *ang = nullptr;          // angular tables, one per angular digit
      const ctype *mx = nullptr, *mw = nullptr;     // Matsubara rule (positive nodes)
      int nx = 0, na = 0, nm = 0;
      int nb = 0;                                   // an inner axis of its own order (Cartesian qz, q0)
      ctype x_extent = 0, m_T = 0;
      ctype q_min = 0, q_ext = 0;                   // Cartesian: range of the first axis
    };

    template <class ctype> constexpr ctype two_pi = ctype(2) * ctype(3.14159265358979323846);

    // A new batcher state for an integrator whose nodes changed (any integrator, any change), so no
    // batch of requests mixes two node sets.
    inline unsigned long long next_node_state()
    {
      static std::atomic<unsigned long long> s{1};
      return s.fetch_add(1, std::memory_order_relaxed);
    }

    // Each shape: ND digits (radial first, the fastest last, as the reference's flat loops),
    // element(q, k) of the radial node, and the node term given the staged radial factor rf.
    // IntegratorGPU<d> (integrator_gpu.hh:34-60)
    template <int d> struct Radial {
      static constexpr int ND = 1, NA = 0;
      template <class C> static __host__ __device__ void radix(const Nodes<C> &n, int (&r)[ND]) { r[0] = n.nx; }
      template <class C> static __host__ __device__ __forceinline__ C element(C q, C k)
      {
        return S_d_prec<C>(d) * (powr<d - 2>(q) / 2 * powr<2>(k)) / powr<d>(two_pi<C>);
      }
      template <class NT, class KERNEL, class C, class... T>
      static __device__ __forceinline__ NT term(const Nodes<C> &, C q, C rf, const int (&)[ND], C k, const T &...t)
      {
        return NT(rf * KERNEL::kernel(q, k, t...));
      }
    };

    // IntegratorAngleGPU<d> (integrator_angle_gpu.hh:21-40): (q, cos)
    template <int d> struct Angle1 {
      static constexpr int ND = 2, NA = 1;
      template <class C> static __host__ __device__ void radix(const Nodes<C> &n, int (&r)[ND])
      {
        r[0] = n.nx;
        r[1] = n.na;
      }
      template <class C> static __host__ __device__ __forceinline__ C element(C q, C k)
      {
        return S_d_prec<C>(d) * (powr<d - 2>(q) / C(2) * powr<2>(k)) * (C(1) / C(2)) / powr<d>(two_pi<C>);
      }
      template <class NT, class KERNEL, class C, class... T>
      static __device__ __forceinline__ NT term(const Nodes<C> &n, C q, C rf, const int (&g)[ND], C k, const T &...t)
      {
        const AngNode<C> a = gld(n.ang + g[1]);
        return NT((rf * a.a) * KERNEL::kernel(q, a.c, k, t...));
      }
    };

    // Integrator4D2AngGPU (integrator_4D_2ang_gpu.hh:20-45): (q, cos1, cos2), the phi integral as 2 pi
    struct Angle2 {
      static constexpr int ND = 3, NA = 2;
      template <class C> static __host__ __device__ void radix(const Nodes<C> &n, int (&r)[ND])
      {
        r[0] = n.nx;
        r[1] = r[2] = n.na;
      }
      template <class C> static __host__ __device__ __forceinline__ C element(C q, C k)
      {
        return (powr<2>(q) * C(0.5) * powr<2>(k)) * two_pi<C> / powr<4>(two_pi<C>);
      }
      template <class NT, class KERNEL, class C, class... T>
      static __device__ __forceinline__ NT term(const Nodes<C> &n, C q, C rf, const int (&g)[ND], C k, const T &...t)
      {
        const AngNode<C> a1 = gld(n.ang + g[1]), a2 = gld(n.ang + n.na + g[2]);
        return NT((rf * (a1.a * a2.a)) * KERNEL::kernel(q, a1.c, a2.c, k, t...));
      }
    };

    // Integrator4DGPU (integrator_4D_gpu.hh:20-49): (q, cos1, phi, cos2)
    struct Angle3 {
      static constexpr int ND = 4, NA = 3;
      template <class C> static __host__ __device__ void radix(const Nodes<C> &n, int (&r)[ND])
      {
        r[0] = n.nx;
        r[1] = r[2] = r[3] = n.na;
      }
      template <class C> static __host__ __device__ __forceinline__ C element(C q, C k)
      {
        return (powr<2>(q) * C(0.5) * powr<2>(k)) / powr<4>(two_pi<C>);
      }
      template <class NT, class KERNEL, class C, class... T>
      static __device__ __forceinline__ NT term(const Nodes<C> &n, C q, C rf, const int (&g)[ND], C k, const T &...t)
      {
        const AngNode<C> a1 = gld(n.ang + g[1]), ap = gld(n.ang + n.na + g[2]), a2 = gld(n.ang + 2 * n.na + g[3]);
        return NT((rf * (a1.a * ap.a * a2.a)) * KERNEL::kernel(q, a1.c, a2.c, ap.c, k, t...));
      }
    };

    // Integrator3DGPU (integrator_3D_gpu.hh:21-43): (q, cos, phi); tables cos, then phi
    struct Angle1Phi {
      static constexpr int ND = 3, NA = 2;
      template <class C> static __host__ __device__ void radix(const Nodes<C> &n, int (&r)[ND])
      {
        r[0] = n.nx;
        r[1] = r[2] = n.na;
      }
      template <class C> static __host__ __device__ __forceinline__ C element(C q, C k)
      {
        return (q / C(2) * powr<2>(k)) / powr<3>(two_pi<C>);
      }
      template <class NT, class KERNEL, class C, class... T>
      static __device__ __forceinline__ NT term(const Nodes<C> &n, C q, C rf, const int (&g)[ND], C k, const T &...t)
      {
        const AngNode<C> a = gld(n.ang + g[1]), ph = gld(n.ang + n.na + g[2]);
        return NT((rf * (a.a * ph.a)) * KERNEL::kernel(q, a.c, ph.c, k, t...));
      }
    };

    // Zero-temperature frequency integrals over q0 in [-q0_ext, q0_ext] as f(q0) + f(-q0) on a
    // Gauss-Legendre rule on [0, 1] mapped to [0, q0_ext] (table {x q0_ext, w q0_ext}, nb nodes).
    // PX0: the rule is in x0 = q0/k (table {x x0_ext, w x0_ext}), so q0 = k x0 and the measure
    // gains a factor k.
    // Integrator2Dpq0GPU (integrator_2Dpq0_gpu.hh:21-46), Integrator2Dpx0GPU (integrator_2Dpx0_gpu.hh:
    // 28-54): (q, cos, q0) in 2+1 dimensions; tables cos, then q0
    template <bool PX0> struct Angle1Q0 {
      static constexpr int ND = 3, NA = 2;
      template <class C> static __host__ __device__ void radix(const Nodes<C> &n, int (&r)[ND])
      {
        r[0] = n.nx;
        r[1] = n.na;
        r[2] = n.nb;
      }
      template <class C> static __host__ __device__ __forceinline__ C element(C, C k)
      {
        const C e = (powr<2>(k) / C(2)) * two_pi<C> * (C(1) / C(2)) / powr<3>(two_pi<C>);
        return PX0 ? e * k : e;
      }
      template <class NT, class KERNEL, class C, class... T>
      static __device__ __forceinline__ NT term(const Nodes<C> &n, C q, C rf, const int (&g)[ND], C k, const T &...t)
      {
        const AngNode<C> a = gld(n.ang + g[1]), z = gld(n.ang + n.na + g[2]);
        const C q0 = PX0 ? k * z.c : z.c;
        return NT((rf * (a.a * z.a)) * (KERNEL::kernel(q, a.c, q0, k, t...) + KERNEL::kernel(q, a.c, -q0, k, t...)));
      }
    };
    // Integrator3Dpq0GPU (integrator_3Dpq0_gpu.hh:21-46), Integrator3Dpx0GPU (integrator_3Dpx0_gpu.hh:
    // 21-47): (q, cos, phi, q0) in 3+1 dimensions; tables cos, phi, then q0
    template <bool PX0> struct Angle1PhiQ0 {
      static constexpr int ND = 4, NA = 3;
      template <class C> static __host__ __device__ void radix(const Nodes<C> &n, int (&r)[ND])
      {
        r[0] = n.nx;
        r[1] = r[2] = n.na;
        r[3] = n.nb;
      }
      template <class C> static __host__ __device__ __forceinline__ C element(C q, C k)
      {
        const C e = (q / C(2) * powr<2>(k)) / powr<4>(two_pi<C>);
        return PX0 ? e * k : e;
      }
      template <class NT, class KERNEL, class C, class... T>
      static __device__ __forceinline__ NT term(const Nodes<C> &n, C q, C rf, const int (&g)[ND], C k, const T &...t)
      {
        const AngNode<C> a = gld(n.ang + g[1]), ph = gld(n.ang + n.na + g[2]), z = gld(n.ang + 2 * n.na + g[3]);
        const C q0 = PX0 ? k * z.c : z.c;
        return NT((rf * (a.a * ph.a * z.a)) *
                  (KERNEL::kernel(q, a.c, ph.c, q0, k, t...) + KERNEL::kernel(q, a.c, ph.c, -q0, k, t...)));
      }
    };

    // Matsubara sum T sum_n f(q0_n) = T f(0) + sum_i w_i (f(x_i) + f(-x_i)) (matsubara.hh:102-108)
    // IntegratorFiniteTq0GPU<d> (integrator_finiteTq0_gpu.hh:20-44): (q, n)
    template <int d> struct FiniteT {
      static constexpr int ND = 2, NA = 0;
      template <class C> static __host__ __device__ void radix(const Nodes<C> &n, int (&r)[ND])
      {
        r[0] = n.nx;
        r[1] = n.nm;
      }
      template <class C> static __host__ __device__ __forceinline__ C element(C q, C k)
      {
        return S_d_prec<C>(d - 1) * (powr<d - 3>(q) / C(2) * powr<2>(k)) / powr<d - 1>(two_pi<C>);
      }
      template <class NT, class KERNEL, class C, class... T>
      static __device__ __forceinline__ NT term(const Nodes<C> &n, C q, C rf, const int (&g)[ND], C k, const T &...t)
      {
        const C q0 = gld(n.mx + g[1]);
        NT r = NT((rf * gld(n.mw + g[1])) * (KERNEL::kernel(q, q0, k, t...) + KERNEL::kernel(q, -q0, k, t...)));
        if (n.m_T > 0 && g[1] == 0) r += NT((rf * n.m_T) * KERNEL::kernel(q, C(0), k, t...));
        return r;
      }
    };

    // IntegratorAngleFiniteTq0GPU<d> (integrator_angle_finiteTq0_gpu.hh:20-50): (q, cos, n)
    template <int d> struct Angle1FiniteT {
      static constexpr int ND = 3, NA = 1;
      template <class C> static __host__ __device__ void radix(const Nodes<C> &n, int (&r)[ND])
      {
        r[0] = n.nx;
        r[1] = n.na;
        r[2] = n.nm;
      }
      template <class C> static __host__ __device__ __forceinline__ C element(C q, C k)
      {
        return S_d_prec<C>(d - 1) * (powr<d - 3>(q) / C(2) * powr<2>(k)) * (C(1) / C(2)) / powr<d - 1>(two_pi<C>);
      }
      template <class NT, class KERNEL, class C, class... T>
      static __device__ __forceinline__ NT term(const Nodes<C> &n, C q, C rf, const int (&g)[ND], C k, const T &...t)
      {
        const AngNode<C> a = gld(n.ang + g[1]);
        const C q0 = gld(n.mx + g[2]), ra = rf * a.a;
        NT r = NT((ra * gld(n.mw + g[2])) *
                  (KERNEL::kernel(q, a.c, q0, k, t...) + KERNEL::kernel(q, a.c, -q0, k, t...)));
        if (n.m_T > 0 && g[2] == 0) r += NT((ra * n.m_T) * KERNEL::kernel(q, a.c, C(0), k, t...));
        return r;
      }
    };

    // Integrator4DFiniteTq0GPU (integrator_4D_finiteTq0_gpu.hh:20-55): (q, cos, phi, n)
    struct Angle2FiniteT {
      static constexpr int ND = 4, NA = 2;
      template <class C> static __host__ __device__ void radix(const Nodes<C> &n, int (&r)[ND])
      {
        r[0] = n.nx;
        r[1] = r[2] = n.na;
        r[3] = n.nm;
      }
      template <class C> static __host__ __device__ __forceinline__ C element(C q, C k)
      {
        return (q / C(2) * powr<2>(k)) / powr<3>(two_pi<C>);
      }
      template <class NT, class KERNEL, class C, class... T>
      static __device__ __forceinline__ NT term(const Nodes<C> &n, C q, C rf, const int (&g)[ND], C k, const T &...t)
      {
        const AngNode<C> a = gld(n.ang + g[1]), ap = gld(n.ang + n.na + g[2]);
        const C q0 = gld(n.mx + g[3]), ra = rf * (a.a * ap.a);
        NT r = NT((ra * gld(n.mw + g[3])) *
                  (KERNEL::kernel(q, a.c, ap.c, q0, k, t...) + KERNEL::kernel(q, a.c, ap.c, -q0, k, t...)));
        if (n.m_T > 0 && g[3] == 0) r += NT((ra * n.m_T) * KERNEL::kernel(q, a.c, ap.c, C(0), k, t...));
        return r;
      }
    };

    // Cartesian momentum integrals over a box, int d^Dq/(2 pi)^D f(qx[, qy[, qz]]): every axis is a
    // Gauss-Legendre rule on [0,1] mapped linearly, q = q_min + q_ext x with weight q_ext w.  The
    // first axis takes the radial digit's place (Nodes::q_min, q_ext); qy and qz come from tables
    // (coordinate, weight) of ny = na and nz = nb entries, qz's after qy's.
    // Integrator1DCartesianGPU (integrator_1D_cartesian_gpu.hh:36-60): (qx)
    struct Cartesian1 {
      static constexpr int ND = 1, NA = 0;
      static constexpr bool cartesian = true;
      template <class C> static __host__ __device__ void radix(const Nodes<C> &n, int (&r)[ND]) { r[0] = n.nx; }
      template <class C> static __host__ __device__ __forceinline__ C element(C, C) { return C(1) / two_pi<C>; }
      template <class NT, class KERNEL, class C, class... T>
      static __device__ __forceinline__ NT term(const Nodes<C> &, C q, C rf, const int (&)[ND], C k, const T &...t)
      {
        return NT(rf * KERNEL::kernel(q, k, t...));
      }
    };
    // Integrator2DCartesianGPU (integrator_2D_cartesian_gpu.hh:21-45): (qx, qy)
    struct Cartesian2 {
      static constexpr int ND = 2, NA = 1;
      static constexpr bool cartesian = true;
      template <class C> static __host__ __device__ void radix(const Nodes<C> &n, int (&r)[ND])
      {
        r[0] = n.nx;
        r[1] = n.na;
      }
      template <class C> static __host__ __device__ __forceinline__ C element(C, C)
      {
        return C(1) / powr<2>(two_pi<C>);
      }
      template <class NT, class KERNEL, class C, class... T>
      static __device__ __forceinline__ NT term(const Nodes<C> &n, C q, C rf, const int (&g)[ND], C k, const T &...t)
      {
        const AngNode<C> y = gld(n.ang + g[1]);
        return NT((rf * y.a) * KERNEL::kernel(q, y.c, k, t...));
      }
    };
    // Integrator3DCartesianGPU (integrator_3D_cartesian_gpu.hh:20-46): (qx, qy, qz)
    struct Cartesian3 {
      static constexpr int ND = 3, NA = 2;
      static constexpr bool cartesian = true;
      template <class C> static __host__ __device__ void radix(const Nodes<C> &n, int (&r)[ND])
      {
        r[0] = n.nx;
        r[1] = n.na;
        r[2] = n.nb;
      }
      template <class C> static __host__ __device__ __forceinline__ C element(C, C)
      {
        return C(1) / powr<3>(two_pi<C>);
      }
      template <class NT, class KERNEL, class C, class... T>
      static __device__ __forceinline__ NT term(const Nodes<C> &n, C q, C rf, const int (&g)[ND], C k, const T &...t)
      {
        const AngNode<C> y = gld(n.ang + g[1]), z = gld(n.ang + n.na + g[2]);
        return NT((rf * (y.a * z.a)) * KERNEL::kernel(q, y.c, z.c, k, t...));
      }
    };

    // Radial node ix of shape S: its momentum q and factor rf (element times weight).  Cartesian
    // shapes map their first axis linearly; the others take q = k sqrt(x x_ext) and the factor
    // element(q, k) w x_ext.
    template <class S, class C>
    __device__ __forceinline__ void radial_node(const Nodes<C> &n, int ix, C k, C &q, C &rf)
    {
      if constexpr (requires { S::cartesian; }) {
        q = n.q_min + n.q_ext * n.xq[ix];
        rf = S::element(q, k) * (n.q_ext * n.wq[ix]);
      } else {
        q = k * sqrt(n.xq[ix] * n.x_extent);
        rf = S::element(q, k) * (n.wq[ix] * n.x_extent);
      }
    }

    // A Cartesian axis table {q_min + q_ext x_i, q_ext w_i} from the rule on [0,1].
    template <class ctype>
    std::vector<AngNode<ctype>> cartesian_table(const std::vector<ctype> &x, const std::vector<ctype> &w, ctype q_min,
                                                ctype q_ext)
    {
      std::vector<AngNode<ctype>> t(x.size());
      for (size_t i = 0; i < x.size(); ++i) t[i] = {q_min + q_ext * x[i], q_ext * w[i]};
      return t;
    }

    // Finite temperature in x0 = q0/k: the first nm Matsubara frequencies q0 = 2 pi n T as a sum
    // (T f(0) + T sum_{0<n<nm} (f(q0_n) + f(-q0_n))), the rest of the frequency axis as an
    // integral over x0 from 2 pi nm T / k to x0_extent on a Gauss-Legendre rule in log x0 (nb nodes
    // in mx, mw), f(q0) + f(-q0).  At T = 0 (nm = 0) the log map starts at 1e-3, shifted back, with
    // the reference's weight w log_ext q0 / k.  Digits (q, [cos, [phi,]] n), the frequency index
    // fastest, n < nm the sum.  Nodes: m_T = T, nm summands, q_ext = x0_extent.
    // KIND 0: IntegratorFiniteTx0GPU<d> (integrator_finiteTx0_gpu.hh:21-62), (q, q0);
    // KIND 1: IntegratorAngleFiniteTx0GPU<d> (integrator_angle_finiteTx0_gpu.hh:21-69), (q, cos, q0);
    // KIND 2: Integrator4DFiniteTx0GPU (integrator_4D_finiteTx0_gpu.hh:21-74), (q, cos, phi, q0).
    template <int d, int KIND> struct FiniteTx0 {
      static constexpr int ND = 2 + KIND, NA = KIND;
      template <class C> static __host__ __device__ void radix(const Nodes<C> &n, int (&r)[ND])
      {
        r[0] = n.nx;
        for (int i = 1; i <= KIND; ++i) r[i] = n.na;
        r[ND - 1] = n.nm + n.nb;
      }
      // the part common to the sum and the integral: S_{d-1} q^{d-3}/2 k^2 (times 1/2 for the cos
      // average of KIND 1; KIND 2 has its angles explicit)
      template <class C> static __host__ __device__ __forceinline__ C element(C q, C k)
      {
        const C r = powr<d - 3>(q) / C(2) * powr<2>(k);
        if constexpr (KIND == 2)
          return r;
        else if constexpr (KIND == 1)
          return S_d_prec<C>(d - 1) * r * (C(1) / C(2));
        else
          return S_d_prec<C>(d - 1) * r;
      }
      template <class NT, class KERNEL, class C, class... T>
      static __device__ __forceinline__ NT term(const Nodes<C> &n, C q, C rf, const int (&g)[ND], C k, const T &...t)
      {
        C af = C(1), c1 = C(0), ph = C(0);
        if constexpr (KIND >= 1) {
          const AngNode<C> a = gld(n.ang + g[1]);
          af = a.a;
          c1 = a.c;
        }
        if constexpr (KIND == 2) {
          const AngNode<C> p = gld(n.ang + n.na + g[2]);
          af *= p.a;
          ph = p.c;
        }
        auto K = [&](C q0) {
          if constexpr (KIND == 0)
            return KERNEL::kernel(q, q0, k, t...);
          else if constexpr (KIND == 1)
            return KERNEL::kernel(q, c1, q0, k, t...);
          else
            return KERNEL::kernel(q, c1, ph, q0, k, t...);
        };
        const int y = g[ND - 1];
        if (y < n.nm) {
          const C f = (rf * af) * (n.m_T / powr<d - 1>(two_pi<C>));
          if (y == 0) return NT(f * K(C(0)));
          const C q0 = two_pi<C> * n.m_T * C(y);
          return NT(f * (K(q0) + K(-q0)));
        }
        const C z = n.m_T == C(0) ? C(1e-3) : C(0);
        const C start = (C(2) * C(n.nm) * C(3.14159265358979323846) * n.m_T) / k;
        const C ls = log(start + z), le = log(n.q_ext / (start + z));
        const int j = y - n.nm;
        const C q0 = k * (exp(ls + le * gld(n.mx + j)) - z);
        const C w = gld(n.mw + j) * le * q0 / k;
        return NT(((rf * af) * (k / powr<d>(two_pi<C>)) * w) * (K(q0) + K(-q0)));
      }
    };

    // Angular tables of a shape from the Gauss-Legendre rule on [0,1] of order na (host nodes):
    //   cos = 2(x - 1/2), weight 2w; cos1 of the 4D forms also carries the Jacobian sqrt(1 - cos1^2);
    //   phi = 2 pi x, weight 2 pi w.
    enum class AngKind { cos, cos_jacobian, phi };
    template <class ctype>
    std::vector<AngNode<ctype>> angle_table(const std::vector<ctype> &x, const std::vector<ctype> &w, AngKind kind)
    {
      std::vector<AngNode<ctype>> t(x.size());
      for (size_t i = 0; i < x.size(); ++i) {
        if (kind == AngKind::phi) {
          t[i] = {two_pi<ctype> * x[i], two_pi<ctype> * w[i]};
        } else {
          const ctype c = 2 * (x[i] - ctype(0.5));
          t[i] = {c, kind == AngKind::cos_jacobian ? 2 * w[i] * std::sqrt(1 - c * c) : 2 * w[i]};
        }
      }
      return t;
    }

    // The tables of several angular digits, one after the other (na entries each).
    template <class ctype>
    std::vector<AngNode<ctype>> angle_tables(const std::vector<ctype> &x, const std::vector<ctype> &w,
                                             std::initializer_list<AngKind> kinds)
    {
      std::vector<AngNode<ctype>> t;
      for (AngKind k : kinds) {
        const auto one = angle_table(x, w, k);
        t.insert(t.end(), one.begin(), one.end());
      }
      return t;
    }

    // A device copy of the angular tables, shared by the copies of an integrator.
    template <class ctype> std::shared_ptr<AngNode<ctype>> upload_angles(const std::vector<AngNode<ctype>> &t)
    {
      void *p = nullptr;
      check(dfrg_malloc(&p, sizeof(AngNode<ctype>) * std::max<size_t>(t.size(), 1)), "angular node table");
      std::shared_ptr<AngNode<ctype>> d((AngNode<ctype> *)p, [](AngNode<ctype> *q) { dfrg_free(q); });
      if (!t.empty()) {
        check(dfrg_memcpy_h2d(p, t.data(), sizeof(AngNode<ctype>) * t.size(), nullptr), "angular node table");
        check(dfrg_stream_sync(nullptr), "angular node table");
      }
      return d;
    }

    // ------------------------------------------------------------------------------------------
    // the unit plan.  An integral's flat node space (radial digit slowest, as the reference's flat
    // loops) is cut into units that depend on its node counts only:
    //   rows of ni >= kRowMin inner nodes: one unit per radial node (rows of more than kMaxUnit
    //     nodes in equal parts).  The lanes walk the row's inner nodes with q fixed, so every
    //     subexpression of the KERNEL that depends on q (and the arguments) only is loop-invariant:
    //     after inlining the compiler evaluates it once per lane and row, not per node;
    //   rows of fewer inner nodes: units of `rows` whole rows (about kRowMin nodes), their radial
    //     factors staged in LDS, walked flat.
    // A unit has a fixed block size and lane assignment and its own partial slot, and the integral
    // is the sum of its unit partials by a tree fixed by the unit count, so how units are grouped
    // into launches and blocks does not enter the result.
    // ------------------------------------------------------------------------------------------
    constexpr int kRowMin = 256, kMaxUnit = 65536;
    constexpr int kRowNodesPerLane = 16;

    struct Plan {
      int nx = 1, ni = 1;
      int row_mode = 0;
      int rows = 1;  // radial nodes per unit (flat mode)
      int parts = 1; // units per row (row mode)
      int unit = 1;  // nodes per unit (row mode: inner nodes)
      int upr = 1;   // units per integral
      int block = 64;
    };

    inline int device_cus()
    {
      static const int cus = [] {
        int c = 0;
        if (dfrg_device_props(nullptr, 0, &c, nullptr, nullptr) != DFRG_OK || c <= 0) c = 256;
        return c;
      }();
      return cus;
    }

    inline Plan make_plan(int nx, int ni)
    {
      Plan p;
      p.nx = nx;
      p.ni = ni;
      if (ni >= kRowMin) {
        p.row_mode = 1;
        p.parts = (ni + kMaxUnit - 1) / kMaxUnit;
        p.unit = (ni + p.parts - 1) / p.parts;
        p.upr = nx * p.parts;
        int b = 64; // at least kRowNodesPerLane nodes per lane
        while (b < 256 && 2 * b * kRowNodesPerLane <= p.unit) b *= 2;
        p.block = b;
      } else {
        p.rows = std::min({(kRowMin + ni - 1) / ni, kMaxQ, nx});
        p.unit = p.rows * ni;
        p.upr = (nx + p.rows - 1) / p.rows;
        p.block = 64;
      }
      return p;
    }

    template <class S, class C> Plan plan_of(const Nodes<C> &n)
    {
      int rad[S::ND];
      S::radix(n, rad);
      long long ni = 1;
      for (int i = 1; i < S::ND; ++i) ni *= rad[i];
      if (ni > (1LL << 30) || rad[0] < 1 || ni < 1) throw std::invalid_argument("integrator: node counts out of range");
      return make_plan(rad[0], (int)ni);
    }

    // The mixed-radix digits of a flat node index (slowest first), advanced by a fixed stride S with
    // carries instead of a division per node.
    // I: the index type (int where the flat index of a unit fits, as in row mode: 32-bit divisions)
    template <int N, class I = long long> struct NodeWalk {
      int dig[N], step[N], rad[N];
      __device__ __forceinline__ NodeWalk(I i, I S, const int *r)
      {
#pragma unroll
        for (int j = N - 1; j >= 1; --j) {
          rad[j] = r[j];
          dig[j] = (int)(i % r[j]);
          i /= r[j];
          step[j] = (int)(S % r[j]);
          S /= r[j];
        }
        rad[0] = r[0];
        dig[0] = (int)i;
        step[0] = (int)S;
      }
      __device__ __forceinline__ void next()
      {
        int carry = 0;
#pragma unroll
        for (int j = N - 1; j >= 1; --j) {
          const int v = dig[j] + step[j] + carry; // < 2 rad[j]
          carry = v >= rad[j];
          dig[j] = carry ? v - rad[j] : v;
        }
        dig[0] += step[0] + carry;
      }
    };

    // ------------------------------------------------------------------------------------------
    // one unit u of integral r, then the block's share of the cross-unit sum
    // ------------------------------------------------------------------------------------------
    constexpr int kMaxBlock = 256;

    template <class S, class NT, class KERNEL, bool ROW, class C, class Args>
    __device__ __forceinline__ void unit_body(const Nodes<C> &n, const Plan &p, int r, int u, const Args &args,
                                              double *partials, unsigned *counters, NT *results)
    {
      constexpr int NC = Acc<NT>::N;
      __shared__ C s_q[kMaxQ], s_rf[kMaxQ];
      __shared__ double s_red[NC][kMaxBlock / 64];
      __shared__ bool s_last;
      const int B = (int)blockDim.x, tid = (int)threadIdx.x;
      const C k = std::get<0>(args);
      int rad[S::ND];
      S::radix(n, rad);
      Acc<NT> acc;
      if constexpr (S::ND > 1 && ROW) {
        {
          const int ix = u / p.parts, part = u - ix * p.parts;
          const int m0 = part * p.unit, m1 = m0 + p.unit < p.ni ? m0 + p.unit : p.ni;
          C q, rf;
          radial_node<S>(n, ix, k, q, rf);
          auto eval = [&](const NodeWalk<S::ND - 1, int> &w) {
            int g[S::ND];
            g[0] = ix;
#pragma unroll
            for (int i = 1; i < S::ND; ++i) g[i] = w.dig[i - 1];
            return std::apply(
                [&](const C &kk, const auto &...t) { return S::template term<NT, KERNEL>(n, q, rf, g, kk, t...); },
                args);
          };
          NodeWalk<S::ND - 1, int> w(m0 + tid, B, rad + 1);
          // two nodes per iteration: their evaluations are independent chains the scheduler can
          // interleave (the sum keeps node order, so results do not change)
#pragma unroll 2
          for (int m = m0 + tid; m < m1; m += B, w.next()) acc.add(eval(w));
        }
      }
      if constexpr (!ROW || S::ND == 1) {
        const int ix0 = u * p.rows, ix1 = ix0 + p.rows < p.nx ? ix0 + p.rows : p.nx;
        for (int j = tid; j < ix1 - ix0; j += B) {
          C q, rf;
          radial_node<S>(n, ix0 + j, k, q, rf);
          s_q[j] = q;
          s_rf[j] = rf;
        }
        __syncthreads();
        const long long f1 = (long long)ix1 * p.ni;
        NodeWalk<S::ND> w((long long)ix0 * p.ni + tid, B, rad);
        for (long long f = (long long)ix0 * p.ni + tid; f < f1; f += B, w.next()) {
          const int j = w.dig[0] - ix0;
          acc.add(std::apply(
              [&](const C &kk, const auto &...t) {
                return S::template term<NT, KERNEL>(n, s_q[j], s_rf[j], w.dig, kk, t...);
              },
              args));
        }
      }

      // block: wave sums, then wave 0's totals in wave order
      const int lane = tid & 63, wid = tid >> 6;
#pragma unroll
      for (int i = 0; i < NC; ++i) {
        const double v = dev::wave_sum(acc.v[i]);
        if (lane == 0) s_red[i][wid] = v;
      }
      __syncthreads();
      double tot[NC];
#pragma unroll
      for (int i = 0; i < NC; ++i) {
        tot[i] = s_red[i][0];
        for (int j = 1; j < B / 64; ++j) tot[i] += s_red[i][j];
      }
      if (p.upr == 1) {
        if (tid == 0) results[r] = Acc<NT>::value(tot);
      } else {
        // units of the integral: publish, ticket; the last block adds the partials (fixed tree)
        // and resets the counter
        double *slots = partials + (long long)r * p.upr * dev::kSlotStride;
        unsigned *counter = counters + (long long)r * dev::kCounterStride;
        if (tid == 0) {
#pragma unroll
          for (int i = 0; i < NC; ++i) dev::publish(&slots[(long long)u * dev::kSlotStride + i], tot[i]);
          dev::wait_stores();
          s_last = dev::take_ticket(counter) == (unsigned)(p.upr - 1);
        }
        __syncthreads();
        if (s_last && tid < 64) {
          // 64 partials per round, one per lane, added by the fixed wave tree (the rounds in unit
          // order): a few dependent steps instead of a serial chain over the lanes, which had cost
          // the launch microseconds at the 64 units of a row-mode integral
          double sum[NC];
#pragma unroll
          for (int i = 0; i < NC; ++i) sum[i] = 0.0;
          for (int c0 = 0; c0 < p.upr; c0 += 64) {
            const int jj = c0 + tid;
#pragma unroll
            for (int i = 0; i < NC; ++i) {
              const double v = jj < p.upr ? dev::fetch(&slots[(long long)jj * dev::kSlotStride + i]) : 0.0;
              sum[i] += dev::wave_sum(v);
            }
          }
          if (tid == 0) {
            results[r] = Acc<NT>::value(sum);
            dev::reset_counter(counter);
          }
        }
      }
      __syncthreads(); // LDS is free for the block's next unit
    }

    // a batch of integrals whose (k, t...) tuples sit in HBM; blocks stride over all units
    // ROW: the plan's row mode (one instantiation per mode, so each has its own register budget)
    template <class S, class NT, class KERNEL, bool ROW, class C, class Args>
    __global__ __launch_bounds__(kMaxBlock) void batch_kernel(
        Nodes<C> n, Plan p, const Args *args, int n_req, double *partials, unsigned *counters, NT *results)
    {
      const long long total = (long long)n_req * p.upr;
      for (long long g = blockIdx.x; g < total; g += gridDim.x) {
        const int r = (int)(g / p.upr), u = (int)(g - (long long)r * p.upr);
        unit_body<S, NT, KERNEL, ROW>(n, p, r, u, args[r], partials, counters, results);
      }
    }

    // Blocks of `block` threads one launch keeps resident on the device: at most that many are
    // launched, each striding over the units.
    template <class F> int resident_blocks(F kernel, int block)
    {
      int per_cu = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
      return per_cu * device_cus();
    }
    // the same, queried once per kernel and block size (64, 128 or 256)
    template <auto kernel> int resident_cached(int block)
    {
      static std::atomic<int> cache[3] = {0, 0, 0};
      const int i = block == 64 ? 0 : block == 128 ? 1 : 2;
      int v = cache[i].load(std::memory_order_relaxed);
      if (v == 0) {
        v = resident_blocks(kernel, block);
        cache[i].store(v, std::memory_order_relaxed);
      }
      return v;
    }

    // ------------------------------------------------------------------------------------------
    // host: the thread's scratch arena
    // ------------------------------------------------------------------------------------------
    class Arena
    {
      struct Block {
        void *d = nullptr;
        size_t cap_req = 0, cap_data = 0;
        hipStream_t s = nullptr; // the stream that last used it
      };
      // Arenas of threads that have ended.  A thread's destructor makes no HIP call (it may run at
      // process exit); the next arena that needs memory takes one of these instead, after waiting
      // for the stream that used it.
      static std::mutex &orphan_mutex()
      {
        static std::mutex m;
        return m;
      }
      static std::vector<Block> &orphans()
      {
        static std::vector<Block> v;
        return v;
      }

    public:
      ~Arena()
      {
        if (b.d) {
          std::lock_guard<std::mutex> lock(orphan_mutex());
          orphans().push_back(b);
        }
      }
      // Room for n_req counters and `data` bytes; counters are zero at the start of every launch.
      void ensure(size_t n_req, size_t data)
      {
        if (!s) {
          void *sv = nullptr;
          check(dfrg_thread_stream(&sv), "thread stream");
          s = (hipStream_t)sv;
        }
        if (b.d && n_req <= b.cap_req && data <= b.cap_data) return;
        const size_t want_req = std::max<size_t>({n_req, 2 * b.cap_req, 64});
        const size_t want_data = std::max<size_t>({data, 2 * b.cap_data, size_t(1) << 16});
        if (b.d) {
          if (hipStreamSynchronize(s) != hipSuccess) throw std::runtime_error("integrator: stream error");
          dfrg_pool_free(b.d);
          b.d = nullptr;
        }
        {
          std::lock_guard<std::mutex> lock(orphan_mutex());
          auto &o = orphans();
          for (size_t i = 0; i < o.size(); ++i)
            if (o[i].cap_req >= n_req && o[i].cap_data >= data) {
              b = o[i];
              o.erase(o.begin() + (long)i);
              break;
            }
        }
        if (b.d) {
          if (hipStreamSynchronize(b.s) != hipSuccess) throw std::runtime_error("integrator: stream error");
        } else {
          b.cap_req = want_req;
          b.cap_data = want_data;
          check(dfrg_pool_alloc(&b.d, counter_bytes() + b.cap_data), "integrator scratch");
        }
        b.s = s;
        if (hipMemsetAsync(b.d, 0, counter_bytes(), s) != hipSuccess)
          throw std::runtime_error("integrator: scratch initialisation failed");
      }
      unsigned *counters() const { return (unsigned *)b.d; }
      char *data() const { return (char *)b.d + counter_bytes(); }
      hipStream_t stream() const { return s; }

    private:
      size_t counter_bytes() const { return b.cap_req * dev::kCounterStride * sizeof(unsigned); }
      Block b;
      hipStream_t s = nullptr;
    };
    inline Arena &thread_arena()
    {
      thread_local Arena a;
      return a;
    }
    inline size_t up256(size_t b) { return (b + 255) / 256 * 256; }

    template <class S, class NT, class KERNEL, class C, class Args>
    void launch_batch(const Nodes<C> &n, const Plan &p, const Args *d_args, int n_req, double *partials,
                      unsigned *counters, NT *results, hipStream_t s)
    {
      const long long total = (long long)n_req * p.upr;
      auto go = [&](auto kern, int resident) {
        const int grid = (int)std::min<long long>(total, resident);
        hipLaunchKernelGGL(kern, dim3(grid), dim3(p.block), 0, s, n, p, d_args, n_req, partials, counters, results);
      };
      if (p.row_mode && S::ND > 1) {
        constexpr auto kern = &batch_kernel<S, NT, KERNEL, (S::ND > 1), C, Args>;
        go(kern, resident_cached<kern>(p.block));
      } else {
        constexpr auto kern = &batch_kernel<S, NT, KERNEL, false, C, Args>;
        go(kern, resident_cached<kern>(p.block));
      }
    }

    // ------------------------------------------------------------------------------------------
    // batches of requests
    // ------------------------------------------------------------------------------------------
    constexpr size_t kMaxBatch = size_t(1) << 16;

    struct BatchBase {
      virtual ~BatchBase() = default;
      virtual void flush() = 0;
      const void *type = nullptr; // identifies the derived batch type (see batch_type)
    };
    // A distinct address per batch type, so the open batch's type is checked without RTTI.
    template <class B> inline const void *batch_type()
    {
      static const char tag = 0;
      return &tag;
    }

    // Launch mode of request().  Off (the default): requests collect into batches, launched when a
    // future of the thread is read, when request_data() ends, or when the batch is full.  On: every
    // request is launched at once as a batch of one, as the reference's request() launches its
    // kernel at the call (integrator_4D_2ang_gpu.hh:170-197): the GPU starts with the first request
    // and overlaps the caller's host work from there, at one launch per request.  The values are
    // the same bit for bit either way.  Process-wide; starts from DFRG_LAUNCH_AT_REQUEST=1 in the
    // environment.
    inline std::atomic<bool> &launch_at_request_flag()
    {
      static std::atomic<bool> f{[] {
        const char *e = std::getenv("DFRG_LAUNCH_AT_REQUEST");
        return e != nullptr && e[0] == '1' && e[1] == '\0';
      }()};
      return f;
    }
    inline void set_launch_at_request(bool on) { launch_at_request_flag().store(on, std::memory_order_relaxed); }
    inline bool launch_at_request() { return launch_at_request_flag().load(std::memory_order_relaxed); }

    // Batches opened on this thread that are not launched yet.
    inline std::vector<std::weak_ptr<BatchBase>> &thread_open()
    {
      thread_local std::vector<std::weak_ptr<BatchBase>> v;
      return v;
    }
    // Launch every open batch of the calling thread (in the order they were opened).
    inline void flush_thread()
    {
      auto &v = thread_open();
      if (v.empty()) return;
      std::vector<std::weak_ptr<BatchBase>> w;
      w.swap(v);
      for (auto &x : w)
        if (auto b = x.lock()) b->flush();
    }

    template <class NT, class Args> struct RequestBatch : BatchBase {
      using Launch = std::function<void(const Args *d_args, int n, const Plan &p, double *partials,
                                        unsigned *counters, NT *results, hipStream_t s)>;
      std::mutex m;
      std::vector<Args> args;
      std::vector<NT> constants, results;
      Launch launch;
      Plan plan;
      bool launched = false, ready = false;
      std::exception_ptr error;
      hipEvent_t ev = nullptr;
      hipStream_t stream = nullptr; // where the batch was enqueued
      void *h = nullptr;            // pinned: tuples, then results (stored by the kernel)

      ~RequestBatch() override
      {
        // the staging returns to the pool only after the work that writes into it has finished
        if (ev) {
          (void)hipEventSynchronize(ev);
          (void)hipEventDestroy(ev);
        } else if (stream) {
          (void)hipStreamSynchronize(stream);
        }
        if (h) dfrg_pool_host_free(h);
      }

      void flush() override
      {
        std::lock_guard<std::mutex> lock(m);
        if (!launched) start();
      }

      NT get(size_t i)
      {
        flush_thread();
        std::lock_guard<std::mutex> lock(m);
        if (!launched) start();
        if (!ready && !error) {
          if (hipEventSynchronize(ev) != hipSuccess)
            error = std::make_exception_ptr(std::runtime_error("integrator: batched request failed"));
          else {
            results.resize(constants.size());
            std::memcpy((void *)results.data(), (char *)h + up256(sizeof(Args) * constants.size()),
                        sizeof(NT) * constants.size());
            ready = true;
          }
        }
        if (error) std::rethrow_exception(error);
        return constants[i] + results[i];
      }

    private:
      // Enqueue on this thread's stream: the kernel, then the event.  The tuples and the results
      // share the pinned staging: the kernel reads each unit's tuple over the host link instead of
      // waiting for a copy to HBM (a copy costs the host ~3.5 us to enqueue and the device a blit
      // launch, more than the units' reads add to the kernel), and its finishing blocks store the
      // results there.
      void start()
      {
        launched = true;
        const int n = (int)args.size();
        try {
          const Plan &p = plan;
          const size_t bytes = up256(sizeof(double) * dev::kSlotStride * (size_t)n * (p.upr > 1 ? p.upr : 0));
          Arena &A = thread_arena();
          A.ensure((size_t)n, bytes);
          const hipStream_t s = A.stream();
          const size_t h_res = up256(sizeof(Args) * n);
          check(dfrg_pool_host_alloc(&h, h_res + sizeof(NT) * n), "batch staging");
          std::memcpy(h, (const void *)args.data(), sizeof(Args) * n);
          void *dh = nullptr; // the staging's device address
          stream = s;
          bool ok = hipHostGetDevicePointer(&dh, h, 0) == hipSuccess;
          if (ok) {
            launch((const Args *)dh, n, p, (double *)A.data(), A.counters(), (NT *)((char *)dh + h_res), s);
            ok = hipGetLastError() == hipSuccess;
            if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
              ev = nullptr;
              ok = false;
            } else if (hipEventRecord(ev, s) != hipSuccess) { // the destructor then waits on the stream
              (void)hipEventDestroy(ev);
              ev = nullptr;
              ok = false;
            }
          }
          if (!ok) throw std::runtime_error("integrator: batched request launch failed");
        } catch (...) {
          error = std::current_exception();
        }
        args.clear();
        args.shrink_to_fit();
      }
    };

    // Collects the requests of one integrator into batches.  A new batch opens when the argument
    // types or the integrator state the launch depends on (`state`, e.g. the Matsubara rule)
    // change, when the open batch has been launched, or at kMaxBatch requests (the full batch is
    // launched at once).
    template <class NT> class RequestBatcher
    {
    public:
      template <class Args, class MakeLaunch>
      std::future<NT> enqueue(unsigned long long state, const Plan &plan, NT constant, Args &&args,
                              MakeLaunch &&make_launch)
      {
        using B = RequestBatch<NT, std::decay_t<Args>>;
        std::shared_ptr<B> b;
        size_t idx = 0;
        bool full = false;
        {
          std::lock_guard<std::mutex> lock(m);
          {
            std::shared_ptr<BatchBase> o = open.lock();
            if (o && o->type == batch_type<B>()) b = std::static_pointer_cast<B>(o);
          }
          for (int attempt = 0;; ++attempt) {
            if (!b || state != open_state || attempt > 0) {
              b = std::make_shared<B>();
              b->type = batch_type<B>();
              b->args.reserve(64); // a request_data grid without reallocations
              b->constants.reserve(64);
              b->launch = make_launch();
              b->plan = plan;
              open = b;
              open_state = state;
              thread_open().push_back(b);
            }
            std::lock_guard<std::mutex> lb(b->m);
            if (b->launched || b->args.size() >= kMaxBatch) continue; // launched already, or full
            idx = b->args.size();
            b->args.push_back(std::forward<Args>(args));
            b->constants.push_back(constant);
            full = b->args.size() == kMaxBatch;
            break;
          }
        }
        if (full || launch_at_request()) b->flush();
        return std::async(std::launch::deferred, [b, idx]() { return b->get(idx); });
      }

    private:
      std::mutex m;
      std::weak_ptr<BatchBase> open;
      unsigned long long open_state = 0;
    };

    // ------------------------------------------------------------------------------------------
    // the integrator side: get() and request() for a shape
    // ------------------------------------------------------------------------------------------
    // get(): a batch of one, launched at once and waited for -- the same kernel as a batch of
    // requests, so get() and request() give bitwise the same value
    template <class S, class NT, class KERNEL, class C, class... T>
    NT get(const Nodes<C> &n, C k, const T &...t)
    {
      flush_thread(); // this thread's open batches go first
      using Args = std::tuple<C, T...>;
      const Plan p = plan_of<S>(n);
      const NT constant = KERNEL::constant(k, t...);
      Arena &A = thread_arena();
      A.ensure(1, up256(sizeof(double) * dev::kSlotStride * (p.upr > 1 ? p.upr : 0)));
      const hipStream_t s = A.stream();
      // pinned staging of this thread: the tuple, which the kernel reads from there, then the
      // result, which the finishing block stores through the staging's device address
      thread_local void *h = nullptr, *dh = nullptr;
      thread_local size_t h_size = 0;
      const size_t h_res = up256(sizeof(Args)), need = h_res + sizeof(NT);
      if (h_size < need) {
        if (h) dfrg_pool_host_free(h);
        h = dh = nullptr;
        h_size = 0;
        check(dfrg_pool_host_alloc(&h, need), "get() staging");
        if (hipHostGetDevicePointer(&dh, h, 0) != hipSuccess) {
          dfrg_pool_host_free(h);
          h = dh = nullptr;
          throw std::runtime_error("integrator: staging not mapped");
        }
        h_size = need;
      }
      const Args args(k, t...);
      std::memcpy(h, (const void *)&args, sizeof(Args)); // the previous get() has finished with h
      launch_batch<S, NT, KERNEL>(n, p, (const Args *)dh, 1, (double *)A.data(), A.counters(),
                                  (NT *)((char *)dh + h_res), s);
      check_hip("integrator kernel launch");
      if (hipStreamSynchronize(s) != hipSuccess) throw std::runtime_error("integrator: stream error");
      NT v;
      std::memcpy((void *)&v, (char *)h + h_res, sizeof(NT));
      return constant + v;
    }

    // hold: device memory the nodes point into that the integrator may replace (the Cartesian axis
    // tables after a range change), kept alive by the batch until it is released
    template <class S, class NT, class KERNEL, class C, class... T>
    std::future<NT> request_holding(RequestBatcher<NT> &batcher, unsigned long long state,
                                    std::shared_ptr<const void> hold, const Nodes<C> &n, C k, const T &...t)
    {
      using Args = std::tuple<C, T...>;
      return batcher.enqueue(state, plan_of<S>(n), KERNEL::constant(k, t...), Args(k, t...), [n, hold]() {
        return [n, hold](const Args *a, int r, const Plan &p, double *part, unsigned *cnt, NT *res, hipStream_t s) {
          launch_batch<S, NT, KERNEL>(n, p, a, r, part, cnt, res, s);
        };
      });
    }
    template <class S, class NT, class KERNEL, class C, class... T>
    std::future<NT> request(RequestBatcher<NT> &batcher, unsigned long long state, const Nodes<C> &n, C k,
                            const T &...t)
    {
      return request_holding<S, NT, KERNEL>(batcher, state, nullptr, n, k, t...);
    }
  } // namespace tmpl
} // namespace dfrg
