// dfrg/integrate.hh -- generic gfx950 kernels behind the DiFfRG-compatible integrator templates.
//
// A user KERNEL (static __host__ __device__ kernel()/constant()) is device code of the caller's
// translation unit, so these kernels are templates compiled there by hipcc.  Each call of get()
// or request() is one launch on the calling thread's stream: every lane evaluates a strided set
// of quadrature nodes (element * weight * KERNEL::kernel), lanes are reduced with a wave64 DPP
// tree, waves in LDS, and blocks through per-block partials whose last-arriving block sums them
// in block order.  The result (one NT) is copied into pinned host memory on the same stream.
// Per-call scratch comes from libdfrg.so's caching pools, so no call allocates after warm-up.
//
// Summation order is fixed by the grid (a function of the node count only), so a result is
// bitwise reproducible run to run.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <future>
#include <stdexcept>
#include <vector>
#include <mutex>
#include <memory>
#include <functional>
#include <tuple>
#include <type_traits>
#include <utility>

#include <DiFfRG/common/complex_math.hh>
#include <DiFfRG/common/dual_traits.hh>
#include <DiFfRG/common/hip_prefix.hh>
#include <DiFfRG/common/math.hh>

#include "dfrg/device.hh"

namespace dfrg
{
  namespace tmpl
  {
    constexpr int kBlock = 256;
    constexpr int kMaxBlocks = 1024;
    constexpr int kNodesPerLane = 4;

    // ------------------------------------------------------------------------------------------
    // reductions over the number type NT: float, double, or a two-component type reduced component
    // by component -- DiFfRG::complex<float|double> (re, im) or a dual number (value, derivative;
    // DiFfRG/common/dual_traits.hh)
    // ------------------------------------------------------------------------------------------
    template <class NT> struct Parts {
      static constexpr bool dual = DiFfRG::internal::DualTraits<NT>::is_dual;
      static __host__ __device__ __forceinline__ auto first(const NT &v)
      {
        if constexpr (dual)
          return DiFfRG::internal::DualTraits<NT>::value(v);
        else
          return v.re;
      }
      static __host__ __device__ __forceinline__ auto second(const NT &v)
      {
        if constexpr (dual)
          return DiFfRG::internal::DualTraits<NT>::derivative(v);
        else
          return v.im;
      }
      template <class P> static __host__ __device__ __forceinline__ NT make(P a, P b)
      {
        if constexpr (dual)
          return DiFfRG::internal::DualTraits<NT>::make(a, b);
        else
          return NT(a, b);
      }
      using P = decltype(first(std::declval<NT>()));
    };
    template <class NT> constexpr bool is_scalar_nt = std::is_same_v<NT, double> || std::is_same_v<NT, float>;

    __device__ __forceinline__ float wave_sum_f32(float v)
    {
      v += dev::dpp_f32<0xB1>(v);
      v += dev::dpp_f32<0x4E>(v);
      v += dev::dpp_f32<0x141>(v);
      v += dev::dpp_f32<0x140>(v);
      v += __shfl_xor(v, 16, 64);
      v += __shfl_xor(v, 32, 64);
      return v;
    }
    template <class NT> __device__ __forceinline__ NT wave_reduce(NT v)
    {
      if constexpr (std::is_same_v<NT, double>)
        return dev::wave_sum(v);
      else if constexpr (std::is_same_v<NT, float>)
        return wave_sum_f32(v);
      else
        return Parts<NT>::make(wave_reduce(Parts<NT>::first(v)), wave_reduce(Parts<NT>::second(v)));
    }

    template <class NT> __device__ __forceinline__ NT block_reduce(NT v)
    {
      // raw storage: NT may have a user-provided constructor (complex)
      __shared__ __attribute__((aligned(16))) unsigned char smem_raw[sizeof(NT) * (kBlock / 64)];
      NT *smem = reinterpret_cast<NT *>(smem_raw);
      const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
      v = wave_reduce(v);
      if (lane == 0) smem[wid] = v;
      __syncthreads();
      NT t = smem[0];
      for (int i = 1; i < kBlock / 64; ++i) t += smem[i];
      return t;
    }

    // Partials of type NT through device-scope atomics, component-wise (dfrg::dev::publish).  A
    // two-component partial is stored as its two components at the start of its NT slot.
    template <class NT> __device__ __forceinline__ void publish_nt(NT *p, const NT &v)
    {
      if constexpr (is_scalar_nt<NT>)
        dev::publish(p, v);
      else {
        using P = typename Parts<NT>::P;
        static_assert(sizeof(NT) >= 2 * sizeof(P), "two-component number type smaller than its components");
        P *raw = reinterpret_cast<P *>(p);
        dev::publish(raw, Parts<NT>::first(v));
        dev::publish(raw + 1, Parts<NT>::second(v));
      }
    }
    template <class NT> __device__ __forceinline__ NT fetch_nt(const NT *p)
    {
      if constexpr (is_scalar_nt<NT>)
        return dev::fetch(p);
      else {
        using P = typename Parts<NT>::P;
        const P *raw = reinterpret_cast<const P *>(p);
        return Parts<NT>::make(dev::fetch(raw), dev::fetch(raw + 1));
      }
    }

    // partial slots of a single call: one 128-byte line each (dfrg::dev::kSlotStride)
    template <class NT> constexpr int slot_stride = sizeof(NT) >= 128 ? 1 : int(128 / sizeof(NT));

    template <class NT> __device__ __forceinline__ NT lane_nt(const NT &v, int i)
    {
      if constexpr (is_scalar_nt<NT>)
        return dev::lane_value(v, i);
      else
        return Parts<NT>::make(dev::lane_value(Parts<NT>::first(v), i), dev::lane_value(Parts<NT>::second(v), i));
    }

    // Publish this block's partial in slot `slot` (at partials[slot * stride]) of a group of n_slots;
    // the last block of the group adds the partials in slot order and writes *result (then resets
    // the counter for the next use of the scratch block).  The first wave of that block fetches the partials one per lane, 64 at
    // a time, and adds them lane by lane in slot order.
    template <class NT>
    __device__ __forceinline__ void finish_slots(NT part, NT *partials, int slot, int n_slots, unsigned *counter,
                                                 NT *result, int stride = 1)
    {
      __shared__ bool last;
      if (threadIdx.x == 0) {
        publish_nt(&partials[(long long)slot * stride], part);
        dev::wait_stores();
        last = dev::take_ticket(counter) == (unsigned)(n_slots - 1);
      }
      __syncthreads();
      if (last && threadIdx.x < 64) {
        NT s = NT(0);
        for (int c = 0; c < n_slots; c += 64) {
          const int j = c + (int)threadIdx.x;
          const NT v = j < n_slots ? fetch_nt(&partials[(long long)j * stride]) : NT(0);
          const int m = n_slots - c < 64 ? n_slots - c : 64;
          if (c == 0) {
            s = lane_nt(v, 0);
            for (int l = 1; l < m; ++l) s += lane_nt(v, l);
          } else
            for (int l = 0; l < m; ++l) s += lane_nt(v, l);
        }
        if (threadIdx.x == 0) {
          *result = s;
          dev::reset_counter(counter);
        }
      }
    }

    // ------------------------------------------------------------------------------------------
    // per-thread node sums.  Block b of nb takes nodes b*kBlock + tid, stepping nb*kBlock; a
    // single call launches nb blocks, a batch launches nb blocks per request with the same
    // partition, so a batched result is bitwise the single call's.
    // ------------------------------------------------------------------------------------------
    // IntegratorGPU<d> (integrator_gpu.hh:34-60)
    // The digits of a flat node index (slowest first) in mixed radix r, walked with a fixed
    // stride S: a thread's nodes i, i + S, i + 2S, ... of the flat loop, with the stride's digits
    // added with carries instead of a division per node.  The slowest digit is clamped to its
    // radix at the start, so a thread whose first node is past the end does not iterate.
    template <int N> struct NodeWalk {
      int dig[N], step[N], rad[N];
      __device__ __forceinline__ NodeWalk(long long i, long long S, const int (&r)[N])
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
        dig[0] = (int)(i < r[0] ? i : r[0]);
        step[0] = (int)S;
      }
      __device__ __forceinline__ bool valid() const { return dig[0] < rad[0]; }
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

    template <int d, class ctype, class NT, class KERNEL, class... T>
    __device__ __forceinline__ NT radial_sum(const ctype *__restrict__ xq, const ctype *__restrict__ wq, int nx,
                                             ctype x_extent, int b, int nb, ctype k, const T &...t)
    {
      NT acc = NT(0);
      constexpr ctype S_d = DiFfRG::S_d_prec<ctype>(d);
      for (int i = b * kBlock + threadIdx.x; i < nx; i += nb * kBlock) {
        const ctype weight = wq[i] * x_extent;
        const ctype q = k * sqrt(xq[i] * x_extent);
        const ctype int_element = S_d * (DiFfRG::powr<d - 2>(q) / 2 * DiFfRG::powr<2>(k)) /
                                  DiFfRG::powr<d>(2 * (ctype)M_PI);
        const NT res = KERNEL::kernel(q, k, t...);
        acc += int_element * res * weight;
      }
      return acc;
    }

    // Integrator4D2AngGPU (integrator_4D_2ang_gpu.hh:20-45): node (q, cos1, cos2), cos2 fastest
    template <class ctype, class NT, class KERNEL, class... T>
    __device__ __forceinline__ NT angle2_sum(const ctype *__restrict__ xq, const ctype *__restrict__ wq,
                                             const ctype *__restrict__ xa, const ctype *__restrict__ wa, int nx,
                                             int na, ctype x_extent, int b, int nb, ctype k, const T &...t)
    {
      NT acc = NT(0);
      for (NodeWalk<3> w((long long)b * kBlock + threadIdx.x, (long long)nb * kBlock, {nx, na, na}); w.valid();
           w.next()) {
        const int ix = w.dig[0], i1 = w.dig[1], i2 = w.dig[2];
        const ctype q = k * sqrt(xq[ix] * x_extent);
        const ctype cos1 = 2 * (xa[i1] - (ctype)0.5);
        const ctype int_element = (DiFfRG::powr<2>(q) * (ctype)0.5 * DiFfRG::powr<2>(k)) // x = p^2 / k^2 integral
                                  * sqrt(1. - DiFfRG::powr<2>(cos1))                      // cos1 Jacobian
                                  * 2 * (ctype)M_PI                                       // phi integral
                                  / DiFfRG::powr<4>(2 * (ctype)M_PI);                     // Fourier factor
        const ctype weight = 2 * wa[i1] * wq[ix] * x_extent * 2 * wa[i2];
        const ctype cos2 = 2 * (xa[i2] - (ctype)0.5);
        acc += KERNEL::kernel(q, cos1, cos2, k, t...) * int_element * weight;
      }
      return acc;
    }

    // Integrator4DGPU (integrator_4D_gpu.hh:20-49): node (q, cos1, phi, cos2), cos2 fastest
    template <class ctype, class NT, class KERNEL, class... T>
    __device__ __forceinline__ NT angle3_sum(const ctype *__restrict__ xq, const ctype *__restrict__ wq,
                                             const ctype *__restrict__ xa, const ctype *__restrict__ wa, int nx,
                                             int na, ctype x_extent, int b, int nb, ctype k, const T &...t)
    {
      NT acc = NT(0);
      for (NodeWalk<4> w((long long)b * kBlock + threadIdx.x, (long long)nb * kBlock, {nx, na, na, na}); w.valid();
           w.next()) {
        const int ix = w.dig[0], i1 = w.dig[1], i3 = w.dig[2], i2 = w.dig[3]; // q, cos1, phi, cos2
        const ctype q = k * sqrt(xq[ix] * x_extent);
        const ctype cos1 = 2 * (xa[i1] - (ctype)0.5);
        const ctype phi = 2 * (ctype)M_PI * xa[i3];
        const ctype int_element = (DiFfRG::powr<2>(q) * (ctype)0.5 * DiFfRG::powr<2>(k)) *
                                  sqrt(1. - DiFfRG::powr<2>(cos1)) / DiFfRG::powr<4>(2 * (ctype)M_PI);
        const ctype weight = 2 * (ctype)M_PI * wa[i3] * 2 * wa[i1] * wq[ix] * x_extent;
        const ctype cos2 = 2 * (xa[i2] - (ctype)0.5);
        acc += KERNEL::kernel(q, cos1, cos2, phi, k, t...) * int_element * weight * 2 * wa[i2];
      }
      return acc;
    }

    // IntegratorFiniteTq0GPU<d> (integrator_finiteTq0_gpu.hh:20-44): node (q, Matsubara n)
    template <int d, class ctype, class NT, class KERNEL, class... T>
    __device__ __forceinline__ NT finiteTq0_sum(const ctype *__restrict__ xq, const ctype *__restrict__ wq,
                                                const ctype *__restrict__ mx, const ctype *__restrict__ mw, int nx,
                                                int nm, ctype x_extent, ctype m_T, int b, int nb, ctype k,
                                                const T &...t)
    {
      NT acc = NT(0);
      constexpr ctype S_dm1 = DiFfRG::S_d_prec<ctype>(d - 1);
      for (NodeWalk<2> w((long long)b * kBlock + threadIdx.x, (long long)nb * kBlock, {nx, nm}); w.valid(); w.next()) {
        const int ix = w.dig[0], iy = w.dig[1];
        const ctype q = k * sqrt(xq[ix] * x_extent);
        const ctype x_weight = wq[ix] * x_extent;
        const ctype int_element = S_dm1 * (DiFfRG::powr<d - 3>(q) / (ctype)2 * DiFfRG::powr<2>(k)) /
                                  DiFfRG::powr<d - 1>(2 * (ctype)M_PI);
        const ctype q0 = mx[iy];
        const ctype weight = x_weight * mw[iy];
        NT res = int_element * weight * (KERNEL::kernel(q, q0, k, t...) + KERNEL::kernel(q, -q0, k, t...));
        if (m_T > 0 && iy == 0) res += int_element * x_weight * m_T * KERNEL::kernel(q, (ctype)0, k, t...);
        acc += res;
      }
      return acc;
    }

    // IntegratorAngleGPU<d> (integrator_angle_gpu.hh:21-40): node (q, cos), cos fastest
    template <int d, class ctype, class NT, class KERNEL, class... T>
    __device__ __forceinline__ NT angle1_sum(const ctype *__restrict__ xq, const ctype *__restrict__ wq,
                                             const ctype *__restrict__ xa, const ctype *__restrict__ wa, int nx,
                                             int na, ctype x_extent, int b, int nb, ctype k, const T &...t)
    {
      NT acc = NT(0);
      constexpr ctype S_d = DiFfRG::S_d_prec<ctype>(d);
      for (NodeWalk<2> w((long long)b * kBlock + threadIdx.x, (long long)nb * kBlock, {nx, na}); w.valid(); w.next()) {
        const int ix = w.dig[0], iy = w.dig[1];
        const ctype q = k * sqrt(xq[ix] * x_extent);
        const ctype cos = 2 * (xa[iy] - (ctype)0.5);
        const ctype int_element = S_d                                                // solid angle
                                  * (DiFfRG::powr<d - 2>(q) / (ctype)2 * DiFfRG::powr<2>(k)) // x = q^2/k^2
                                  * (1 / (ctype)2)                                   // cos integral divided out
                                  / DiFfRG::powr<d>(2 * (ctype)M_PI);                // Fourier factor
        const ctype weight = 2 * wa[iy] * wq[ix] * x_extent;
        acc += int_element * weight * KERNEL::kernel(q, cos, k, t...);
      }
      return acc;
    }

    // Integrator4DFiniteTq0GPU (integrator_4D_finiteTq0_gpu.hh:20-55): node (q, cos, phi, Matsubara n)
    template <class ctype, class NT, class KERNEL, class... T>
    __device__ __forceinline__ NT angle2_finiteT_sum(const ctype *__restrict__ xq, const ctype *__restrict__ wq,
                                                     const ctype *__restrict__ xa, const ctype *__restrict__ wa,
                                                     const ctype *__restrict__ mx, const ctype *__restrict__ mw,
                                                     int nx, int na, int nm, ctype x_extent, ctype m_T, int b, int nb,
                                                     ctype k, const T &...t)
    {
      NT acc = NT(0);
      for (NodeWalk<4> w((long long)b * kBlock + threadIdx.x, (long long)nb * kBlock, {nx, na, na, nm}); w.valid();
           w.next()) {
        const int ix = w.dig[0], iy = w.dig[1], iz = w.dig[2], in = w.dig[3]; // q, cos, phi, n
        const ctype q = k * sqrt(xq[ix] * x_extent);
        const ctype int_element = (DiFfRG::powr<1>(q) / (ctype)2 * DiFfRG::powr<2>(k)) // x = q^2/k^2, 3 dims
                                  / DiFfRG::powr<3>(2 * (ctype)M_PI);                   // Fourier factor
        const ctype cos = 2 * (xa[iy] - (ctype)0.5);
        const ctype phi = 2 * (ctype)M_PI * xa[iz];
        const ctype x_weight = 2 * (ctype)M_PI * wa[iz] * 2 * wa[iy] * wq[ix] * x_extent;
        const ctype q0 = mx[in];
        const ctype weight = x_weight * mw[in];
        NT res = int_element * weight *
                 (KERNEL::kernel(q, cos, phi, q0, k, t...) + KERNEL::kernel(q, cos, phi, -q0, k, t...));
        if (m_T > 0 && in == 0) res += int_element * x_weight * m_T * KERNEL::kernel(q, cos, phi, (ctype)0, k, t...);
        acc += res;
      }
      return acc;
    }

    // IntegratorAngleFiniteTq0GPU<d> (integrator_angle_finiteTq0_gpu.hh:20-50): node (q, cos, Matsubara n)
    template <int d, class ctype, class NT, class KERNEL, class... T>
    __device__ __forceinline__ NT angle1_finiteT_sum(const ctype *__restrict__ xq, const ctype *__restrict__ wq,
                                                     const ctype *__restrict__ xa, const ctype *__restrict__ wa,
                                                     const ctype *__restrict__ mx, const ctype *__restrict__ mw,
                                                     int nx, int na, int nm, ctype x_extent, ctype m_T, int b, int nb,
                                                     ctype k, const T &...t)
    {
      NT acc = NT(0);
      constexpr ctype S_dm1 = DiFfRG::S_d_prec<ctype>(d - 1);
      for (NodeWalk<3> w((long long)b * kBlock + threadIdx.x, (long long)nb * kBlock, {nx, na, nm}); w.valid();
           w.next()) {
        const int ix = w.dig[0], iy = w.dig[1], in = w.dig[2];
        const ctype q = k * sqrt(xq[ix] * x_extent);
        const ctype int_element = S_dm1                                                // solid angle
                                  * (DiFfRG::powr<d - 3>(q) / (ctype)2 * DiFfRG::powr<2>(k)) // x = q^2/k^2
                                  * (1 / (ctype)2)                                     // cos integral divided out
                                  / DiFfRG::powr<d - 1>(2 * (ctype)M_PI);              // Fourier factor
        const ctype cos = 2 * (xa[iy] - (ctype)0.5);
        const ctype x_weight = 2 * wa[iy] * wq[ix] * x_extent;
        const ctype q0 = mx[in];
        const ctype weight = x_weight * mw[in];
        NT res = int_element * weight * (KERNEL::kernel(q, cos, q0, k, t...) + KERNEL::kernel(q, cos, -q0, k, t...));
        if (m_T > 0 && in == 0) res += int_element * x_weight * m_T * KERNEL::kernel(q, cos, (ctype)0, k, t...);
        acc += res;
      }
      return acc;
    }

    // ------------------------------------------------------------------------------------------
    // kernels: one call (arguments by value), or a batch of requests whose (k, t...) tuples sit in
    // device memory, nb blocks per request
    // ------------------------------------------------------------------------------------------
    template <int d, class ctype, class NT, class KERNEL, class... T>
    __global__ __launch_bounds__(kBlock) void radial_kernel(const ctype *__restrict__ xq, const ctype *__restrict__ wq,
                                                            int nx, ctype x_extent, ctype k, NT *partials,
                                                            unsigned *counter, NT *result, T... t)
    {
      const NT acc = radial_sum<d, ctype, NT, KERNEL>(xq, wq, nx, x_extent, blockIdx.x, gridDim.x, k, t...);
      finish_slots(block_reduce(acc), partials, blockIdx.x, gridDim.x, counter, result, slot_stride<NT>);
    }
    template <int d, class ctype, class NT, class KERNEL, class Args>
    __global__ __launch_bounds__(kBlock) void radial_batch_kernel(const ctype *__restrict__ xq,
                                                                  const ctype *__restrict__ wq, int nx, ctype x_extent,
                                                                  const Args *args, int nb, NT *partials,
                                                                  unsigned *counters, NT *results)
    {
      const int r = blockIdx.x / nb, b = blockIdx.x - r * nb;
      const NT acc = std::apply(
          [&](const ctype &k, const auto &...t) {
            return radial_sum<d, ctype, NT, KERNEL>(xq, wq, nx, x_extent, b, nb, k, t...);
          },
          args[r]);
      finish_slots(block_reduce(acc), partials + (long long)r * nb, b, nb,
                   counters + (long long)r * dev::kCounterStride, results + r);
    }

    template <class ctype, class NT, class KERNEL, class... T>
    __global__ __launch_bounds__(kBlock) void angle2_kernel(const ctype *__restrict__ xq, const ctype *__restrict__ wq,
                                                            const ctype *__restrict__ xa, const ctype *__restrict__ wa,
                                                            int nx, int na, ctype x_extent, ctype k, NT *partials,
                                                            unsigned *counter, NT *result, T... t)
    {
      const NT acc = angle2_sum<ctype, NT, KERNEL>(xq, wq, xa, wa, nx, na, x_extent, blockIdx.x, gridDim.x, k, t...);
      finish_slots(block_reduce(acc), partials, blockIdx.x, gridDim.x, counter, result, slot_stride<NT>);
    }
    template <class ctype, class NT, class KERNEL, class Args>
    __global__ __launch_bounds__(kBlock) void angle2_batch_kernel(const ctype *__restrict__ xq,
                                                                  const ctype *__restrict__ wq,
                                                                  const ctype *__restrict__ xa,
                                                                  const ctype *__restrict__ wa, int nx, int na,
                                                                  ctype x_extent, const Args *args, int nb,
                                                                  NT *partials, unsigned *counters, NT *results)
    {
      const int r = blockIdx.x / nb, b = blockIdx.x - r * nb;
      const NT acc = std::apply(
          [&](const ctype &k, const auto &...t) {
            return angle2_sum<ctype, NT, KERNEL>(xq, wq, xa, wa, nx, na, x_extent, b, nb, k, t...);
          },
          args[r]);
      finish_slots(block_reduce(acc), partials + (long long)r * nb, b, nb,
                   counters + (long long)r * dev::kCounterStride, results + r);
    }

    template <class ctype, class NT, class KERNEL, class... T>
    __global__ __launch_bounds__(kBlock) void angle3_kernel(const ctype *__restrict__ xq, const ctype *__restrict__ wq,
                                                            const ctype *__restrict__ xa, const ctype *__restrict__ wa,
                                                            int nx, int na, ctype x_extent, ctype k, NT *partials,
                                                            unsigned *counter, NT *result, T... t)
    {
      const NT acc = angle3_sum<ctype, NT, KERNEL>(xq, wq, xa, wa, nx, na, x_extent, blockIdx.x, gridDim.x, k, t...);
      finish_slots(block_reduce(acc), partials, blockIdx.x, gridDim.x, counter, result, slot_stride<NT>);
    }
    template <class ctype, class NT, class KERNEL, class Args>
    __global__ __launch_bounds__(kBlock) void angle3_batch_kernel(const ctype *__restrict__ xq,
                                                                  const ctype *__restrict__ wq,
                                                                  const ctype *__restrict__ xa,
                                                                  const ctype *__restrict__ wa, int nx, int na,
                                                                  ctype x_extent, const Args *args, int nb,
                                                                  NT *partials, unsigned *counters, NT *results)
    {
      const int r = blockIdx.x / nb, b = blockIdx.x - r * nb;
      const NT acc = std::apply(
          [&](const ctype &k, const auto &...t) {
            return angle3_sum<ctype, NT, KERNEL>(xq, wq, xa, wa, nx, na, x_extent, b, nb, k, t...);
          },
          args[r]);
      finish_slots(block_reduce(acc), partials + (long long)r * nb, b, nb,
                   counters + (long long)r * dev::kCounterStride, results + r);
    }

    template <int d, class ctype, class NT, class KERNEL, class... T>
    __global__ __launch_bounds__(kBlock) void finiteTq0_kernel(const ctype *__restrict__ xq,
                                                               const ctype *__restrict__ wq,
                                                               const ctype *__restrict__ mx,
                                                               const ctype *__restrict__ mw, int nx, int nm,
                                                               ctype x_extent, ctype m_T, ctype k, NT *partials,
                                                               unsigned *counter, NT *result, T... t)
    {
      const NT acc =
          finiteTq0_sum<d, ctype, NT, KERNEL>(xq, wq, mx, mw, nx, nm, x_extent, m_T, blockIdx.x, gridDim.x, k, t...);
      finish_slots(block_reduce(acc), partials, blockIdx.x, gridDim.x, counter, result, slot_stride<NT>);
    }
    template <int d, class ctype, class NT, class KERNEL, class Args>
    __global__ __launch_bounds__(kBlock) void finiteTq0_batch_kernel(const ctype *__restrict__ xq,
                                                                     const ctype *__restrict__ wq,
                                                                     const ctype *__restrict__ mx,
                                                                     const ctype *__restrict__ mw, int nx, int nm,
                                                                     ctype x_extent, ctype m_T, const Args *args,
                                                                     int nb, NT *partials, unsigned *counters,
                                                                     NT *results)
    {
      const int r = blockIdx.x / nb, b = blockIdx.x - r * nb;
      const NT acc = std::apply(
          [&](const ctype &k, const auto &...t) {
            return finiteTq0_sum<d, ctype, NT, KERNEL>(xq, wq, mx, mw, nx, nm, x_extent, m_T, b, nb, k, t...);
          },
          args[r]);
      finish_slots(block_reduce(acc), partials + (long long)r * nb, b, nb,
                   counters + (long long)r * dev::kCounterStride, results + r);
    }

    template <int d, class ctype, class NT, class KERNEL, class... T>
    __global__ __launch_bounds__(kBlock) void angle1_kernel(const ctype *__restrict__ xq, const ctype *__restrict__ wq,
                                                            const ctype *__restrict__ xa, const ctype *__restrict__ wa,
                                                            int nx, int na, ctype x_extent, ctype k, NT *partials,
                                                            unsigned *counter, NT *result, T... t)
    {
      const NT acc =
          angle1_sum<d, ctype, NT, KERNEL>(xq, wq, xa, wa, nx, na, x_extent, blockIdx.x, gridDim.x, k, t...);
      finish_slots(block_reduce(acc), partials, blockIdx.x, gridDim.x, counter, result, slot_stride<NT>);
    }
    template <int d, class ctype, class NT, class KERNEL, class Args>
    __global__ __launch_bounds__(kBlock) void angle1_batch_kernel(const ctype *__restrict__ xq,
                                                                  const ctype *__restrict__ wq,
                                                                  const ctype *__restrict__ xa,
                                                                  const ctype *__restrict__ wa, int nx, int na,
                                                                  ctype x_extent, const Args *args, int nb,
                                                                  NT *partials, unsigned *counters, NT *results)
    {
      const int r = blockIdx.x / nb, b = blockIdx.x - r * nb;
      const NT acc = std::apply(
          [&](const ctype &k, const auto &...t) {
            return angle1_sum<d, ctype, NT, KERNEL>(xq, wq, xa, wa, nx, na, x_extent, b, nb, k, t...);
          },
          args[r]);
      finish_slots(block_reduce(acc), partials + (long long)r * nb, b, nb,
                   counters + (long long)r * dev::kCounterStride, results + r);
    }

    template <class ctype, class NT, class KERNEL, class... T>
    __global__ __launch_bounds__(kBlock) void angle2_finiteT_kernel(
        const ctype *__restrict__ xq, const ctype *__restrict__ wq, const ctype *__restrict__ xa,
        const ctype *__restrict__ wa, const ctype *__restrict__ mx, const ctype *__restrict__ mw, int nx, int na,
        int nm, ctype x_extent, ctype m_T, ctype k, NT *partials, unsigned *counter, NT *result, T... t)
    {
      const NT acc = angle2_finiteT_sum<ctype, NT, KERNEL>(xq, wq, xa, wa, mx, mw, nx, na, nm, x_extent, m_T,
                                                           blockIdx.x, gridDim.x, k, t...);
      finish_slots(block_reduce(acc), partials, blockIdx.x, gridDim.x, counter, result, slot_stride<NT>);
    }
    template <class ctype, class NT, class KERNEL, class Args>
    __global__ __launch_bounds__(kBlock) void angle2_finiteT_batch_kernel(
        const ctype *__restrict__ xq, const ctype *__restrict__ wq, const ctype *__restrict__ xa,
        const ctype *__restrict__ wa, const ctype *__restrict__ mx, const ctype *__restrict__ mw, int nx, int na,
        int nm, ctype x_extent, ctype m_T, const Args *args, int nb, NT *partials, unsigned *counters, NT *results)
    {
      const int r = blockIdx.x / nb, b = blockIdx.x - r * nb;
      const NT acc = std::apply(
          [&](const ctype &k, const auto &...t) {
            return angle2_finiteT_sum<ctype, NT, KERNEL>(xq, wq, xa, wa, mx, mw, nx, na, nm, x_extent, m_T, b, nb, k,
                                                         t...);
          },
          args[r]);
      finish_slots(block_reduce(acc), partials + (long long)r * nb, b, nb,
                   counters + (long long)r * dev::kCounterStride, results + r);
    }

    template <int d, class ctype, class NT, class KERNEL, class... T>
    __global__ __launch_bounds__(kBlock) void angle1_finiteT_kernel(
        const ctype *__restrict__ xq, const ctype *__restrict__ wq, const ctype *__restrict__ xa,
        const ctype *__restrict__ wa, const ctype *__restrict__ mx, const ctype *__restrict__ mw, int nx, int na,
        int nm, ctype x_extent, ctype m_T, ctype k, NT *partials, unsigned *counter, NT *result, T... t)
    {
      const NT acc = angle1_finiteT_sum<d, ctype, NT, KERNEL>(xq, wq, xa, wa, mx, mw, nx, na, nm, x_extent, m_T,
                                                              blockIdx.x, gridDim.x, k, t...);
      finish_slots(block_reduce(acc), partials, blockIdx.x, gridDim.x, counter, result, slot_stride<NT>);
    }
    template <int d, class ctype, class NT, class KERNEL, class Args>
    __global__ __launch_bounds__(kBlock) void angle1_finiteT_batch_kernel(
        const ctype *__restrict__ xq, const ctype *__restrict__ wq, const ctype *__restrict__ xa,
        const ctype *__restrict__ wa, const ctype *__restrict__ mx, const ctype *__restrict__ mw, int nx, int na,
        int nm, ctype x_extent, ctype m_T, const Args *args, int nb, NT *partials, unsigned *counters, NT *results)
    {
      const int r = blockIdx.x / nb, b = blockIdx.x - r * nb;
      const NT acc = std::apply(
          [&](const ctype &k, const auto &...t) {
            return angle1_finiteT_sum<d, ctype, NT, KERNEL>(xq, wq, xa, wa, mx, mw, nx, na, nm, x_extent, m_T, b, nb,
                                                            k, t...);
          },
          args[r]);
      finish_slots(block_reduce(acc), partials + (long long)r * nb, b, nb,
                   counters + (long long)r * dev::kCounterStride, results + r);
    }

    // ------------------------------------------------------------------------------------------
    // host side: per-call scratch, launch and result
    // ------------------------------------------------------------------------------------------
    inline int grid_for(long long nodes)
    {
      const long long per_block = (long long)kBlock * kNodesPerLane;
      long long g = (nodes + per_block - 1) / per_block;
      if (g < 1) g = 1;
      if (g > kMaxBlocks) g = kMaxBlocks;
      return (int)g;
    }

    // One in-flight evaluation: scratch (partials | counter | result) in HBM and a pinned result
    // slot, enqueued on `stream`.
    template <class NT> struct Pending {
      hipStream_t stream = nullptr;
      void *d_scratch = nullptr;
      void *h_result = nullptr;
      NT constant = NT(0);

      static constexpr size_t counter_offset = ((kMaxBlocks * slot_stride<NT> * sizeof(NT) + 255) / 256) * 256;
      static constexpr size_t result_offset = counter_offset + 256;
      static constexpr size_t bytes = result_offset + 256;

      NT *partials() const { return (NT *)d_scratch; }
      unsigned *counter() const { return (unsigned *)((char *)d_scratch + counter_offset); }
      NT *result() const { return (NT *)((char *)d_scratch + result_offset); }

      static Pending make()
      {
        Pending p;
        void *s = nullptr;
        DiFfRG::dfrg_check(dfrg_thread_stream(&s), "thread stream");
        p.stream = (hipStream_t)s;
        DiFfRG::dfrg_check(dfrg_pool_alloc(&p.d_scratch, bytes), "scratch");
        DiFfRG::dfrg_check(dfrg_pool_host_alloc(&p.h_result, sizeof(NT)), "result slot");
        return p;
      }
      // after the kernel launch: copy the result back on the same stream
      void enqueue_result()
      {
        DiFfRG::check_hip("integrator kernel launch");
        if (hipMemcpyAsync(h_result, result(), sizeof(NT), hipMemcpyDeviceToHost, stream) != hipSuccess)
          throw std::runtime_error("integrator: result copy failed");
      }
      NT wait()
      {
        if (hipStreamSynchronize(stream) != hipSuccess) throw std::runtime_error("integrator: stream error");
        NT v;
        std::memcpy((void *)&v, h_result, sizeof(NT));
        dfrg_pool_free(d_scratch);
        dfrg_pool_host_free(h_result);
        return constant + v;
      }
    };

    // ------------------------------------------------------------------------------------------
    // batched requests.  request() calls of one integrator are collected: their (k, t...) tuples
    // wait on the host, and reading any of their futures launches one kernel for all of them (nb
    // blocks per request, the single-call partition) and copies all results back at once.  This
    // turns a request_data loop over a grid of external points into one launch and one sync.
    // ------------------------------------------------------------------------------------------
    constexpr size_t kMaxBatch = size_t(1) << 16;

    struct BatchBase {
      virtual ~BatchBase() = default;
    };

    template <class NT, class Args> struct RequestBatch : BatchBase {
      using Launch = std::function<void(const Args *d_args, int n, int nb, NT *partials, unsigned *counters,
                                        NT *results, hipStream_t stream)>;
      std::mutex m;
      std::vector<Args> args;
      std::vector<NT> constants, results;
      int nb = 1;
      bool done = false;
      Launch launch;

      NT get(size_t i)
      {
        std::lock_guard<std::mutex> lock(m);
        if (!done) {
          run();
          done = true;
        }
        return constants[i] + results[i];
      }

    private:
      static size_t up(size_t b) { return (b + 255) / 256 * 256; }
      void run()
      {
        const int n = (int)args.size();
        void *sv = nullptr;
        DiFfRG::dfrg_check(dfrg_thread_stream(&sv), "thread stream");
        const hipStream_t s = (hipStream_t)sv;
        const size_t o_part = up(sizeof(Args) * n);
        const size_t o_cnt = o_part + up(sizeof(NT) * (size_t)n * nb);
        const size_t o_res = o_cnt + up(sizeof(unsigned) * dev::kCounterStride * n);
        const size_t bytes = o_res + up(sizeof(NT) * n);
        const size_t h_res_off = up(sizeof(Args) * n);
        void *d = nullptr, *h = nullptr;
        DiFfRG::dfrg_check(dfrg_pool_alloc(&d, bytes), "batch scratch");
        DiFfRG::dfrg_check(dfrg_pool_host_alloc(&h, h_res_off + sizeof(NT) * n), "batch staging");
        char *dc = (char *)d, *hc = (char *)h;
        std::memcpy(hc, (const void *)args.data(), sizeof(Args) * n);
        bool ok = hipMemcpyAsync(dc, hc, sizeof(Args) * n, hipMemcpyHostToDevice, s) == hipSuccess &&
                  hipMemsetAsync(dc + o_cnt, 0, sizeof(unsigned) * dev::kCounterStride * n, s) == hipSuccess;
        if (ok) {
          launch((const Args *)dc, n, nb, (NT *)(dc + o_part), (unsigned *)(dc + o_cnt), (NT *)(dc + o_res), s);
          ok = hipGetLastError() == hipSuccess &&
               hipMemcpyAsync(hc + h_res_off, dc + o_res, sizeof(NT) * n, hipMemcpyDeviceToHost, s) == hipSuccess &&
               hipStreamSynchronize(s) == hipSuccess;
        }
        if (ok) {
          results.resize(n);
          std::memcpy((void *)results.data(), hc + h_res_off, sizeof(NT) * n);
        }
        dfrg_pool_free(d);
        dfrg_pool_host_free(h);
        args.clear();
        args.shrink_to_fit();
        if (!ok) throw std::runtime_error("integrator: batched request launch failed");
      }
    };

    // Collects the requests of one integrator into batches.  A new batch opens when the argument
    // types or the integrator state that the launch depends on (`state`, e.g. the Matsubara rule)
    // change, when the open batch has been read, or at kMaxBatch requests.
    template <class NT> class RequestBatcher
    {
    public:
      template <class Args, class MakeLaunch>
      std::future<NT> enqueue(unsigned long long state, int nb, NT constant, Args &&args, MakeLaunch &&make_launch)
      {
        using B = RequestBatch<NT, std::decay_t<Args>>;
        std::shared_ptr<B> b;
        size_t idx = 0;
        {
          std::lock_guard<std::mutex> lock(m);
          b = std::dynamic_pointer_cast<B>(open);
          for (int attempt = 0;; ++attempt) {
            if (!b || state != open_state || attempt > 0) {
              b = std::make_shared<B>();
              b->nb = nb;
              b->launch = make_launch();
              open = b;
              open_state = state;
            }
            std::lock_guard<std::mutex> lb(b->m);
            if (b->done || b->args.size() >= kMaxBatch) continue; // read already, or full: open another
            idx = b->args.size();
            b->args.push_back(std::forward<Args>(args));
            b->constants.push_back(constant);
            break;
          }
        }
        return std::async(std::launch::deferred, [b, idx]() { return b->get(idx); });
      }

    private:
      std::mutex m;
      std::shared_ptr<BatchBase> open;
      unsigned long long open_state = 0;
    };
  } // namespace tmpl
} // namespace dfrg
