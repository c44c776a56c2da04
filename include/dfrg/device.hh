// dfrg/device.hh -- device building blocks shared by the C-ABI kernels (diffrg_amd/csrc)
// and the drop-in integrator templates (include/DiFfRG/...).  gfx950 (CDNA4), wave64.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace dfrg
{
  namespace dev
  {
    constexpr int wave = 64;
    constexpr double pi = 3.14159265358979323846;

    // ------------------------------------------------------------------------------------------
    // wave64 sum: DPP within each 16-lane row (quad_perm xor1, xor2, row_half_mirror,
    // row_mirror), then two cross-row swaps.  Partners always add the same two operands,
    // so every lane ends with the bitwise-identical total (a fixed, data-independent tree).
    // ------------------------------------------------------------------------------------------
    template <int ctrl> __device__ __forceinline__ double dpp_f64(double v)
    {
      const uint64_t u = __double_as_longlong(v);
      const int lo = __builtin_amdgcn_mov_dpp((int)(uint32_t)u, ctrl, 0xF, 0xF, false);
      const int hi = __builtin_amdgcn_mov_dpp((int)(uint32_t)(u >> 32), ctrl, 0xF, 0xF, false);
      return __longlong_as_double(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
    }
    template <int ctrl> __device__ __forceinline__ float dpp_f32(float v)
    {
      return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), ctrl, 0xF, 0xF, false));
    }

    __device__ __forceinline__ double wave_sum(double v)
    {
      v += dpp_f64<0xB1>(v);  // quad_perm(1,0,3,2): lane ^ 1
      v += dpp_f64<0x4E>(v);  // quad_perm(2,3,0,1): lane ^ 2
      v += dpp_f64<0x141>(v); // row_half_mirror: pairs the two quads of each 8-lane half
      v += dpp_f64<0x140>(v); // row_mirror: pairs the two halves of each 16-lane row
      v += __shfl_xor(v, 16, wave);
      v += __shfl_xor(v, 32, wave);
      return v;
    }

    // Sums of eight wave-wide vectors in one pass, by recursive halving: pairs of vectors swap
    // register halves across lanes (v_permlane32_swap, then v_permlane16_swap) so that one add
    // combines two vectors, an 8-lane rotation merges the last pair, and DPP finishes inside
    // 8-lane groups.  34 VALU instructions for the eight sums instead of eight wave_sum calls.
    // Lane l returns the total of vector 4 b3 + 2 b4 + b5 (b = the bits of l); sum8_lane(v) is
    // the first of the eight lanes that hold vector v's total.  Every vector is added with the same
    // tree (lane l with l + 32, then + 16, then l ^ 8, then within eight lanes, each add
    // commutative), so its total does not depend on its position or on the other seven vectors.
    constexpr int sum8_lane(int v) { return 8 * (v >> 2) + 16 * ((v >> 1) & 1) + 32 * (v & 1); }

    // Lanes 0-31 return a_l + a_{l+32}, lanes 32-63 b_{l-32} + b_l.
    __device__ __forceinline__ double pair_sum32(double a, double b)
    {
      const unsigned long long ua = __double_as_longlong(a), ub = __double_as_longlong(b);
      const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)ua, (unsigned)ub, false, false);
      const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(ua >> 32), (unsigned)(ub >> 32), false, false);
      const double a2 = __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0]));
      const double b2 = __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
      return a2 + b2;
    }
    // Within each 32-lane half: lanes 0-15 return a_l + a_{l+16}, lanes 16-31 b_{l-16} + b_l.
    __device__ __forceinline__ double pair_sum16(double a, double b)
    {
      const unsigned long long ua = __double_as_longlong(a), ub = __double_as_longlong(b);
      const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)ua, (unsigned)ub, false, false);
      const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(ua >> 32), (unsigned)(ub >> 32), false, false);
      const double a2 = __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0]));
      const double b2 = __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
      return a2 + b2;
    }
    __device__ __forceinline__ double wave_sum8(const double (&v)[8])
    {
      const double w0 = pair_sum32(v[0], v[1]), w1 = pair_sum32(v[2], v[3]);
      const double w2 = pair_sum32(v[4], v[5]), w3 = pair_sum32(v[6], v[7]);
      const double x0 = pair_sum16(w0, w1), x1 = pair_sum16(w2, w3);
      const bool up = (threadIdx.x & 8) != 0;
      double y = (up ? x1 : x0) + dpp_f64<0x128>(up ? x0 : x1); // row_ror:8 pairs lane l with l ^ 8
      y += dpp_f64<0xB1>(y);                                     // lane ^ 1
      y += dpp_f64<0x4E>(y);                                     // lane ^ 2
      y += dpp_f64<0x141>(y);                                    // the two quads of each 8 lanes
      return y;
    }

    // Sums of four wave-wide vectors: wave_sum8's first two levels leave each vector in one 16-lane
    // row, then the same DPP steps finish inside the rows.  Every vector goes through wave_sum8's
    // add tree, so its total is bitwise the one wave_sum8 gives it; vector v's total is in the 16
    // lanes from sum8_lane(v) (v < 4).  21 VALU instructions for the four sums.
    __device__ __forceinline__ double wave_sum4(const double (&v)[4])
    {
      const double x = pair_sum16(pair_sum32(v[0], v[1]), pair_sum32(v[2], v[3]));
      double y = x + dpp_f64<0x128>(x); // row_ror:8 pairs lane l with l ^ 8
      y += dpp_f64<0xB1>(y);            // lane ^ 1
      y += dpp_f64<0x4E>(y);            // lane ^ 2
      y += dpp_f64<0x141>(y);           // the two quads of each 8 lanes
      return y;
    }

    // Block sum in a fixed order: wave sums to LDS, then wave 0 adds them in index order.
    // `smem` needs blockDim.x / 64 doubles.  Returns the total to every thread of the block.
    __device__ __forceinline__ double block_sum(double v, double *smem)
    {
      const int lane = threadIdx.x & (wave - 1);
      const int wid = threadIdx.x / wave;
      const int nw = (blockDim.x + wave - 1) / wave;
      v = wave_sum(v);
      __syncthreads();
      if (lane == 0) smem[wid] = v;
      __syncthreads();
      double t = 0.0;
      for (int i = 0; i < nw; ++i) t += smem[i];
      __syncthreads();
      return t;
    }

    // ------------------------------------------------------------------------------------------
    // Cross-block publication without cache-wide fences.  Each XCD has its own L2, so a release
    // fence at device scope writes back the whole L2 and an acquire fence invalidates it, which
    // costs microseconds per block.  Partials are instead written and read with device-scope
    // relaxed atomics, which are coherent across the XCDs by themselves; waiting for the store
    // counter before the ticket increment makes the published value complete before any block
    // can observe the ticket.  No ordinary (non-atomic) store needs to be published.
    // ------------------------------------------------------------------------------------------
    __device__ __forceinline__ void publish(double *p, double v)
    {
      __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ __forceinline__ void publish(float *p, float v)
    {
      __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ __forceinline__ double fetch(const double *p)
    {
      return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ __forceinline__ float fetch(const float *p)
    {
      return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // Wait until this wave's stores (and every other vector memory operation) have completed;
    // the memory clobber also keeps the compiler from moving memory accesses across it.
    __device__ __forceinline__ void wait_stores() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }
    __device__ __forceinline__ unsigned take_ticket(unsigned *counter)
    {
      const unsigned t = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      asm volatile("" ::: "memory"); // reads of partials stay after the ticket
      return t;
    }
    __device__ __forceinline__ void reset_counter(unsigned *counter)
    {
      __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }

    // Cross-block scratch takes one 128-byte line per partial slot and per counter.  Device-scope
    // atomics and write-through stores that hit one line are serialised where the line is
    // coherent, so slots and counters that many blocks touch at the same time get lines of their
    // own.
    constexpr int kSlotStride = 16;    // doubles per partial slot
    constexpr int kCounterStride = 32; // unsigned per counter

    // Lane i's value of v, for a wave-uniform i (two 32-bit lane reads into scalar registers).
    __device__ __forceinline__ double lane_value(double v, int i)
    {
      const long long b = __double_as_longlong(v);
      const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, i);
      const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), i);
      return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
    }
    __device__ __forceinline__ float lane_value(float v, int i)
    {
      return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), i));
    }

    // ------------------------------------------------------------------------------------------
    // Deterministic cross-block reduction.  Each block publishes its partial and takes a ticket;
    // the block that takes the last ticket of a group sums all partials of that group with a tree
    // fixed by the group size and resets the group's counter so the next launch (or graph replay) starts from zero.
    // Slot s lives at partials[s * kSlotStride].  The finishing block's first wave fetches the
    // partials 64 at a time, one per lane, so the fetches of a group cost one memory round trip
    // per 64 slots instead of one per slot, and adds each round with the wave's fixed DPP tree
    // (rounds in slot order) rather than a serial chain over the lanes.  Returns true in the
    // finishing block, with the group total in *total (valid in thread 0).
    // ------------------------------------------------------------------------------------------
    __device__ __forceinline__ bool finish_group(double partial, double *partials, int slot, int first_slot,
                                                 int n_slots, unsigned *counter, double *total)
    {
      __shared__ bool last;
      if (threadIdx.x == 0) {
        publish(&partials[(long long)slot * kSlotStride], partial);
        wait_stores();
        last = take_ticket(counter) == (unsigned)(n_slots - 1);
      }
      __syncthreads();
      if (!last) return false;
      if (threadIdx.x < 64) {
        double t = 0.0;
        for (int c = 0; c < n_slots; c += 64) {
          const int j = c + (int)threadIdx.x;
          const double v = j < n_slots ? fetch(&partials[(long long)(first_slot + j) * kSlotStride]) : 0.0;
          t += wave_sum(v);
        }
        if (threadIdx.x == 0) {
          *total = t;
          reset_counter(counter);
        }
      }
      return true;
    }

    // ------------------------------------------------------------------------------------------
    // Exponential regulator R = k^2 exp(-(q^2/k^2)^b), b = 1 specialisation used by the
    // built-in integrands (regulators.hh:144-191 form).
    // ------------------------------------------------------------------------------------------
    __device__ __forceinline__ double exp_RB_b1(double k2, double q2) { return k2 * exp(-(q2 / k2)); }
    __device__ __forceinline__ double exp_RBdot_b1(double k2, double q2)
    {
      const double xb = q2 / k2;
      return 2.0 * exp(-xb) * k2 * (1.0 + xb);
    }

    // Log-coordinate lerp: idx = log1p((y - c) * inv_b) * scale, clamped to [0, n-1]
    // (coordinates.hh:182-186 backward, linear_interpolation_1D.hh:75-92 lerp).
    __device__ __forceinline__ double lerp_log_f64(const double *tab, int n, double c, double inv_b, double scale,
                                                   double y)
    {
      double idx = log1p((y - c) * inv_b) * scale;
      idx = fmax(0.0, fmin(idx, (double)(n - 1)));
      const double fl = floor(idx);
      const int i0 = (int)fl;
      const int i1 = (int)ceil(idx);
      return tab[i0] * (1.0 - idx + fl) + tab[i1] * (idx - fl);
    }

    // FP32 variant: hardware log2 / floor, table in FP32.
    __device__ __forceinline__ float lerp_log_f32(const float *tab, int n, float c, float inv_b, float scale_ln2,
                                                  float y)
    {
      // log1p(t) * scale = log2(1 + t) * (ln2 * scale)
      float idx = __builtin_amdgcn_logf(1.0f + (y - c) * inv_b) * scale_ln2;
      idx = fmaxf(0.0f, fminf(idx, (float)(n - 1))); // NaN-safe clamp
      const float fl = __builtin_floorf(idx);
      const float fr = idx - fl;
      const int i0 = (int)fl;
      const int i1 = i0 + (fr > 0.0f ? 1 : 0);
      return tab[i0] + fr * (tab[i1] - tab[i0]);
    }

    __device__ __forceinline__ float fast_exp2f(float v) { return __builtin_amdgcn_exp2f(v); }
    // Reciprocal with one Newton step: the hardware estimate is biased by about -0.06 ulp;
    // after the step the error is an unbiased rounding (<= 1 ulp).
    __device__ __forceinline__ float rcp_refined(float d)
    {
      float r = __builtin_amdgcn_rcpf(d);
      const float e = fmaf(-d, r, 1.0f);
      return fmaf(r, e, r);
    }
    __device__ __forceinline__ float fast_rcpf(float v) { return __builtin_amdgcn_rcpf(v); }
    // FP64 reciprocal from an FP32 hardware seed and one FP64 Newton step (relative error about
    // (2^-23)^2, i.e. a few FP64 ulp; d must be within the FP32 range).
    __device__ __forceinline__ double rcp_seeded(double d)
    {
      const double r = (double)__builtin_amdgcn_rcpf((float)d);
      return fma(r, fma(-d, r, 1.0), r);
    }
    __device__ __forceinline__ float fast_sqrtf(float v) { return __builtin_amdgcn_sqrtf(v); }

    // exp(-t), t >= 0, to FP64 accuracy: t on a 1/64 grid (exp2t[j] = 2^(j/64), j < 64, a table the
    // caller keeps in LDS) and a degree-6 polynomial of the remainder |r| <= ln2/128 (truncation
    // below 1e-16 relative); t beyond 1020 ln2 gives 0 through the exponent.
    __device__ __forceinline__ double exp_neg_tab(double t, const double *exp2t)
    {
      constexpr double kLog2e = 1.4426950408889634074, kLn2 = 0.69314718055994530942;
      const double t64 = fmax(-t * (64.0 * kLog2e), -64.0 * 1020.0);
      const double tn = rint(t64);
      const int jn = (int)tn;
      const double r = (t64 - tn) * (kLn2 / 64.0);
      double pe = fma(r, 1.0 / 720.0, 1.0 / 120.0);
      pe = fma(pe, r, 1.0 / 24.0);
      pe = fma(pe, r, 1.0 / 6.0);
      pe = fma(pe, r, 0.5);
      pe = fma(pe, r, 1.0);
      pe = fma(pe, r, 1.0);
      return ldexp(pe * exp2t[jn & 63], jn >> 6);
    }

    // 1/sqrt(s) for s > 0 in the FP32 range: FP32 hardware seed and one FP64 Newton step on
    // f(y) = 1/y^2 - s, y <- y + y (1 - s y^2)/2 (relative error about the square of the seed's).
    __device__ __forceinline__ double rsqrt_seeded(double s)
    {
      const double y = (double)__builtin_amdgcn_rsqf((float)s);
      return fma(0.5 * y, fma(-s * y, y, 1.0), y);
    }
  } // namespace dev
} // namespace dfrg
