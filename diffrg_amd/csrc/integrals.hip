// integrals.hip -- batched momentum-loop integrals with the built-in integrands.
//
//   dfrg_threshold_batch   radial d-dim LPA threshold integrals (IntegratorGPU<d> form)
//   dfrg_se_batch_*        two-angle self-energy integrals of uneven orders in one launch
//                          (Integrator4D2AngGPU form); tile engine, autotuner, hipGraph
//   dfrg_ft_batch          finite-T two-propagator integrals (IntegratorFiniteTq0GPU form)
//
// Plain mode evaluates everything in FP64.  Mixed mode evaluates the integrand's transcendental
// and reciprocal parts in FP32 (hardware exp2/log2/rcp/sqrt), keeps nodes, weights, integration
// elements, the propagator denominators (which carry the cancellation near m^2 -> -k^2) and all
// accumulation in FP64, and reduces in a fixed tree (lane -> wave -> block -> tile, compensated).
#include "internal.hh"
#include "dfrg/device.hh"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <fstream>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

using namespace dfrg;

namespace
{
  constexpr double kPi = 3.14159265358979323846;
  constexpr double kLog2e = 1.4426950408889634074;
  constexpr double kLn2 = 0.69314718055994530942;

  double S_d_host(int d)
  {
    switch (d) {
    case 1: return 2.0;
    case 2: return 2.0 * kPi;
    case 3: return 4.0 * kPi;
    case 4: return 2.0 * kPi * kPi;
    case 5: return 8.0 * kPi * kPi / 3.0;
    case 6: return kPi * kPi * kPi;
    case 7: return 16.0 * kPi * kPi * kPi / 15.0;
    }
    return NAN;
  }

  __device__ __forceinline__ double ipow_dev(double x, int n)
  {
    double r = 1.0;
    for (int i = 0; i < n; ++i) r *= x;
    return r;
  }

  // ============================================================================================
  // C0: radial threshold integrals, one wave per integral
  // ============================================================================================
  template <int MODE>
  __global__ __launch_bounds__(256) void threshold_kernel(int d, const double *__restrict__ m2, int n_int,
                                                           const double *__restrict__ x, const double *__restrict__ w,
                                                           int nx, double x_extent, double k, double pref,
                                                           double *__restrict__ out)
  {
    const int lane = threadIdx.x & 63;
    const int integral = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (integral >= n_int) return; // whole wave exits together
    const double m = m2[integral];
    const double k2 = k * k;
    double acc = 0.0;
    for (int j = lane; j < nx; j += 64) {
      const double q = k * sqrt(x[j] * x_extent);
      const double q2 = q * q;
      // S_d q^{d-2}/2 k^2/(2pi)^d * w x_ext  (pref = S_d/(2pi)^d)
      const double el = pref * (ipow_dev(q, d - 2) / 2.0 * k2) * (w[j] * x_extent);
      const double xr = q2 / k2;
      if (MODE == DFRG_FP64) {
        const double e = exp(-xr);
        const double den = q2 + k2 * e + m;
        acc += el * (2.0 * e * k2 * (1.0 + xr)) / (den * den);
      } else {
        // the regulator enters the pole denominator q^2 + R + m2, which cancels as m2 -> -k^2:
        // R and the denominator stay FP64; reciprocal and products run in FP32
        // 64-node integrals cannot average FP32 roundings below 1e-8: the reciprocal is an FP32
        // hardware seed refined by one FP64 Newton step, everything else stays FP64
        const double e = exp(-xr);
        const double den = q2 + k2 * e + m;
        const double g = dev::rcp_seeded(den);
        acc += el * (2.0 * e * k2 * (1.0 + xr)) * (g * g);
      }
    }
    acc = dev::wave_sum(acc);
    if (lane == 0) out[integral] = acc;
  }

  // ============================================================================================
  // C2: finite-T two-propagator integrals, one block per integral
  // ============================================================================================
  struct FTArgs {
    int d, radial_mode, n_int, nx;
    const double *T, *p0, *pabs, *m1, *m2, *mx, *mw;
    const int *m_off, *m_n;
    const double *x, *w;
    double extent, k, pref;
    double *out;
  };

  template <int MODE, int BLOCK> __global__ __launch_bounds__(BLOCK) void ft_kernel(FTArgs A)
  {
    extern __shared__ double sm[];
    const int nx = A.nx;
    double *s_elw = sm;          // element * weight (plain) or element * weight * RBdot (mixed)
    double *s_rdot = sm + nx;    // RBdot(q^2)
    double *s_a1 = sm + 2 * nx;  // q^2 + R(q^2) + m1
    double *s_a2 = sm + 3 * nx;  // q^2 + |p|^2 + R(q^2 + |p|^2) + m2
    double *s_red = sm + 4 * nx; // block reduction
    const int i = blockIdx.x;
    const double k = A.k, k2 = k * k;
    const double T = A.T[i], p0 = A.p0[i], pa = A.pabs[i], m1 = A.m1[i], m2 = A.m2[i];
    const int off = A.m_off[i], nm = A.m_n[i];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    // this lane's first Matsubara node, loaded before the q-only terms so the two loads overlap
    double wn0 = 0.0, wgt0 = 0.0;
    if (lane < nm) {
      wn0 = A.mx[off + lane];
      wgt0 = A.mw[off + lane];
    }
    double acc = 0.0;
    // q-only terms, and the zero mode T f(q, 0) of sum = T f(0) + sum_n w_n (f(x_n) + f(-x_n))
    for (int j = threadIdx.x; j < nx; j += BLOCK) {
      double q, el;
      if (A.radial_mode == 0) {
        q = k * sqrt(A.x[j] * A.extent);
        el = A.pref * (ipow_dev(q, A.d - 2) / q / 2.0 * k2);
      } else {
        q = A.extent * A.x[j];
        el = A.pref * ipow_dev(q, A.d - 2);
      }
      const double q2 = q * q;
      const double elw = el * (A.w[j] * A.extent);
      const double xr = q2 / k2;
      const double rdot = 2.0 * exp(-xr) * k2 * (1.0 + xr);
      const double a1 = q2 + k2 * exp(-xr) + m1;
      const double qp2 = q2 + pa * pa;
      const double a2 = qp2 + k2 * exp(-(qp2 / k2)) + m2;
      s_elw[j] = MODE == DFRG_FP64 ? elw : elw * rdot;
      s_rdot[j] = rdot;
      s_a1[j] = a1;
      s_a2[j] = a2;
      if (T > 0) {
        double g10, g20;
        if (MODE == DFRG_FP64) {
          g10 = 1.0 / a1;
          g20 = 1.0 / (p0 * p0 + a2);
        } else {
          g10 = dev::rcp_seeded(a1);
          g20 = dev::rcp_seeded(p0 * p0 + a2);
        }
        acc += elw * (T * (rdot * g10 * g10 * g20));
      }
    }
    __syncthreads();
    // Matsubara nodes across the lanes (fixed per lane), radial nodes across the waves; the
    // q-only terms are wave-uniform LDS reads
    constexpr int NW = BLOCK / 64;
    for (int n0 = 0; n0 < nm; n0 += 64) {
      const int n = n0 + lane;
      if (n >= nm) break;
      const double wn = n0 == 0 ? wn0 : A.mx[off + n], wgt = n0 == 0 ? wgt0 : A.mw[off + n];
      const double dp = wn + p0, dm = -wn + p0;
      if (MODE == DFRG_FP64) {
        // the reference's expression tree: three FP64 divisions per pair
        double s = 0.0;
#pragma unroll 2
        for (int j = wv; j < nx; j += NW) {
          const double a1 = s_a1[j], a2 = s_a2[j], rd = s_rdot[j];
          const double g1 = 1.0 / (wn * wn + a1), gp = 1.0 / (dp * dp + a2), gm = 1.0 / (dm * dm + a2);
          s += s_elw[j] * (wgt * (rd * g1 * g1 * gp + rd * g1 * g1 * gm));
        }
        acc += s;
      } else {
        // f(q,w) + f(q,-w) = RBdot (d2p + d2m) / (d1^2 d2p d2m): one reciprocal per pair, from an
        // FP32 hardware seed refined in FP64; denominators and products in FP64
        const double w2 = wn * wn, dp2 = dp * dp, dm2 = dm * dm;
        double s = 0.0;
#pragma unroll 4
        for (int j = wv; j < nx; j += NW) {
          const double a2 = s_a2[j];
          const double d1 = w2 + s_a1[j];
          const double d2p = dp2 + a2, d2m = dm2 + a2;
          const double D2 = d2p * d2m;
          const double r = dev::rcp_seeded(d1 * D2);
          s = fma(s_elw[j], (d2p + d2m) * (r * (r * D2)), s);
        }
        acc = fma(wgt, s, acc);
      }
    }
    const double tot = dev::block_sum(acc, s_red);
    if (threadIdx.x == 0) A.out[i] = tot;
  }

  // ============================================================================================
  // C1/C3: two-angle self-energy tile engine
  // ============================================================================================
  // A tile is (integral, range of radial nodes).  Everything a block needs about its tile sits in
  // one record, so the block starts with a single load; the table, angular terms and the tile's
  // radial nodes then come in one round of independent loads into LDS.  When every integral of a
  // batch shares nodes, table and tile count, the record follows from the tile index and one
  // prototype in the kernel arguments (SEArgs::shared), and that round is the block's first.
  struct SETile {
    double p, c, inv_b, scale; // external momentum; log coordinates: start c, 1/b, (n-1)/bias
    int integral, q0, nqt, slot;
    int na, l2na, xq_off, xa_off; // node offsets: x at off, w at off + order
    int nq, tab_off, ntab, first_slot;
    int n_slots, narrow, pad1, pad2; // narrow: the mixed path may use se_G_mixed<true>
  };

  struct SECellZ {
    double z, dz; // Z_i and Z_{i+1} - Z_i
  };

  // mixed-path cell data of node i: y_i^2 (FP64), 1/(b + y_i - c) and y_i (FP32)
  struct SECellY {
    double y2;
    float inv_span, y;
  };

  struct SEArgs {
    int n_tiles;
    const SETile *tiles;
    const double *nodes;
    const double *tab64;  // FP64 tables
    const SECellY *tabA;  // mixed-path cell data: (y_i^2, 1/(b + y_i - c), y_i)
    const SECellZ *tabB;  // mixed-path cell data: (Z_i, Z_{i+1} - Z_i)
    double x_extent, k, pref; // pref = 2pi/(2pi)^4
    double *partials;   // one line per tile (dev::kSlotStride)
    unsigned *counters; // one line per integral (dev::kCounterStride)
    unsigned *queue;    // [0] next tile, [kCounterStride] exited blocks (persistent mode)
    int persistent;
    int xcd_map; // one tile per block: give each XCD a contiguous run of tiles
    double *out;
    // shared layout: every integral has the same nodes, table and tile count, so tile t is proto
    // with integral t / tiles_per_int, radial start (t mod tiles_per_int) nqt, slot t, and p from
    // p_ext: a block issues its table and node loads without waiting for a tile record
    int shared, tiles_per_int;
    const double *p_ext;
    SETile proto;
  };

  // Blocks are dispatched to the eight XCDs in turn (block b to XCD b mod 8).  With xcd_map, the
  // blocks of XCD x take the x-th contiguous run of the tile list, so the tiles of an integral --
  // consecutive in the list -- share one L2 (its table staged once, its partials local).  A bijection
  // of [0, G): XCD x owns per + (x < rem) blocks and as many consecutive tiles.
  __device__ __forceinline__ int xcd_tile(int b, int G)
  {
    const int per = G >> 3, rem = G & 7, x = b & 7, k = b >> 3;
    return x * per + (x < rem ? x : rem) + k;
  }

  constexpr int kMaxTab = 256;
  constexpr int kMaxAng = 64;
  constexpr int kMaxQ = 128;

  // G(|p-q|) = 1/(Z(|p-q|) |p-q|^2 + R(|p-q|^2)) in FP64, the reference's expression tree.
  __device__ __forceinline__ double se_G_fp64(const double *tab, int ntab, double c, double inv_b, double scale,
                                              double k2, double pq2)
  {
    const double y = sqrt(pq2);
    const double Z = dev::lerp_log_f64(tab, ntab, c, inv_b, scale, y);
    return 1.0 / (Z * pq2 + k2 * exp(-(pq2 / k2)));
  }

  struct SEMixed {
    const SECellY *tabA;
    const SECellZ *tabB;
    const double *exp2t; // 2^(j/64), j = 0..63
    float cf, ibf, scl, sf;
    double ek64;
    float imaxf; // n - 2
  };

  // G(|p-q|) in mixed precision: FP32 hardware seeds replace the FP64 library sequences, FP64 is
  // kept where the error budget needs it.  y = |p-q| as an FP32 seed finds the Z cell through the
  // FP32 log2; the in-cell offset y - y_i = (|p-q|^2 - y_i^2)/(y + y_i) takes the difference in
  // FP64 (no cancellation error) and the rest in FP32; the in-cell position s*log1p(u)
  // (|u| <~ 0.05) in FP32 with the rounding of 1+u restored, Z assembled in FP64 from the FP64 node
  // value; R = k^2 2^t with t split on a 1/64 grid (table of k^2 2^(j/64)) and a degree-3
  // polynomial remainder in FP64 (|r| <= ln2/128: truncation below 4e-11 relative; 2^t for t
  // below -2^25/64 underflows to 0 through the exponent);
  // the denominator in FP64; the reciprocal from an FP32 seed refined in FP64.  Per-evaluation
  // error ~1e-9 relative.
  // NARROW (chosen per integral when its table nodes are >= 1e-30 and |64 t| < 2^30 at all its
  // quadrature points): y + y_i needs no guard, and 64 t is rounded to an integer by adding
  // 1.5 2^52 (the sum's low word is the integer), one FMA instead of round-and-convert.
  template <bool NARROW> __device__ __forceinline__ double se_G_mixed(const SEMixed &m, double pq2d)
  {
    const float ys = __builtin_amdgcn_sqrtf((float)pq2d);
    // cell index: clamp to [0, n-2] first (med3 of a NaN gives 0), then truncation is floor
    const float coarse = __builtin_amdgcn_logf(fmaf(ys - m.cf, m.ibf, 1.0f)) * m.scl;
    const int i0 = (int)__builtin_amdgcn_fmed3f(coarse, 0.0f, m.imaxf); // NaN gives 0
    const SECellY ca = m.tabA[i0];
    const SECellZ cz = m.tabB[i0];
    const float num = (float)(pq2d - ca.y2);
    const float u = num * (ca.inv_span * __builtin_amdgcn_rcpf(NARROW ? ys + ca.y : fmaxf(ys + ca.y, 1e-30f)));
    const float w1 = 1.0f + u;
    const float ru = u - (w1 - 1.0f); // rounding error of 1 + u
    float fr = fmaf(__builtin_amdgcn_logf(w1), m.scl, ru * m.sf);
    fr = __builtin_amdgcn_fmed3f(fr, 0.0f, 1.0f); // realises the [0, n-1] clamp
    const double Z = fma((double)fr, cz.dz, cz.z);
    // 64 t = pq2 ek64, t = -|p-q|^2/k^2 log2(e); jn = round(64 t), f = 64 t - jn in [-1/2, 1/2]
    int jn;
    double f;
    if (NARROW) {
      constexpr double kRound = 6755399441055744.0; // 1.5 * 2^52
      const double sh = fma(pq2d, m.ek64, kRound);
      jn = (int)(unsigned)__double_as_longlong(sh);
      f = fma(pq2d, m.ek64, -(sh - kRound));
    } else {
      const double t64 = pq2d * m.ek64;
      const double tn = rint(t64);
      jn = (int)tn;
      f = t64 - tn;
    }
    // 2^(f/64) = exp(f ln2/64) as a cubic in f
    constexpr double c1 = kLn2 / 64.0, c2 = c1 * c1 / 2.0, c3 = c1 * c1 * c1 / 6.0;
    const double pe = fma(fma(fma(f, c3, c2), f, c1), f, 1.0);
    const double R = ldexp(pe * m.exp2t[jn & 63], jn >> 6);
    return dev::rcp_seeded(fma(Z, pq2d, R));
  }

  // log1p(t) to FP64 accuracy from the FP32 hardware log2 and one Newton step on e^l = 1 + t,
  // with 1 + t carried exactly as w + err (TwoSum).  Quadratic convergence from the ~1e-7 seed
  // leaves ~1e-14 absolute.  t <= -1 gives NaN, which the callers' clamps map to index 0.
  __device__ __forceinline__ double log1p_seeded(double t, const double *exp2t)
  {
    const double w = 1.0 + t;
    const double bv = w - 1.0;
    const double err = (1.0 - (w - bv)) + (t - bv);
    const double l0 = (double)(__builtin_amdgcn_logf((float)w) * (float)kLn2);
    const double e = dev::exp_neg_tab(l0, exp2t);
    return l0 + fma(w, e, -1.0) + err * e;
  }

  // Sum over the block's share of (cos1, cos2) pairs and radial nodes of one tile.  Lanes own
  // pairs; with two or more pairs per lane, each radial step evaluates two pairs that share the
  // LDS reads of the q terms; with one pair per lane, lanes split the radial nodes and each step
  // evaluates two of them.  Two independent chains per step hide the LDS latency.  UNROLL = 2
  // takes two radial nodes per step of the paired loop (four chains), for launches with few waves
  // per SIMD; the accumulation order, and so the result, is the same.
  template <int BLOCK, int UNROLL, typename G>
  __device__ __forceinline__ double se_tile_sum(const G &g, int na, int l2, int nqt, const double2 *s_ang,
                                                const double *s_w2, const double *s_qs, const double *s_qc,
                                                const double *s_qt)
  {
    const int npairs = na * na;
    const int pstride = npairs < BLOCK ? npairs : BLOCK;
    const int qlanes = BLOCK / pstride;
    const int pr0 = threadIdx.x & (pstride - 1);
    const int qs0 = threadIdx.x / pstride;
    double acc = 0.0;
    if (2 * pstride <= npairs) {
      for (int pr = pr0; pr < npairs; pr += 2 * pstride) {
        const int prb = pr + pstride;
        const double2 a = s_ang[pr >> l2], b = s_ang[prb >> l2];
        double sa = 0.0, sb = 0.0;
        int j = 0;
        if constexpr (UNROLL > 1)
          for (; j + UNROLL - 1 < nqt; j += UNROLL) {
            double ga[UNROLL], gb[UNROLL];
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) {
              const double qs = s_qs[j + u], qc = s_qc[j + u];
              ga[u] = g(fma(qc, a.x, qs));
              gb[u] = g(fma(qc, b.x, qs));
            }
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) {
              const double qt = s_qt[j + u];
              sa = fma(qt, ga[u], sa);
              sb = fma(qt, gb[u], sb);
            }
          }
        for (; j < nqt; ++j) {
          const double qs = s_qs[j], qc = s_qc[j], qt = s_qt[j];
          const double ga = g(fma(qc, a.x, qs));
          const double gb = g(fma(qc, b.x, qs));
          sa = fma(qt, ga, sa);
          sb = fma(qt, gb, sb);
        }
        acc += a.y * s_w2[pr & (na - 1)] * sa;
        acc += b.y * s_w2[prb & (na - 1)] * sb;
      }
    } else {
      const double2 a = s_ang[pr0 >> l2];
      double sa = 0.0, sb = 0.0;
      int j = qs0;
      for (; j + qlanes < nqt; j += 2 * qlanes) {
        const int j2 = j + qlanes;
        const double ga = g(fma(s_qc[j], a.x, s_qs[j]));
        const double gb = g(fma(s_qc[j2], a.x, s_qs[j2]));
        sa = fma(s_qt[j], ga, sa);
        sb = fma(s_qt[j2], gb, sb);
      }
      if (j < nqt) sa = fma(s_qt[j], g(fma(s_qc[j], a.x, s_qs[j])), sa);
      acc = a.y * s_w2[pr0 & (na - 1)] * (sa + sb);
    }
    return acc;
  }

  template <int MODE, int BLOCK, int UNROLL> __global__ __launch_bounds__(BLOCK) void se_kernel(SEArgs A)
  {
    constexpr bool F64 = MODE == DFRG_FP64;
    __shared__ double s_tab64[F64 ? kMaxTab : 1];
    __shared__ SECellY s_tabA[F64 ? 1 : kMaxTab];
    __shared__ SECellZ s_tabB[F64 ? 1 : kMaxTab];
    __shared__ double s_exp2[F64 ? 1 : 64];  // 2^(j/64)
    __shared__ double s_exp2k[F64 ? 1 : 64]; // k^2 2^(j/64)
    __shared__ double2 s_ang[kMaxAng]; // (cos1, 2 w1 sqrt(1 - cos1^2))
    __shared__ double s_w2[kMaxAng];   // 2 w2
    __shared__ double s_qs[kMaxQ], s_qc[kMaxQ], s_qt[kMaxQ];
    __shared__ double2 s_qw[kMaxQ]; // (q, w_q x_extent)
    __shared__ double s_red[BLOCK / 64];
    __shared__ int s_tile;

    // 2^(j/64) in the block's last wave, which stages nothing, so wave 0's loads are not delayed
    if (!F64) {
      const int j = (int)threadIdx.x - (BLOCK - 64);
      if (j >= 0) {
        s_exp2[j] = exp2(j * (1.0 / 64.0));
        s_exp2k[j] = A.k * A.k * s_exp2[j];
      }
    }
    const double k = A.k, k2 = k * k;
    int cur = -1; // integral whose table and angular terms are in LDS (block-uniform)
    for (int iter = 0;; ++iter) {
      int t;
      if (A.persistent) {
        __syncthreads(); // the previous tile is finished with s_tile
        if (threadIdx.x == 0) s_tile = (int)atomicAdd(&A.queue[0], 1u);
        __syncthreads();
        t = __builtin_amdgcn_readfirstlane(s_tile);
      } else {
        t = iter == 0 ? (A.xcd_map ? xcd_tile((int)blockIdx.x, (int)gridDim.x) : (int)blockIdx.x) : A.n_tiles;
      }
      if (t >= A.n_tiles) break;
      SETile T;
      if (A.shared) {
        T = A.proto;
        T.integral = t / A.tiles_per_int;
        T.q0 = (t - T.integral * A.tiles_per_int) * T.nqt;
        T.slot = t;
        T.first_slot = T.integral * A.tiles_per_int;
      } else {
        T = A.tiles[t];
      }
      const int i = T.integral;
      // one round of independent loads: table and angular terms (when the integral changes) and
      // the radial nodes of this tile
      if (i != cur) {
        if (F64) {
          for (int j = threadIdx.x; j < T.ntab; j += BLOCK) s_tab64[j] = A.tab64[T.tab_off + j];
        } else {
          for (int j = threadIdx.x; j < T.ntab; j += BLOCK) {
            s_tabA[j] = A.tabA[T.tab_off + j];
            s_tabB[j] = A.tabB[T.tab_off + j];
          }
        }
        const double *xa = A.nodes + T.xa_off;
        for (int j = threadIdx.x; j < T.na; j += BLOCK) {
          const double c1 = 2.0 * (xa[j] - 0.5);
          const double wa = xa[T.na + j];
          s_ang[j] = make_double2(c1, 2.0 * wa * sqrt(1.0 - c1 * c1)); // cos1 weight and Jacobian
          s_w2[j] = 2.0 * wa;                                           // cos2 weight
        }
      }
      const double *xq = A.nodes + T.xq_off;
      const double pe = A.shared ? A.p_ext[i] : T.p;
      for (int j = threadIdx.x; j < T.nqt; j += BLOCK) {
        const double q = k * sqrt(xq[T.q0 + j] * A.x_extent);
        s_qs[j] = pe * pe + q * q;
        s_qc[j] = -2.0 * pe * q;
        s_qw[j] = make_double2(q, xq[T.nq + T.q0 + j] * A.x_extent);
      }
      __syncthreads();
      // q-only factor from the staged table: el * w_q x_ext * RBdot(q) G(q)^2.  Plain mode: the
      // reference's FP64 expression tree.  Mixed mode: FP64-accurate seeded forms (FP32 hardware
      // log2 / reciprocal seeds with FP64 Newton steps, table exp), since this factor multiplies
      // every evaluation of its radial node and its error would not average out.
      for (int j = threadIdx.x; j < T.nqt; j += BLOCK) {
        const double2 qw = s_qw[j];
        const double q = qw.x, q2 = q * q;
        double qt;
        if (F64) {
          const double xr = q2 / k2;
          const double ex = exp(-xr);
          const double Zq = dev::lerp_log_f64(s_tab64, T.ntab, T.c, T.inv_b, T.scale, q);
          const double Gq = 1.0 / (Zq * q2 + k2 * ex);
          const double el = A.pref * (q2 * 0.5 * k2);
          qt = el * qw.y * (2.0 * ex * k2 * (1.0 + xr)) * Gq * Gq;
        } else {
          const double xr = q2 * dev::rcp_seeded(k2);
          const double ex = dev::exp_neg_tab(xr, s_exp2);
          double idx = log1p_seeded((q - T.c) * T.inv_b, s_exp2) * T.scale;
          idx = fmax(0.0, fmin(idx, (double)(T.ntab - 1)));
          const double fl = floor(idx);
          const int i0 = (int)fl, i1 = (int)ceil(idx);
          const double Zq = s_tabB[i0].z * (1.0 - idx + fl) + s_tabB[i1].z * (idx - fl);
          const double Gq = dev::rcp_seeded(Zq * q2 + k2 * ex);
          const double el = A.pref * (q2 * 0.5 * k2);
          qt = el * qw.y * (2.0 * ex * k2 * (1.0 + xr)) * Gq * Gq;
        }
        s_qt[j] = qt;
      }
      cur = i;
      __syncthreads();

      double acc;
      if (F64) {
        const auto g = [&](double pq2) { return se_G_fp64(s_tab64, T.ntab, T.c, T.inv_b, T.scale, k2, pq2); };
        acc = se_tile_sum<BLOCK, UNROLL>(g, T.na, T.l2na, T.nqt, s_ang, s_w2, s_qs, s_qc, s_qt);
      } else {
        SEMixed m;
        m.tabA = s_tabA;
        m.tabB = s_tabB;
        m.exp2t = s_exp2k;
        m.cf = (float)T.c;
        m.ibf = (float)T.inv_b;
        m.scl = (float)(T.scale * kLn2);
        m.sf = (float)T.scale;
        m.ek64 = -64.0 * kLog2e / k2;
        m.imaxf = (float)(T.ntab - 2);
        if (T.narrow) { // tile-uniform
          const auto g = [&](double pq2) { return se_G_mixed<true>(m, pq2); };
          acc = se_tile_sum<BLOCK, UNROLL>(g, T.na, T.l2na, T.nqt, s_ang, s_w2, s_qs, s_qc, s_qt);
        } else {
          const auto g = [&](double pq2) { return se_G_mixed<false>(m, pq2); };
          acc = se_tile_sum<BLOCK, UNROLL>(g, T.na, T.l2na, T.nqt, s_ang, s_w2, s_qs, s_qc, s_qt);
        }
      }
      const double part = dev::block_sum(acc, s_red);
      double total = 0.0;
      if (dev::finish_group(part, A.partials, T.slot, T.first_slot, T.n_slots,
                            &A.counters[(long long)i * dev::kCounterStride], &total)) {
        if (threadIdx.x == 0) A.out[i] = total;
      }
    }
    // persistent mode: the last block out resets the queue for the next launch / replay
    // (every earlier queue access of this block is an atomic whose result has been used)
    if (A.persistent && threadIdx.x == 0) {
      const unsigned done = dev::take_ticket(&A.queue[dev::kCounterStride]);
      if (done == gridDim.x - 1) {
        dev::reset_counter(&A.queue[0]);
        dev::reset_counter(&A.queue[dev::kCounterStride]);
      }
    }
  }

  template <int MODE, int UNROLL> hipError_t launch_se(int block, int grid, hipStream_t s, const SEArgs &a)
  {
    switch (block) {
    case 64: hipLaunchKernelGGL((se_kernel<MODE, 64, UNROLL>), dim3(grid), dim3(64), 0, s, a); break;
    case 128: hipLaunchKernelGGL((se_kernel<MODE, 128, UNROLL>), dim3(grid), dim3(128), 0, s, a); break;
    case 256: hipLaunchKernelGGL((se_kernel<MODE, 256, UNROLL>), dim3(grid), dim3(256), 0, s, a); break;
    case 512: hipLaunchKernelGGL((se_kernel<MODE, 512, UNROLL>), dim3(grid), dim3(512), 0, s, a); break;
    case 1024: hipLaunchKernelGGL((se_kernel<MODE, 1024, UNROLL>), dim3(grid), dim3(1024), 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }

  int ilog2(int v)
  {
    int l = 0;
    while ((1 << l) < v) ++l;
    return l;
  }
  bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

  std::atomic<int> g_ft_block{0}; // threads per integral of dfrg_ft_batch, 0 = automatic

  // Zeroes the first word of n_lines consecutive 128-byte lines (ticket counters and queue words).
  __global__ __launch_bounds__(256) void clear_lines_kernel(unsigned *lines, int n_lines)
  {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_lines) lines[(long long)i * dev::kCounterStride] = 0u;
  }
} // namespace

// ----------------------------------------------------------------------------------------------
// self-energy batch object
// ----------------------------------------------------------------------------------------------
struct dfrg_se_batch {
  dfrg_qp *qp = nullptr;
  int n_int = 0;
  double x_extent = 20.0, k = 1.0;
  // per integral (host)
  std::vector<int> nq, na, xq_off, xa_off, tab_off, tab_n;
  std::vector<double> p, coord; // coord: c, 1/b, (n-1)/bias
  unsigned long long evals = 0;
  // device data
  // one allocation: nodes, tab64, tabA, tabB, p, ticket counters, queue (256-byte aligned sections)
  void *d_inputs = nullptr;
  double *d_nodes = nullptr, *d_tab64 = nullptr, *d_p = nullptr;
  SECellY *d_tabA = nullptr;
  SECellZ *d_tabB = nullptr;
  SETile *d_tiles = nullptr;
  double *d_partials = nullptr;
  unsigned *d_counters = nullptr, *d_queue = nullptr;
  size_t tiles_cap = 0;
  // shape
  int block = 256, tile_evals = 8192, grid = 0, n_tiles = 0;
  std::vector<int> narrow; // per integral: the mixed path may use se_G_mixed<true>
  int uniform_tiles = 0;    // every tile has the same evaluation count
  int shared_layout = 0;    // every tile record follows from proto and its index (SEArgs::shared)
  int tiles_per_int = 0;
  SETile proto{};
  std::string signature;
  // graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int g_mode = -1, g_launches = 0, g_block = 0, g_tile = 0, g_grid = -1;
  double *g_out = nullptr;
  hipStream_t g_stream = nullptr;
};

namespace
{
  int se_build_tiles(dfrg_se_batch *b)
  {
    const int n = b->n_int;
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    // longest integrals first (LPT), ties by index: deterministic
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
      return (long long)b->nq[a] * b->na[a] * b->na[a] > (long long)b->nq[c] * b->na[c] * b->na[c];
    });
    std::vector<SETile> tiles;
    int slot = 0;
    for (int i : order) {
      const int per_q = b->na[i] * b->na[i];
      int qpt = std::max(1, std::min(b->nq[i], b->tile_evals / per_q));
      qpt = std::min(qpt, kMaxQ);
      const int first = slot;
      const int cnt = (b->nq[i] + qpt - 1) / qpt;
      for (int q0 = 0; q0 < b->nq[i]; q0 += qpt) {
        SETile t{};
        t.p = b->p[i];
        t.c = b->coord[3 * i];
        t.inv_b = b->coord[3 * i + 1];
        t.scale = b->coord[3 * i + 2];
        t.integral = i;
        t.q0 = q0;
        t.nqt = std::min(qpt, b->nq[i] - q0);
        t.slot = slot++;
        t.na = b->na[i];
        t.l2na = ilog2(b->na[i]);
        t.xq_off = b->xq_off[i];
        t.xa_off = b->xa_off[i];
        t.nq = b->nq[i];
        t.tab_off = b->tab_off[i];
        t.ntab = b->tab_n[i];
        t.first_slot = first;
        t.n_slots = cnt;
        t.narrow = b->narrow[i];
        tiles.push_back(t);
      }
    }
    b->n_tiles = (int)tiles.size();
    b->uniform_tiles = 1;
    for (const SETile &t : tiles)
      if ((long long)t.nqt * t.na * t.na != (long long)tiles[0].nqt * tiles[0].na * tiles[0].na) b->uniform_tiles = 0;
    // shared layout: the kernel's tile from proto and the index equals every record here
    b->shared_layout = 0;
    if (!tiles.empty()) {
      SETile pr = tiles[0];
      pr.p = 0.0;
      pr.integral = pr.q0 = pr.slot = pr.first_slot = 0;
      const int tpi = pr.n_slots;
      bool same = tpi > 0 && (long long)tpi * n == (long long)tiles.size();
      for (int t = 0; same && t < (int)tiles.size(); ++t) {
        SETile c = pr;
        c.integral = t / tpi;
        c.q0 = (t - c.integral * tpi) * pr.nqt;
        c.slot = t;
        c.first_slot = c.integral * tpi;
        c.p = b->p[c.integral];
        same = std::memcmp(&c, &tiles[t], sizeof(SETile)) == 0;
      }
      if (same) {
        b->shared_layout = 1;
        b->tiles_per_int = tpi;
        b->proto = pr;
      }
    }
    if ((size_t)b->n_tiles > b->tiles_cap) {
      hipFree(b->d_tiles);
      hipFree(b->d_partials);
      b->d_tiles = nullptr;
      b->d_partials = nullptr;
      b->tiles_cap = b->n_tiles;
      DFRG_HIP(hipMalloc((void **)&b->d_tiles, sizeof(SETile) * b->tiles_cap));
      DFRG_HIP(hipMalloc((void **)&b->d_partials, sizeof(double) * dev::kSlotStride * b->tiles_cap));
    }
    DFRG_HIP(hipMemcpy(b->d_tiles, tiles.data(), sizeof(SETile) * tiles.size(), hipMemcpyHostToDevice));
    return DFRG_OK;
  }

  SEArgs se_args(dfrg_se_batch *b, double *d_out)
  {
    SEArgs a;
    a.n_tiles = b->n_tiles;
    a.tiles = b->d_tiles;
    a.nodes = b->d_nodes;
    a.tab64 = b->d_tab64;
    a.tabA = b->d_tabA;
    a.tabB = b->d_tabB;
    a.x_extent = b->x_extent;
    a.k = b->k;
    a.pref = 2.0 * kPi / ((2.0 * kPi) * ((2.0 * kPi) * ((2.0 * kPi) * (2.0 * kPi))));
    a.partials = b->d_partials;
    a.counters = b->d_counters;
    a.queue = b->d_queue;
    a.persistent = b->grid > 0 ? 1 : 0;
    // equal tiles only: with uneven tiles the longest-first order balances the XCDs better
    a.xcd_map = !a.persistent && b->uniform_tiles;
    a.out = d_out;
    a.shared = b->shared_layout;
    a.tiles_per_int = b->tiles_per_int > 0 ? b->tiles_per_int : 1;
    a.p_ext = b->d_p;
    a.proto = b->proto;
    return a;
  }

  // Radial unroll of the mixed paired loop: 2 when the launch puts at most 16 waves on a CU (4 per
  // SIMD), where four independent evaluation chains per lane hide more latency than two; 1 for
  // fuller launches, which hide it with waves.  Results do not depend on it.
  int se_unroll(const dfrg_se_batch *b)
  {
    const long long cus = context().cu_count > 0 ? context().cu_count : 256;
    const long long blocks = b->grid > 0 ? std::min(b->grid, b->n_tiles) : b->n_tiles;
    return blocks * (b->block / 64) <= 16 * cus ? 2 : 1;
  }

  // Cross-block scratch of one launch: partial slots, ticket counters and the work queue.
  struct SEScratch {
    double *partials;
    unsigned *counters, *queue;
  };

  int se_launch(dfrg_se_batch *b, int mode, double *d_out, hipStream_t s, const SEScratch *scr = nullptr)
  {
    SEArgs a = se_args(b, d_out);
    if (scr) {
      a.partials = scr->partials;
      a.counters = scr->counters;
      a.queue = scr->queue;
    }
    const int grid = b->grid > 0 ? std::min(b->grid, b->n_tiles) : b->n_tiles;
    hipError_t e = mode == DFRG_FP64    ? launch_se<DFRG_FP64, 1>(b->block, grid, s, a)
                   : se_unroll(b) == 2 ? launch_se<DFRG_MIXED, 2>(b->block, grid, s, a)
                                       : launch_se<DFRG_MIXED, 1>(b->block, grid, s, a);
    if (e != hipSuccess) {
      set_error("launch of se_kernel failed: %s", hipGetErrorString(e));
      return DFRG_ERR_HIP;
    }
    return DFRG_OK;
  }

  void se_drop_graph(dfrg_se_batch *b)
  {
    if (b->graph_exec) hipGraphExecDestroy(b->graph_exec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->graph_exec = nullptr;
    b->graph = nullptr;
    b->g_mode = -1;
  }
} // namespace

extern "C" {

int dfrg_threshold_batch(int d, const double *d_m2, int n_int, const double *d_x, const double *d_w, int nx,
                         double x_extent, double k, int mode, double *d_out, void *stream)
{
  DFRG_TRY(require_init());
  DFRG_CHECK_ARG(d >= 2 && d <= 7, "dimension %d out of range [2,7]", d);
  DFRG_CHECK_ARG(n_int >= 0 && nx >= 1, "invalid sizes n_int=%d nx=%d", n_int, nx);
  DFRG_CHECK_ARG(mode == DFRG_FP64 || mode == DFRG_MIXED, "invalid mode %d", mode);
  if (n_int == 0) return DFRG_OK;
  DFRG_CHECK_ARG(d_m2 && d_x && d_w && d_out, "NULL device pointer");
  const double pref = S_d_host(d) / std::pow(2.0 * kPi, d);
  const int block = 256;
  const long long threads = (long long)n_int * 64;
  const int grid = (int)((threads + block - 1) / block);
  hipStream_t s = resolve_stream(stream);
  if (mode == DFRG_FP64)
    hipLaunchKernelGGL(threshold_kernel<DFRG_FP64>, dim3(grid), dim3(block), 0, s, d, d_m2, n_int, d_x, d_w, nx,
                       x_extent, k, pref, d_out);
  else
    hipLaunchKernelGGL(threshold_kernel<DFRG_MIXED>, dim3(grid), dim3(block), 0, s, d, d_m2, n_int, d_x, d_w, nx,
                       x_extent, k, pref, d_out);
  DFRG_LAUNCHED("threshold_kernel");
  return DFRG_OK;
}

int dfrg_ft_set_block(int block)
{
  DFRG_CHECK_ARG(block == 0 || block == 64 || block == 128 || block == 256 || block == 512,
                 "block %d not in {0,64,128,256,512}", block);
  g_ft_block = block;
  return DFRG_OK;
}

int dfrg_ft_batch(int d, int radial_mode, int n_int, const double *d_T, const double *d_p0, const double *d_pabs,
                  const double *d_m1, const double *d_m2, const double *d_mx, const double *d_mw, const int *d_m_off,
                  const int *d_m_n, const double *d_x, const double *d_w, int nx, double extent, double k, int mode,
                  double *d_out, void *stream)
{
  DFRG_TRY(require_init());
  DFRG_CHECK_ARG(d >= 2 && d <= 7, "dimension %d out of range [2,7]", d);
  DFRG_CHECK_ARG(radial_mode == 0 || radial_mode == 1, "radial_mode must be 0 or 1");
  DFRG_CHECK_ARG(nx >= 1 && nx <= 1024, "radial order %d out of range [1,1024]", nx);
  DFRG_CHECK_ARG(mode == DFRG_FP64 || mode == DFRG_MIXED, "invalid mode %d", mode);
  DFRG_CHECK_ARG(n_int >= 0, "invalid size n_int=%d", n_int);
  if (n_int == 0) return DFRG_OK;
  DFRG_CHECK_ARG(d_T && d_p0 && d_pabs && d_m1 && d_m2 && d_mx && d_mw && d_m_off && d_m_n && d_x && d_w && d_out,
                 "NULL device pointer");
  FTArgs a{d, radial_mode, n_int, nx, d_T, d_p0, d_pabs, d_m1, d_m2, d_mx, d_mw, d_m_off, d_m_n, d_x, d_w,
           extent, k, 0.0, d_out};
  a.pref = S_d_host(d - 1) / std::pow(2.0 * kPi, d - 1);
  // threads per integral: the caller's choice (dfrg_ft_set_block), else enough waves for about
  // two per SIMD over the whole batch, 1 to 8 per integral and not more than radial nodes
  int B = g_ft_block;
  if (B == 0) {
    const int simds = 4 * (context().cu_count > 0 ? context().cu_count : 256);
    int nw = 1;
    while (nw < 8 && nw * 2 <= nx && (long long)n_int * nw < 2LL * simds) nw *= 2;
    B = 64 * nw;
  }
  hipStream_t s = resolve_stream(stream);
  const size_t shm = sizeof(double) * (4 * nx + B / 64);
  if (mode == DFRG_FP64) {
    switch (B) {
    case 64: hipLaunchKernelGGL((ft_kernel<DFRG_FP64, 64>), dim3(n_int), dim3(64), shm, s, a); break;
    case 128: hipLaunchKernelGGL((ft_kernel<DFRG_FP64, 128>), dim3(n_int), dim3(128), shm, s, a); break;
    case 256: hipLaunchKernelGGL((ft_kernel<DFRG_FP64, 256>), dim3(n_int), dim3(256), shm, s, a); break;
    default: hipLaunchKernelGGL((ft_kernel<DFRG_FP64, 512>), dim3(n_int), dim3(512), shm, s, a); break;
    }
  } else {
    switch (B) {
    case 64: hipLaunchKernelGGL((ft_kernel<DFRG_MIXED, 64>), dim3(n_int), dim3(64), shm, s, a); break;
    case 128: hipLaunchKernelGGL((ft_kernel<DFRG_MIXED, 128>), dim3(n_int), dim3(128), shm, s, a); break;
    case 256: hipLaunchKernelGGL((ft_kernel<DFRG_MIXED, 256>), dim3(n_int), dim3(256), shm, s, a); break;
    default: hipLaunchKernelGGL((ft_kernel<DFRG_MIXED, 512>), dim3(n_int), dim3(512), shm, s, a); break;
    }
  }
  DFRG_LAUNCHED("ft_kernel");
  return DFRG_OK;
}

int dfrg_se_batch_create(dfrg_se_batch **out, dfrg_qp *qp, int n_int, const double *p, const int *nq, const int *na,
                         const int *tab_off, const int *tab_n, const double *start, const double *stop,
                         const double *bias, const double *tables, int tables_len, double x_extent, double k)
{
  DFRG_TRY(require_init());
  DFRG_CHECK_ARG(out && qp, "NULL handle");
  DFRG_CHECK_ARG(n_int >= 1, "n_int must be positive");
  DFRG_CHECK_ARG(p && nq && na && tab_off && tab_n && start && stop && bias && tables, "NULL input array");
  DFRG_CHECK_ARG(x_extent > 0 && k > 0, "x_extent and k must be positive");
  for (int i = 0; i < n_int; ++i) {
    DFRG_CHECK_ARG(nq[i] >= 1 && nq[i] <= 1024, "integral %d: radial order %d out of range", i, nq[i]);
    DFRG_CHECK_ARG(is_pow2(na[i]) && na[i] >= 2 && na[i] <= kMaxAng,
                   "integral %d: angular order %d must be a power of two in [2,%d]", i, na[i], kMaxAng);
    DFRG_CHECK_ARG(tab_n[i] >= 2 && tab_n[i] <= kMaxTab, "integral %d: table size %d out of range [2,%d]", i,
                   tab_n[i], kMaxTab);
    DFRG_CHECK_ARG(tab_off[i] >= 0 && tab_off[i] + tab_n[i] <= tables_len, "integral %d: table out of bounds", i);
    DFRG_CHECK_ARG(bias[i] != 0.0 && stop[i] > start[i], "integral %d: invalid coordinates", i);
  }
  // integrals may share table cells (the same tab_off), then with the same size and coordinates;
  // distinct tables must not overlap (the mixed path keeps per-cell coordinate data)
  {
    std::vector<int> idx(n_int);
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](int a, int c) { return tab_off[a] < tab_off[c]; });
    for (int k = 1; k < n_int; ++k) {
      const int a = idx[k - 1], c = idx[k];
      if (tab_off[a] == tab_off[c]) {
        DFRG_CHECK_ARG(tab_n[a] == tab_n[c] && start[a] == start[c] && stop[a] == stop[c] && bias[a] == bias[c],
                       "integrals %d and %d share a table with different sizes or coordinates", a, c);
      } else {
        DFRG_CHECK_ARG(tab_off[a] + tab_n[a] <= tab_off[c], "tables of integrals %d and %d overlap", a, c);
      }
    }
  }
  dfrg_se_batch *b = new dfrg_se_batch;
  b->qp = qp;
  b->n_int = n_int;
  b->x_extent = x_extent;
  b->k = k;
  b->nq.assign(nq, nq + n_int);
  b->na.assign(na, na + n_int);
  b->p.assign(p, p + n_int);
  // quadrature nodes: one FP64 copy of each distinct order, concatenated (x then w)
  std::vector<int> orders;
  for (int i = 0; i < n_int; ++i) {
    orders.push_back(nq[i]);
    orders.push_back(na[i]);
  }
  std::sort(orders.begin(), orders.end());
  orders.erase(std::unique(orders.begin(), orders.end()), orders.end());
  std::vector<int> off_of(1025, -1);
  std::vector<double> nodes;
  for (int o : orders) {
    const void *hx = nullptr, *hw = nullptr;
    int rc = dfrg_qp_gl(qp, o, 0, nullptr, nullptr, &hx, &hw);
    if (rc != DFRG_OK) {
      delete b;
      return rc;
    }
    off_of[o] = (int)nodes.size();
    nodes.insert(nodes.end(), (const double *)hx, (const double *)hx + o);
    nodes.insert(nodes.end(), (const double *)hw, (const double *)hw + o);
  }
  b->xq_off.resize(n_int);
  b->xa_off.resize(n_int);
  b->tab_off.assign(tab_off, tab_off + n_int);
  b->tab_n.assign(tab_n, tab_n + n_int);
  b->coord.resize(3 * n_int);
  for (int i = 0; i < n_int; ++i) {
    b->xq_off[i] = off_of[nq[i]];
    b->xa_off[i] = off_of[na[i]];
    // LogarithmicCoordinates1D: b = (stop - start)/expm1(bias); idx = log1p((y-c)/b) (n-1)/bias
    const double bb = (stop[i] - start[i]) / std::expm1(bias[i]);
    b->coord[3 * i] = start[i];
    b->coord[3 * i + 1] = 1.0 / bb;
    b->coord[3 * i + 2] = (tab_n[i] - 1.0) / bias[i];
    b->evals += (unsigned long long)nq[i] * na[i] * na[i];
  }
  // the narrow mixed form, per integral (so a result does not depend on the batch around it):
  // every table node is at least 1e-30 (the nodes increase from start, and FP32 rounding keeps
  // them >= 1e-30f, the guard's value), and |64 t| <= (|p| + q_max)^2 64 log2(e)/k^2 < 2^30
  {
    const double q_max = k * std::sqrt(x_extent);
    b->narrow.resize(n_int);
    for (int i = 0; i < n_int; ++i) {
      const double pq = std::fabs(p[i]) + q_max;
      b->narrow[i] = start[i] >= 1e-30 && pq * pq * 64.0 * kLog2e / (k * k) < 1073741824.0;
    }
  }
  // cell data of the mixed path (coordinates.hh:141-196 LogarithmicCoordinates1D): node
  // y_i = b expm1(a i/(n-1)) + c, 1/(b + y_i - c), Z_i and Z_{i+1} - Z_i
  std::vector<SECellY> tabA(tables_len);
  std::vector<SECellZ> tabB(tables_len);
  for (int i = 0; i < n_int; ++i) {
    const double a = bias[i], bb = (stop[i] - start[i]) / std::expm1(a), cc = start[i];
    const int n = tab_n[i];
    for (int j = 0; j < n; ++j) {
      const double yj = bb * std::expm1(a * j * (1.0 / (n - 1.0))) + cc;
      const double v = tables[tab_off[i] + j];
      const double nv = (j + 1 < n) ? tables[tab_off[i] + j + 1] : v;
      tabA[tab_off[i] + j] = SECellY{yj * yj, (float)(1.0 / (bb + yj - cc)), (float)yj};
      tabB[tab_off[i] + j] = SECellZ{v, nv - v};
    }
  }
  auto fail = [&](int rc) {
    dfrg_se_batch_destroy(b);
    return rc;
  };
#define DFRG_HIP_B(call)                                                                                       \
  do {                                                                                                         \
    hipError_t e_ = (call);                                                                                    \
    if (e_ != hipSuccess) {                                                                                    \
      set_error("%s failed: %s", #call, hipGetErrorString(e_));                                               \
      return fail(DFRG_ERR_HIP);                                                                               \
    }                                                                                                          \
  } while (0)
  // the inputs and the ticket counters in one allocation, so a block's first loads and its
  // cross-block finish stay within a few pages
  {
    auto up = [](size_t v) { return (v + 255) / 256 * 256; };
    const size_t o_tab64 = up(sizeof(double) * nodes.size()), o_tabA = o_tab64 + up(sizeof(double) * tables_len);
    const size_t o_tabB = o_tabA + up(sizeof(SECellY) * tables_len), o_p = o_tabB + up(sizeof(SECellZ) * tables_len);
    const size_t o_cnt = o_p + up(sizeof(double) * n_int);
    const size_t o_queue = o_cnt + up(sizeof(unsigned) * dev::kCounterStride * n_int);
    DFRG_HIP_B(hipMalloc(&b->d_inputs, o_queue + up(sizeof(unsigned) * 2 * dev::kCounterStride)));
    char *base = (char *)b->d_inputs;
    b->d_nodes = (double *)base;
    b->d_tab64 = (double *)(base + o_tab64);
    b->d_tabA = (SECellY *)(base + o_tabA);
    b->d_tabB = (SECellZ *)(base + o_tabB);
    b->d_p = (double *)(base + o_p);
    b->d_counters = (unsigned *)(base + o_cnt);
    b->d_queue = (unsigned *)(base + o_queue);
  }
  DFRG_HIP_B(hipMemcpy(b->d_nodes, nodes.data(), sizeof(double) * nodes.size(), hipMemcpyHostToDevice));
  DFRG_HIP_B(hipMemcpy(b->d_tab64, tables, sizeof(double) * tables_len, hipMemcpyHostToDevice));
  DFRG_HIP_B(hipMemcpy(b->d_tabA, tabA.data(), sizeof(SECellY) * tables_len, hipMemcpyHostToDevice));
  DFRG_HIP_B(hipMemcpy(b->d_tabB, tabB.data(), sizeof(SECellZ) * tables_len, hipMemcpyHostToDevice));
  DFRG_HIP_B(hipMemcpy(b->d_p, p, sizeof(double) * n_int, hipMemcpyHostToDevice));
  DFRG_HIP_B(hipMemset(b->d_counters, 0, sizeof(unsigned) * dev::kCounterStride * n_int));
  DFRG_HIP_B(hipMemset(b->d_queue, 0, sizeof(unsigned) * 2 * dev::kCounterStride));
#undef DFRG_HIP_B
  // signature for the autotune cache
  {
    std::ostringstream os;
    unsigned long long h = 1469598103934665603ULL;
    for (int i = 0; i < n_int; ++i) {
      h = (h ^ (unsigned long long)nq[i]) * 1099511628211ULL;
      h = (h ^ (unsigned long long)na[i]) * 1099511628211ULL;
    }
    os << "n" << n_int << "_e" << b->evals << "_h" << std::hex << h;
    b->signature = os.str();
  }
  int rc = se_build_tiles(b);
  if (rc != DFRG_OK) return fail(rc);
  DFRG_HIP(hipDeviceSynchronize());
  *out = b;
  return DFRG_OK;
}

int dfrg_se_batch_destroy(dfrg_se_batch *b)
{
  if (!b) return DFRG_OK;
  hipDeviceSynchronize();
  se_drop_graph(b);
  hipFree(b->d_inputs);
  hipFree(b->d_tiles);
  hipFree(b->d_partials);
  delete b;
  return DFRG_OK;
}

int dfrg_se_batch_set_shape(dfrg_se_batch *b, int block, int tile_evals, int grid)
{
  DFRG_TRY(require_init());
  DFRG_CHECK_ARG(b != nullptr, "NULL batch");
  DFRG_CHECK_ARG(block == 64 || block == 128 || block == 256 || block == 512 || block == 1024,
                 "block %d not in {64,128,256,512,1024}", block);
  DFRG_CHECK_ARG(tile_evals >= 1, "tile_evals must be positive");
  DFRG_CHECK_ARG(grid >= 0, "grid must be >= 0");
  DFRG_HIP(hipDeviceSynchronize());
  b->block = block;
  b->tile_evals = tile_evals;
  b->grid = grid;
  se_drop_graph(b);
  return se_build_tiles(b);
}

int dfrg_se_batch_get_shape(dfrg_se_batch *b, int *block, int *tile_evals, int *grid, int *n_tiles)
{
  DFRG_CHECK_ARG(b != nullptr, "NULL batch");
  if (block) *block = b->block;
  if (tile_evals) *tile_evals = b->tile_evals;
  if (grid) *grid = b->grid;
  if (n_tiles) *n_tiles = b->n_tiles;
  return DFRG_OK;
}

int dfrg_se_batch_evals(dfrg_se_batch *b, unsigned long long *evals)
{
  DFRG_CHECK_ARG(b && evals, "NULL argument");
  *evals = b->evals;
  return DFRG_OK;
}

int dfrg_se_batch_variants(dfrg_se_batch *b, int mode, int *variant)
{
  DFRG_CHECK_ARG(b && variant, "NULL argument");
  DFRG_CHECK_ARG(mode == DFRG_FP64 || mode == DFRG_MIXED, "mode must be DFRG_FP64 or DFRG_MIXED");
  for (int i = 0; i < b->n_int; ++i)
    variant[i] = mode == DFRG_FP64 ? 0 : b->narrow[i] ? 2 : 1;
  return DFRG_OK;
}

int dfrg_se_batch_unroll(dfrg_se_batch *b, int *unroll)
{
  DFRG_CHECK_ARG(b && unroll, "NULL argument");
  *unroll = se_unroll(b);
  return DFRG_OK;
}

int dfrg_se_batch_run(dfrg_se_batch *b, int mode, double *d_out, void *stream)
{
  DFRG_TRY(require_init());
  DFRG_CHECK_ARG(b && d_out, "NULL argument");
  DFRG_CHECK_ARG(mode == DFRG_FP64 || mode == DFRG_MIXED, "invalid mode %d", mode);
  return se_launch(b, mode, d_out, resolve_stream(stream));
}

int dfrg_se_batch_run_noalloc(dfrg_se_batch *b, int mode, double *d_out, void *stream)
{
  DFRG_TRY(require_init());
  DFRG_CHECK_ARG(b && d_out, "NULL argument");
  DFRG_CHECK_ARG(stream != nullptr, "dfrg_se_batch_run_noalloc needs a stream (NULL would create one)");
  DFRG_CHECK_ARG(mode == DFRG_FP64 || mode == DFRG_MIXED, "invalid mode %d", mode);
  DFRG_CHECK_ARG(b->n_int <= DFRG_NOALLOC_MAX_INTEGRALS, "%d integrals exceed the static scratch limit of %d",
                 b->n_int, DFRG_NOALLOC_MAX_INTEGRALS);
  DFRG_CHECK_ARG(b->n_tiles <= DFRG_NOALLOC_MAX_TILES,
                 "%d tiles exceed the static scratch limit of %d (a larger tile_evals gives fewer tiles)", b->n_tiles,
                 DFRG_NOALLOC_MAX_TILES);
  char *base = (char *)noalloc_scratch();
  if (base == nullptr) {
    set_error("static scratch of this device was not resolved");
    return DFRG_ERR_RUNTIME;
  }
  // layout: queue (two lines), counters (one line per integral), partials (one line per tile)
  static_assert(2 * dev::kCounterStride * sizeof(unsigned) == 256, "queue takes the first 256 bytes");
  static_assert(dev::kCounterStride * sizeof(unsigned) == 128 && dev::kSlotStride * sizeof(double) == 128,
                "one 128-byte line per counter and per partial slot");
  SEScratch scr;
  scr.queue = (unsigned *)base;
  scr.counters = (unsigned *)(base + 256);
  scr.partials = (double *)(base + 256 + 128 * (size_t)DFRG_NOALLOC_MAX_INTEGRALS);
  hipStream_t s = (hipStream_t)stream;
  // the queue's two lines and the batch's counter lines are consecutive: clear them all
  const int lines = 2 + b->n_int;
  hipLaunchKernelGGL(clear_lines_kernel, dim3((lines + 255) / 256), dim3(256), 0, s, scr.queue, lines);
  DFRG_LAUNCHED("clear_lines_kernel");
  return se_launch(b, mode, d_out, s, &scr);
}

int dfrg_se_batch_graph_launch(dfrg_se_batch *b, int mode, double *d_out, int launches, void *stream)
{
  DFRG_TRY(require_init());
  DFRG_CHECK_ARG(b && d_out, "NULL argument");
  DFRG_CHECK_ARG(mode == DFRG_FP64 || mode == DFRG_MIXED, "invalid mode %d", mode);
  DFRG_CHECK_ARG(launches >= 1 && launches <= 4096, "launches %d out of range [1,4096]", launches);
  hipStream_t s = resolve_stream(stream);
  if (!(b->graph_exec && b->g_mode == mode && b->g_out == d_out && b->g_launches == launches &&
        b->g_block == b->block && b->g_tile == b->tile_evals && b->g_grid == b->grid)) {
    se_drop_graph(b);
    hipStream_t cs;
    DFRG_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    DFRG_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    int rc = DFRG_OK;
    for (int l = 0; l < launches && rc == DFRG_OK; ++l) rc = se_launch(b, mode, d_out, cs);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(cs, &g);
    hipStreamDestroy(cs);
    if (rc != DFRG_OK || e != hipSuccess) {
      if (g) hipGraphDestroy(g);
      if (rc != DFRG_OK) return rc;
      set_error("hipStreamEndCapture failed: %s", hipGetErrorString(e));
      return DFRG_ERR_HIP;
    }
    b->graph = g;
    DFRG_HIP(hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0));
    b->g_mode = mode;
    b->g_out = d_out;
    b->g_launches = launches;
    b->g_block = b->block;
    b->g_tile = b->tile_evals;
    b->g_grid = b->grid;
  }
  DFRG_HIP(hipGraphLaunch(b->graph_exec, s));
  return DFRG_OK;
}

int dfrg_se_batch_autotune(dfrg_se_batch *b, int mode, int reps, const char *cache_path, int *block,
                           int *tile_evals, int *grid, float *ms)
{
  DFRG_TRY(require_init());
  DFRG_CHECK_ARG(b != nullptr, "NULL batch");
  DFRG_CHECK_ARG(mode == DFRG_FP64 || mode == DFRG_MIXED, "invalid mode %d", mode);
  if (reps < 1) reps = 1;
  const std::string key = std::string(context().name) + " se m" + std::to_string(mode) + " " + b->signature;
  // cache lookup: lines "<key> block tile grid ms"
  if (cache_path) {
    std::ifstream in(cache_path);
    std::string line;
    while (std::getline(in, line)) {
      if (line.compare(0, key.size(), key) == 0 && line.size() > key.size() && line[key.size()] == ' ') {
        std::istringstream is(line.substr(key.size()));
        int bl, te, gr;
        float t;
        if (is >> bl >> te >> gr >> t) {
          DFRG_TRY(dfrg_se_batch_set_shape(b, bl, te, gr));
          if (block) *block = bl;
          if (tile_evals) *tile_evals = te;
          if (grid) *grid = gr;
          if (ms) *ms = t;
          return DFRG_OK;
        }
      }
    }
  }
  double *d_out = nullptr;
  DFRG_HIP(hipMalloc((void **)&d_out, sizeof(double) * b->n_int));
  hipStream_t s;
  DFRG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  hipEvent_t e0, e1;
  DFRG_HIP(hipEventCreate(&e0));
  DFRG_HIP(hipEventCreate(&e1));
  const int cus = context().cu_count > 0 ? context().cu_count : 256;
  const int blocks[] = {128, 256, 512, 1024};
  const int tiles[] = {1024, 2048, 4096, 8192, 16384, 32768, 65536, 131072};
  float best = 1e30f;
  bool all_ok = true;
  int bb = b->block, bt = b->tile_evals, bg = b->grid;
  // settle the clocks first: launches of the current shape until 4 ms of device time have passed
  // (at most 2048), so the first candidates are not timed on a GPU still ramping up
  {
    bool ok = hipEventRecord(e0, s) == hipSuccess;
    for (int n = 0; ok && n < 2048; n += 16) {
      for (int r = 0; r < 16 && ok; ++r) ok = se_launch(b, mode, d_out, s) == DFRG_OK;
      float t = 0.f;
      ok = ok && hipEventRecord(e1, s) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
           hipEventElapsedTime(&t, e0, e1) == hipSuccess;
      if (t >= 4.f) break;
    }
    if (!ok) hipGetLastError();
  }
  // time per launch of `n` launches at the current shape after one untimed launch; < 0 when a
  // launch or event call failed
  auto time_shape = [&](int n) -> float {
    bool ok = se_launch(b, mode, d_out, s) == DFRG_OK; // warm-up
    ok = ok && hipEventRecord(e0, s) == hipSuccess;
    for (int r = 0; r < n && ok; ++r) ok = se_launch(b, mode, d_out, s) == DFRG_OK;
    ok = ok && hipEventRecord(e1, s) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
    float t = 0.f;
    ok = ok && hipEventElapsedTime(&t, e0, e1) == hipSuccess && t > 0.f;
    if (!ok) {
      hipGetLastError(); // clear a launch error so it is not reported by a later call
      return -1.f;
    }
    return t / n;
  };
  struct Cand {
    float t;
    int bl, te, gr;
  };
  std::vector<Cand> cands;
  // first sweep: every shape, `reps` launches each; a shape counts only when every launch and event
  // call of its measurement succeeded
  for (int bl : blocks) {
    for (int te : tiles) {
      for (int gmul : {0, 2, 4, 8}) {
        const int gr = gmul == 0 ? 0 : gmul * cus * 512 / bl;
        if (dfrg_se_batch_set_shape(b, bl, te, gr) != DFRG_OK) continue;
        const float t = time_shape(reps);
        if (t < 0.f) {
          all_ok = false;
          continue;
        }
        cands.push_back({t, bl, te, gr});
      }
    }
  }
  // second sweep: the four fastest again, 4 x reps launches each, in three interleaved passes (the
  // best pass counts), so one noisy measurement does not decide between shapes a few % apart
  std::stable_sort(cands.begin(), cands.end(), [](const Cand &x, const Cand &y) { return x.t < y.t; });
  if (cands.size() > 4) cands.resize(4);
  std::vector<float> t2(cands.size(), 1e30f);
  for (int pass = 0; pass < 3; ++pass)
    for (size_t c = 0; c < cands.size(); ++c) {
      const float t = dfrg_se_batch_set_shape(b, cands[c].bl, cands[c].te, cands[c].gr) == DFRG_OK
                          ? time_shape(4 * reps)
                          : -1.f;
      if (t < 0.f) {
        all_ok = false;
        continue;
      }
      t2[c] = std::min(t2[c], t);
    }
  for (size_t c = 0; c < cands.size(); ++c)
    if (t2[c] < best) {
      best = t2[c];
      bb = cands[c].bl;
      bt = cands[c].te;
      bg = cands[c].gr;
    }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipStreamDestroy(s);
  hipFree(d_out);
  DFRG_TRY(dfrg_se_batch_set_shape(b, bb, bt, bg)); // the previous shape when no candidate succeeded
  if (best >= 1e30f) {
    set_error("autotune: no launch shape could be measured");
    return DFRG_ERR_RUNTIME;
  }
  if (cache_path && all_ok) {
    std::ofstream o(cache_path, std::ios::app);
    o << key << " " << bb << " " << bt << " " << bg << " " << best << "\n";
  }
  if (block) *block = bb;
  if (tile_evals) *tile_evals = bt;
  if (grid) *grid = bg;
  if (ms) *ms = best;
  return DFRG_OK;
}

} // extern "C"
