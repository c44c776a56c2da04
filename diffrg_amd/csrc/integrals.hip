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
    double *elw = sm;           // element * weight
    double *rdot = sm + nx;     // RBdot(q^2)
    double *s1 = sm + 2 * nx;   // q^2 + R(q^2) + m1
    double *s2 = sm + 3 * nx;   // q^2 + |p|^2 + R(q^2 + |p|^2) + m2
    double *red = sm + 4 * nx;  // block reduction
    const int i = blockIdx.x;
    const double k = A.k, k2 = k * k;
    const double T = A.T[i], p0 = A.p0[i], pa = A.pabs[i], m1 = A.m1[i], m2 = A.m2[i];
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
      elw[j] = el * (A.w[j] * A.extent);
      const double xr = q2 / k2;
      rdot[j] = 2.0 * exp(-xr) * k2 * (1.0 + xr);
      s1[j] = q2 + k2 * exp(-xr) + m1;
      const double qp2 = q2 + pa * pa;
      s2[j] = qp2 + k2 * exp(-(qp2 / k2)) + m2;
    }
    __syncthreads();
    const double *mx = A.mx + A.m_off[i];
    const double *mw = A.mw + A.m_off[i];
    const int nm = A.m_n[i];
    const int work = nx * nm;
    double acc = 0.0;
    for (int e = threadIdx.x; e < work; e += BLOCK) {
      const int j = e / nm;
      const int n = e - j * nm;
      const double wn = mx[n];
      const double a1 = s1[j], a2 = s2[j];
      const double dp = wn + p0, dm = -wn + p0;
      // denominators in FP64 in both modes; mixed replaces the FP64 division sequences by an FP32
      // hardware reciprocal refined with one FP64 Newton step
      const double d1 = wn * wn + a1, d2p = dp * dp + a2, d2m = dm * dm + a2;
      double g1, gp, gm;
      if (MODE == DFRG_FP64) {
        g1 = 1.0 / d1;
        gp = 1.0 / d2p;
        gm = 1.0 / d2m;
      } else {
        g1 = dev::rcp_seeded(d1);
        gp = dev::rcp_seeded(d2p);
        gm = dev::rcp_seeded(d2m);
      }
      double val = mw[n] * (rdot[j] * g1 * g1 * gp + rdot[j] * g1 * g1 * gm);
      if (T > 0 && n == 0) {
        double g10, g20;
        if (MODE == DFRG_FP64) {
          g10 = 1.0 / a1;
          g20 = 1.0 / (p0 * p0 + a2);
        } else {
          g10 = dev::rcp_seeded(a1);
          g20 = dev::rcp_seeded(p0 * p0 + a2);
        }
        val += T * (rdot[j] * g10 * g10 * g20);
      }
      acc += elw[j] * val;
    }
    const double tot = dev::block_sum(acc, red);
    if (threadIdx.x == 0) A.out[i] = tot;
  }

  // ============================================================================================
  // C1/C3: two-angle self-energy tile engine
  // ============================================================================================
  struct SETile {
    int integral, q0, nq, slot;
  };

  struct SECellZ {
    double z;
    float dz, pad;
  };

  struct SEArgs {
    int n_tiles;
    const SETile *tiles;
    const double *p;
    const int *nq, *na, *log2na;
    const int *xq_off, *xa_off; // offsets into nodes: x at off, w at off + n
    const double *nodes;
    const int *tab_off, *tab_n;
    const double *tab64;  // FP64 tables
    const double2 *tabA;  // mixed-path cell data: (y_i, 1/(b + y_i - c))
    const SECellZ *tabB;  // mixed-path cell data: (Z_i, Z_{i+1} - Z_i)
    const double *coord;  // per integral: c, inv_b, scale (3 doubles)
    const int *first_slot, *n_slots;
    double x_extent, k, pref; // pref = 2pi/(2pi)^4
    double *partials;
    unsigned *counters;
    unsigned *queue; // [0] next tile, [1] exited blocks (persistent mode)
    int persistent;
    double *out;
  };

  constexpr int kMaxTab = 256;
  constexpr int kMaxAng = 64;
  constexpr int kMaxQ = 128;

  template <int MODE, int BLOCK> __global__ __launch_bounds__(BLOCK) void se_kernel(SEArgs A)
  {
    __shared__ double s_tab64[MODE == DFRG_FP64 ? kMaxTab : 1];
    __shared__ double2 s_tabA[MODE == DFRG_FP64 ? 1 : kMaxTab];
    __shared__ SECellZ s_tabB[MODE == DFRG_FP64 ? 1 : kMaxTab];
    __shared__ double s_exp2[MODE == DFRG_FP64 ? 1 : 64]; // 2^(j/64), j = 0..63
    __shared__ double s_cos1[kMaxAng], s_aw1[kMaxAng], s_w2[kMaxAng];
    __shared__ double s_qt[kMaxQ], s_qs[kMaxQ], s_qc[kMaxQ];
    __shared__ double s_red[BLOCK / 64];
    __shared__ int s_tile, s_cur;

    if (threadIdx.x == 0) s_cur = -1;
    if (MODE != DFRG_FP64)
      for (int j = threadIdx.x; j < 64; j += BLOCK) s_exp2[j] = exp2(j * (1.0 / 64.0));
    const double k = A.k, k2 = k * k;
    int iter = 0;
    for (;;) {
      __syncthreads();
      if (threadIdx.x == 0) {
        if (A.persistent)
          s_tile = (int)atomicAdd(&A.queue[0], 1u);
        else
          s_tile = (iter == 0) ? (int)blockIdx.x : A.n_tiles;
      }
      __syncthreads();
      const int t = s_tile;
      ++iter;
      if (t >= A.n_tiles) break;
      const SETile tile = A.tiles[t];
      const int i = tile.integral;
      const int na = A.na[i], l2 = A.log2na[i];
      const double p = A.p[i];
      const double c = A.coord[3 * i], inv_b = A.coord[3 * i + 1], scale = A.coord[3 * i + 2];
      const int ntab = A.tab_n[i];
      const double *xa = A.nodes + A.xa_off[i];
      const double *wa = xa + na;
      if (s_cur != i) {
        // stage this integral's table and angular terms
        if (MODE == DFRG_FP64) {
          for (int j = threadIdx.x; j < ntab; j += BLOCK) s_tab64[j] = A.tab64[A.tab_off[i] + j];
        } else {
          for (int j = threadIdx.x; j < ntab; j += BLOCK) {
            s_tabA[j] = A.tabA[A.tab_off[i] + j];
            s_tabB[j] = A.tabB[A.tab_off[i] + j];
          }
        }
        for (int j = threadIdx.x; j < na; j += BLOCK) {
          const double c1 = 2.0 * (xa[j] - 0.5);
          s_cos1[j] = c1;
          s_aw1[j] = 2.0 * wa[j] * sqrt(1.0 - c1 * c1); // cos1 weight and Jacobian
          s_w2[j] = 2.0 * wa[j];                          // cos2 weight
        }
      }
      // q-only terms of this tile (FP64): el * w_q x_ext * RBdot(q) G(q)^2, p^2 + q^2, -2 p q
      {
        const double *xq = A.nodes + A.xq_off[i];
        const double *wq = xq + A.nq[i];
        for (int j = threadIdx.x; j < tile.nq; j += BLOCK) {
          const int iq = tile.q0 + j;
          const double q = k * sqrt(xq[iq] * A.x_extent);
          const double q2 = q * q;
          const double xr = q2 / k2;
          const double ex = exp(-xr);
          // read from HBM: the LDS copy of the table may still be in flight here
          const double Zq = dev::lerp_log_f64(A.tab64 + A.tab_off[i], ntab, c, inv_b, scale, q);
          const double Gq = 1.0 / (Zq * q2 + k2 * ex);
          const double el = A.pref * (q2 * 0.5 * k2);
          s_qt[j] = el * (wq[iq] * A.x_extent) * (2.0 * ex * k2 * (1.0 + xr)) * Gq * Gq;
          s_qs[j] = p * p + q2;
          s_qc[j] = -2.0 * p * q;
        }
      }
      __syncthreads();
      if (threadIdx.x == 0) s_cur = i;

      const int npairs = na * na;
      const int pstride = npairs < BLOCK ? npairs : BLOCK;
      const int qlanes = BLOCK / pstride;
      const int pr0 = threadIdx.x % pstride;
      const int qs0 = threadIdx.x / pstride;
      const int nqt = tile.nq;
      double acc = 0.0;
      if (MODE == DFRG_FP64) {
        for (int pr = pr0; pr < npairs; pr += pstride) {
          const int i1 = pr >> l2, i2 = pr & (na - 1);
          const double c1 = s_cos1[i1];
          const double aw = s_aw1[i1] * s_w2[i2];
          double sq = 0.0;
          for (int j = qs0; j < nqt; j += qlanes) {
            const double pq2 = fma(s_qc[j], c1, s_qs[j]);
            const double y = sqrt(pq2);
            const double Z = dev::lerp_log_f64(s_tab64, ntab, c, inv_b, scale, y);
            const double G = 1.0 / (Z * pq2 + k2 * exp(-(pq2 / k2)));
            sq += s_qt[j] * G;
          }
          acc += aw * sq;
        }
      } else {
        // Mixed: FP32 hardware seeds replace the FP64 library sequences; FP64 is kept where the
        // error budget needs it.  |p-q|^2 in FP64 (cancellation); sqrt from an FP32 seed and one
        // FP64 Newton step; Z cell found with the FP32 log2, the in-cell position s*log1p(u)
        // (|u| <~ 0.2) in FP32 with the rounding of 1+u restored, Z assembled in FP64 from the
        // FP64 node value; R = k^2 2^t with t split on a 1/64 grid (table) and a degree-4
        // polynomial remainder in FP64; the denominator in FP64; the reciprocal from an FP32 seed
        // refined in FP64.  Per-evaluation error ~1e-9 relative.
        const float cf = (float)c, ibf = (float)inv_b, scl = (float)(scale * kLn2), sf = (float)scale;
        const double ek64 = -64.0 * kLog2e / k2; // t * 64
        const int imax = ntab - 2;
        for (int pr = pr0; pr < npairs; pr += pstride) {
          const int i1 = pr >> l2, i2 = pr & (na - 1);
          const double c1 = s_cos1[i1];
          const double aw = s_aw1[i1] * s_w2[i2];
          double sq = 0.0;
          for (int j = qs0; j < nqt; j += qlanes) {
            const double pq2d = fma(s_qc[j], c1, s_qs[j]);
            const float ys = __builtin_amdgcn_sqrtf((float)pq2d);
            const double ysd = (double)ys;
            const double yd = fma(fma(-ysd, ysd, pq2d), (double)(0.5f * __builtin_amdgcn_rcpf(ys)), ysd);
            const float coarse = __builtin_amdgcn_logf(fmaf(ys - cf, ibf, 1.0f)) * scl;
            const int i0 = min(max((int)__builtin_floorf(coarse), 0), imax);
            const double2 ca = s_tabA[i0];
            const SECellZ cz = s_tabB[i0];
            const float u = (float)((yd - ca.x) * ca.y);
            const float w1 = 1.0f + u;
            const float ru = u - (w1 - 1.0f); // rounding error of 1 + u
            float fr = fmaf(__builtin_amdgcn_logf(w1), scl, ru * sf);
            fr = fmaxf(0.0f, fminf(fr, 1.0f)); // NaN-safe; realises the [0, n-1] clamp
            const double Z = cz.z + (double)(fr * cz.dz);
            // R = k^2 2^(t), t = -pq2/k^2 log2(e) <= 0
            const double t64 = pq2d * ek64;
            double R = 0.0;
            if (t64 > -64.0 * 1000.0) {
              const double tn = rint(t64);
              const int jn = (int)tn;
              const double r = (t64 - tn) * (kLn2 / 64.0);
              const double pe = fma(fma(fma(fma(r, 1.0 / 24.0, 1.0 / 6.0), r, 0.5), r, 1.0), r, 1.0);
              R = k2 * ldexp(pe * s_exp2[jn & 63], jn >> 6);
            }
            const double G = dev::rcp_seeded(fma(Z, pq2d, R));
            sq = fma(s_qt[j], G, sq);
          }
          acc += aw * sq;
        }
      }
      const double part = dev::block_sum(acc, s_red);
      double total = 0.0;
      if (dev::finish_group(part, A.partials, tile.slot, A.first_slot[i], A.n_slots[i], &A.counters[i], &total)) {
        if (threadIdx.x == 0) A.out[i] = total;
      }
    }
    // persistent mode: the last block out resets the queue for the next launch / replay
    if (A.persistent && threadIdx.x == 0) {
      __threadfence();
      const unsigned done = atomicAdd(&A.queue[1], 1u);
      if (done == gridDim.x - 1) {
        atomicExch(&A.queue[0], 0u);
        atomicExch(&A.queue[1], 0u);
      }
    }
  }

  template <int MODE> hipError_t launch_se(int block, int grid, size_t, hipStream_t s, const SEArgs &a)
  {
    switch (block) {
    case 64: hipLaunchKernelGGL((se_kernel<MODE, 64>), dim3(grid), dim3(64), 0, s, a); break;
    case 128: hipLaunchKernelGGL((se_kernel<MODE, 128>), dim3(grid), dim3(128), 0, s, a); break;
    case 256: hipLaunchKernelGGL((se_kernel<MODE, 256>), dim3(grid), dim3(256), 0, s, a); break;
    case 512: hipLaunchKernelGGL((se_kernel<MODE, 512>), dim3(grid), dim3(512), 0, s, a); break;
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
} // namespace

// ----------------------------------------------------------------------------------------------
// self-energy batch object
// ----------------------------------------------------------------------------------------------
struct dfrg_se_batch {
  dfrg_qp *qp = nullptr;
  int n_int = 0;
  double x_extent = 20.0, k = 1.0;
  std::vector<int> nq, na;
  std::vector<double> p;
  unsigned long long evals = 0;
  // device data
  double *d_p = nullptr, *d_nodes = nullptr, *d_tab64 = nullptr, *d_coord = nullptr;
  double2 *d_tabA = nullptr;
  SECellZ *d_tabB = nullptr;
  int *d_ints = nullptr; // nq, na, log2na, xq_off, xa_off, tab_off, tab_n, first_slot, n_slots
  SETile *d_tiles = nullptr;
  double *d_partials = nullptr;
  unsigned *d_counters = nullptr, *d_queue = nullptr;
  size_t tiles_cap = 0;
  // shape
  int block = 256, tile_evals = 8192, grid = 0, n_tiles = 0;
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
    std::vector<int> first(n), nslots(n);
    int slot = 0;
    for (int i : order) {
      const int per_q = b->na[i] * b->na[i];
      int qpt = std::max(1, std::min(b->nq[i], b->tile_evals / per_q));
      qpt = std::min(qpt, kMaxQ);
      first[i] = slot;
      int cnt = 0;
      for (int q0 = 0; q0 < b->nq[i]; q0 += qpt) {
        tiles.push_back({i, q0, std::min(qpt, b->nq[i] - q0), slot++});
        ++cnt;
      }
      nslots[i] = cnt;
    }
    b->n_tiles = (int)tiles.size();
    if ((size_t)b->n_tiles > b->tiles_cap) {
      hipFree(b->d_tiles);
      hipFree(b->d_partials);
      b->tiles_cap = b->n_tiles;
      DFRG_HIP(hipMalloc((void **)&b->d_tiles, sizeof(SETile) * b->tiles_cap));
      DFRG_HIP(hipMalloc((void **)&b->d_partials, sizeof(double) * b->tiles_cap));
    }
    DFRG_HIP(hipMemcpy(b->d_tiles, tiles.data(), sizeof(SETile) * tiles.size(), hipMemcpyHostToDevice));
    DFRG_HIP(hipMemcpy(b->d_ints + 7 * n, first.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    DFRG_HIP(hipMemcpy(b->d_ints + 8 * n, nslots.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    return DFRG_OK;
  }

  SEArgs se_args(dfrg_se_batch *b, double *d_out)
  {
    const int n = b->n_int;
    SEArgs a;
    a.n_tiles = b->n_tiles;
    a.tiles = b->d_tiles;
    a.p = b->d_p;
    a.nq = b->d_ints;
    a.na = b->d_ints + n;
    a.log2na = b->d_ints + 2 * n;
    a.xq_off = b->d_ints + 3 * n;
    a.xa_off = b->d_ints + 4 * n;
    a.tab_off = b->d_ints + 5 * n;
    a.tab_n = b->d_ints + 6 * n;
    a.first_slot = b->d_ints + 7 * n;
    a.n_slots = b->d_ints + 8 * n;
    a.nodes = b->d_nodes;
    a.tab64 = b->d_tab64;
    a.tabA = b->d_tabA;
    a.tabB = b->d_tabB;
    a.coord = b->d_coord;
    a.x_extent = b->x_extent;
    a.k = b->k;
    a.pref = 2.0 * kPi / ((2.0 * kPi) * ((2.0 * kPi) * ((2.0 * kPi) * (2.0 * kPi))));
    a.partials = b->d_partials;
    a.counters = b->d_counters;
    a.queue = b->d_queue;
    a.persistent = b->grid > 0 ? 1 : 0;
    a.out = d_out;
    return a;
  }

  int se_launch(dfrg_se_batch *b, int mode, double *d_out, hipStream_t s)
  {
    const SEArgs a = se_args(b, d_out);
    const int grid = b->grid > 0 ? std::min(b->grid, b->n_tiles) : b->n_tiles;
    hipError_t e = mode == DFRG_FP64 ? launch_se<DFRG_FP64>(b->block, grid, 0, s, a)
                                     : launch_se<DFRG_MIXED>(b->block, grid, 0, s, a);
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
  if (n_int == 0) return DFRG_OK;
  FTArgs a{d, radial_mode, n_int, nx, d_T, d_p0, d_pabs, d_m1, d_m2, d_mx, d_mw, d_m_off, d_m_n, d_x, d_w,
           extent, k, 0.0, d_out};
  a.pref = S_d_host(d - 1) / std::pow(2.0 * kPi, d - 1);
  constexpr int B = 256;
  const size_t shm = sizeof(double) * (4 * nx + B / 64);
  hipStream_t s = resolve_stream(stream);
  if (mode == DFRG_FP64)
    hipLaunchKernelGGL((ft_kernel<DFRG_FP64, B>), dim3(n_int), dim3(B), shm, s, a);
  else
    hipLaunchKernelGGL((ft_kernel<DFRG_MIXED, B>), dim3(n_int), dim3(B), shm, s, a);
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
  std::vector<int> ints(9 * n_int, 0);
  std::vector<double> coord(3 * n_int);
  for (int i = 0; i < n_int; ++i) {
    ints[i] = nq[i];
    ints[n_int + i] = na[i];
    ints[2 * n_int + i] = ilog2(na[i]);
    ints[3 * n_int + i] = off_of[nq[i]];
    ints[4 * n_int + i] = off_of[na[i]];
    ints[5 * n_int + i] = tab_off[i];
    ints[6 * n_int + i] = tab_n[i];
    // LogarithmicCoordinates1D: b = (stop - start)/expm1(bias); idx = log1p((y-c)/b) (n-1)/bias
    const double bb = (stop[i] - start[i]) / std::expm1(bias[i]);
    coord[3 * i] = start[i];
    coord[3 * i + 1] = 1.0 / bb;
    coord[3 * i + 2] = (tab_n[i] - 1.0) / bias[i];
    b->evals += (unsigned long long)nq[i] * na[i] * na[i];
  }
  // cell data of the mixed path (coordinates.hh:141-196 LogarithmicCoordinates1D): node
  // y_i = b expm1(a i/(n-1)) + c, 1/(b + y_i - c), Z_i and Z_{i+1} - Z_i
  std::vector<double2> tabA(tables_len);
  std::vector<SECellZ> tabB(tables_len);
  for (int i = 0; i < n_int; ++i) {
    const double a = bias[i], bb = (stop[i] - start[i]) / std::expm1(a), cc = start[i];
    const int n = tab_n[i];
    for (int j = 0; j < n; ++j) {
      const double yj = bb * std::expm1(a * j * (1.0 / (n - 1.0))) + cc;
      const double v = tables[tab_off[i] + j];
      const double nv = (j + 1 < n) ? tables[tab_off[i] + j + 1] : v;
      tabA[tab_off[i] + j] = make_double2(yj, 1.0 / (bb + yj - cc));
      tabB[tab_off[i] + j] = SECellZ{v, (float)(nv - v), 0.0f};
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
  DFRG_HIP_B(hipMalloc((void **)&b->d_p, sizeof(double) * n_int));
  DFRG_HIP_B(hipMalloc((void **)&b->d_nodes, sizeof(double) * nodes.size()));
  DFRG_HIP_B(hipMalloc((void **)&b->d_tab64, sizeof(double) * tables_len));
  DFRG_HIP_B(hipMalloc((void **)&b->d_tabA, sizeof(double2) * tables_len));
  DFRG_HIP_B(hipMalloc((void **)&b->d_tabB, sizeof(SECellZ) * tables_len));
  DFRG_HIP_B(hipMalloc((void **)&b->d_coord, sizeof(double) * 3 * n_int));
  DFRG_HIP_B(hipMalloc((void **)&b->d_ints, sizeof(int) * 9 * n_int));
  DFRG_HIP_B(hipMalloc((void **)&b->d_counters, sizeof(unsigned) * n_int));
  DFRG_HIP_B(hipMalloc((void **)&b->d_queue, sizeof(unsigned) * 2));
  DFRG_HIP_B(hipMemcpy(b->d_p, p, sizeof(double) * n_int, hipMemcpyHostToDevice));
  DFRG_HIP_B(hipMemcpy(b->d_nodes, nodes.data(), sizeof(double) * nodes.size(), hipMemcpyHostToDevice));
  DFRG_HIP_B(hipMemcpy(b->d_tab64, tables, sizeof(double) * tables_len, hipMemcpyHostToDevice));
  DFRG_HIP_B(hipMemcpy(b->d_tabA, tabA.data(), sizeof(double2) * tables_len, hipMemcpyHostToDevice));
  DFRG_HIP_B(hipMemcpy(b->d_tabB, tabB.data(), sizeof(SECellZ) * tables_len, hipMemcpyHostToDevice));
  DFRG_HIP_B(hipMemcpy(b->d_coord, coord.data(), sizeof(double) * 3 * n_int, hipMemcpyHostToDevice));
  DFRG_HIP_B(hipMemcpy(b->d_ints, ints.data(), sizeof(int) * 9 * n_int, hipMemcpyHostToDevice));
  DFRG_HIP_B(hipMemset(b->d_counters, 0, sizeof(unsigned) * n_int));
  DFRG_HIP_B(hipMemset(b->d_queue, 0, sizeof(unsigned) * 2));
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
  hipFree(b->d_p);
  hipFree(b->d_nodes);
  hipFree(b->d_tab64);
  hipFree(b->d_tabA);
  hipFree(b->d_tabB);
  hipFree(b->d_coord);
  hipFree(b->d_ints);
  hipFree(b->d_tiles);
  hipFree(b->d_partials);
  hipFree(b->d_counters);
  hipFree(b->d_queue);
  delete b;
  return DFRG_OK;
}

int dfrg_se_batch_set_shape(dfrg_se_batch *b, int block, int tile_evals, int grid)
{
  DFRG_TRY(require_init());
  DFRG_CHECK_ARG(b != nullptr, "NULL batch");
  DFRG_CHECK_ARG(block == 64 || block == 128 || block == 256 || block == 512, "block %d not in {64,128,256,512}",
                 block);
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

int dfrg_se_batch_run(dfrg_se_batch *b, int mode, double *d_out, void *stream)
{
  DFRG_TRY(require_init());
  DFRG_CHECK_ARG(b && d_out, "NULL argument");
  DFRG_CHECK_ARG(mode == DFRG_FP64 || mode == DFRG_MIXED, "invalid mode %d", mode);
  return se_launch(b, mode, d_out, resolve_stream(stream));
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
    if (rc != DFRG_OK) return rc;
    if (e != hipSuccess) {
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
  const int blocks[] = {128, 256, 512};
  const int tiles[] = {1024, 2048, 4096, 8192, 16384, 32768};
  float best = 1e30f;
  int bb = b->block, bt = b->tile_evals, bg = b->grid;
  for (int bl : blocks) {
    for (int te : tiles) {
      for (int gmul : {0, 2, 4, 8}) {
        const int gr = gmul == 0 ? 0 : gmul * cus * (512 / bl);
        if (dfrg_se_batch_set_shape(b, bl, te, gr) != DFRG_OK) continue;
        se_launch(b, mode, d_out, s); // warm-up
        hipEventRecord(e0, s);
        for (int r = 0; r < reps; ++r) se_launch(b, mode, d_out, s);
        hipEventRecord(e1, s);
        hipEventSynchronize(e1);
        float t = 0.f;
        hipEventElapsedTime(&t, e0, e1);
        t /= reps;
        if (t < best) {
          best = t;
          bb = bl;
          bt = te;
          bg = gr;
        }
      }
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipStreamDestroy(s);
  hipFree(d_out);
  DFRG_TRY(dfrg_se_batch_set_shape(b, bb, bt, bg));
  if (cache_path) {
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
