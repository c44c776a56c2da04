// quadrature.hip -- FP64 quadrature rules generated on the GPU, and the provider cache.
//
// Gauss-Legendre: one lane per root, Newton iteration on the Legendre three-term recurrence in
// double-double arithmetic, so that nodes near 0 and 1 keep full relative precision after the
// map t = (1 -+ z)/2.  Replaces GSL's gsl_integration_fixed_legendre (quadrature.cu:23-58).
//
// Matsubara (Monien, Math. Comp. 79, 857 (2010); matsubara.cu:33-85): the Jacobi matrix of the
// recurrence a_j, b_j is built on the device; lane i finds the i-th smallest eigenvalue by Sturm
// bisection and the first component of its eigenvector from a twisted factorisation
// (Parlett & Dhillon's twisted LDL^T / UDU^T, the core of MRRR).  The reference diagonalises the
// same matrix with the sequential implicit-QL routine IMTQLX (diagonalization.hh:34-145); the
// parallel form is at least as accurate on this graded matrix.
#include "internal.hh"
#include "dfrg/device.hh"

// The double-double arithmetic below relies on error-free transformations; fusing a product
// into a later sum (FP contraction, on by default for device code) would break them.
#pragma clang fp contract(off)

#include <cmath>
#include <map>
#include <vector>

namespace dfrg
{
  namespace
  {
    // ---------------------------------------------------------------------------------------
    // double-double arithmetic (Dekker / Knuth error-free transformations)
    // ---------------------------------------------------------------------------------------
    struct dd {
      double hi, lo;
    };
    __device__ __forceinline__ dd quick_two_sum(double a, double b)
    {
      const double s = a + b;
      return {s, b - (s - a)};
    }
    __device__ __forceinline__ dd two_sum(double a, double b)
    {
      const double s = a + b;
      const double bb = s - a;
      return {s, (a - (s - bb)) + (b - bb)};
    }
    __device__ __forceinline__ dd dd_add(dd a, dd b)
    {
      dd s = two_sum(a.hi, b.hi);
      const dd t = two_sum(a.lo, b.lo);
      s.lo += t.hi;
      s = quick_two_sum(s.hi, s.lo);
      s.lo += t.lo;
      return quick_two_sum(s.hi, s.lo);
    }
    __device__ __forceinline__ dd dd_neg(dd a) { return {-a.hi, -a.lo}; }
    __device__ __forceinline__ dd dd_sub(dd a, dd b) { return dd_add(a, dd_neg(b)); }
    __device__ __forceinline__ dd dd_mul(dd a, dd b)
    {
      const double p = a.hi * b.hi;
      double e = fma(a.hi, b.hi, -p);
      e += a.hi * b.lo + a.lo * b.hi;
      return quick_two_sum(p, e);
    }
    __device__ __forceinline__ dd dd_mul_d(dd a, double b)
    {
      const double p = a.hi * b;
      double e = fma(a.hi, b, -p);
      e += a.lo * b;
      return quick_two_sum(p, e);
    }
    __device__ __forceinline__ dd dd_div(dd a, dd b)
    {
      const double q1 = a.hi / b.hi;
      dd r = dd_sub(a, dd_mul_d(b, q1));
      const double q2 = r.hi / b.hi;
      r = dd_sub(r, dd_mul_d(b, q2));
      const double q3 = r.hi / b.hi;
      return dd_add(quick_two_sum(q1, q2), dd{q3, 0.0});
    }
    __device__ __forceinline__ dd dd_from(double a) { return {a, 0.0}; }

    // P_n(z) and P_n'(z) by the three-term recurrence, in double-double.
    __device__ void legendre_dd(int n, dd z, dd *pn, dd *dpn)
    {
      dd p0 = dd_from(1.0), p1 = z;
      for (int j = 2; j <= n; ++j) {
        const dd t = dd_sub(dd_mul(dd_mul_d(z, 2.0 * j - 1.0), p1), dd_mul_d(p0, j - 1.0));
        const dd p2 = dd_div(t, dd_from((double)j));
        p0 = p1;
        p1 = p2;
      }
      *pn = p1;
      // n (z P_n - P_{n-1}) / (z^2 - 1)
      *dpn = dd_div(dd_mul_d(dd_sub(dd_mul(z, p1), p0), (double)n), dd_sub(dd_mul(z, z), dd_from(1.0)));
    }

    __global__ void gl_kernel(int n, double *x, double *w)
    {
      const int i = blockIdx.x * blockDim.x + threadIdx.x;
      if (i >= (n + 1) / 2) return;
      // i-th root from the top; Tricomi-type initial guess
      dd z = dd_from(cos(dev::pi * (i + 0.75) / (n + 0.5)));
      dd pn, dpn;
      for (int it = 0; it < 12; ++it) {
        legendre_dd(n, z, &pn, &dpn);
        const dd dz = dd_div(pn, dpn);
        z = dd_sub(z, dz);
        if (fabs(dz.hi) <= 1e-31 * fmax(fabs(z.hi), 1e-3)) break;
      }
      legendre_dd(n, z, &pn, &dpn);
      const dd one = dd_from(1.0);
      // weight on [0,1]: 1 / ((1 - z^2) P_n'(z)^2)
      const dd wz = dd_div(one, dd_mul(dd_sub(one, dd_mul(z, z)), dd_mul(dpn, dpn)));
      const dd xhi = dd_mul_d(dd_add(one, z), 0.5);
      const dd xlo = dd_mul_d(dd_sub(one, z), 0.5);
      x[n - 1 - i] = xhi.hi;
      x[i] = xlo.hi;
      w[n - 1 - i] = wz.hi;
      w[i] = wz.hi;
      if ((n & 1) && i == n / 2) x[i] = 0.5;
    }

    // ---------------------------------------------------------------------------------------
    // Monien rule
    // ---------------------------------------------------------------------------------------
    __device__ __forceinline__ int sturm_count_below(const double *a, const double *e2, int n, double x)
    {
      int c = 0;
      double q = a[0] - x;
      c += q < 0.0;
      for (int i = 1; i < n; ++i) {
        if (q == 0.0) q = 1e-300;
        q = (a[i] - x) - e2[i - 1] / q;
        c += q < 0.0;
      }
      return c;
    }

    // One block; n <= 1024 lanes; scratch holds 2 n^2 doubles (per-lane twisted factorisation).
    __global__ void matsubara_kernel(int n, double T, double *x_out, double *w_out, double *scratch)
    {
      extern __shared__ double sm[];
      double *a = sm;          // diagonal
      double *e = sm + n;      // off-diagonal, e[j] couples j and j+1
      double *e2 = sm + 2 * n; // e^2
      const double pi = dev::pi;
      for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const double j1 = j + 1;
        double aj = 2.0 * (pi * pi) / (4 * j + 1) / (4 * j + 5);
        if (j == 0) aj = (pi * pi) / 15.0;
        a[j] = aj;
        const double bj = (pi * (pi * (pi * pi))) / ((4 * j1 - 1) * (4 * j1 + 3)) / ((4 * j1 + 1) * (4 * j1 + 1));
        const double ej = (j < n - 1) ? sqrt(bj) : 0.0;
        e[j] = ej;
      }
      __syncthreads();
      for (int j = threadIdx.x; j < n; j += blockDim.x) e2[j] = e[j] * e[j];
      __syncthreads();

      const int k = threadIdx.x;
      if (k >= n) return;
      // Gershgorin interval
      double lo = 1e300, hi = -1e300;
      for (int j = 0; j < n; ++j) {
        const double r = (j > 0 ? fabs(e[j - 1]) : 0.0) + (j < n - 1 ? fabs(e[j]) : 0.0);
        lo = fmin(lo, a[j] - r);
        hi = fmax(hi, a[j] + r);
      }
      // k-th smallest eigenvalue: count(x) > k  <=>  lambda_k < x
      for (int it = 0; it < 256; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        if (sturm_count_below(a, e2, n, mid) > k)
          hi = mid;
        else
          lo = mid;
      }
      const double lam = 0.5 * (lo + hi);

      // twisted factorisation at lam: top-down pivots d, bottom-up pivots dp
      double *d = scratch + (size_t)k * 2 * n;
      double *dp = d + n;
      d[0] = a[0] - lam;
      for (int j = 1; j < n; ++j) {
        const double prev = d[j - 1] != 0.0 ? d[j - 1] : 1e-300;
        d[j] = (a[j] - lam) - e2[j - 1] / prev;
      }
      dp[n - 1] = a[n - 1] - lam;
      for (int j = n - 2; j >= 0; --j) {
        const double next = dp[j + 1] != 0.0 ? dp[j + 1] : 1e-300;
        dp[j] = (a[j] - lam) - e2[j] / next;
      }
      int r = 0;
      double gmin = 1e300;
      for (int j = 0; j < n; ++j) {
        const double g = fabs(d[j] + dp[j] - (a[j] - lam));
        if (g < gmin) {
          gmin = g;
          r = j;
        }
      }
      // eigenvector with z_r = 1; components outward from the twist index
      double z = 1.0, s = 1.0, z0 = (r == 0) ? 1.0 : 0.0;
      for (int j = r - 1; j >= 0; --j) {
        z = -(e[j] / (d[j] != 0.0 ? d[j] : 1e-300)) * z;
        s += z * z;
        if (j == 0) z0 = z;
      }
      z = 1.0;
      for (int j = r + 1; j < n; ++j) {
        z = -(e[j - 1] / (dp[j] != 0.0 ? dp[j] : 1e-300)) * z;
        s += z * z;
      }
      const double mu0 = (pi * pi) / 6.0;
      const double wk = mu0 * (z0 * z0 / s);
      // normalise the weights and scale the nodes (matsubara.cu:78-84)
      w_out[k] = T * wk / lam;
      x_out[k] = 2.0 * pi * T / sqrt(lam);
    }

    // T = 0 rule (matsubara.cu:128-161): GL(h) on [0, 2E] and log-mapped GL(h) on [2E, 1e6 E].
    __global__ void matsubara_vacuum_kernel(int h, double E, const double *gx, const double *gw, double *x,
                                            double *w)
    {
      const int i = blockIdx.x * blockDim.x + threadIdx.x;
      if (i >= h) return;
      const double two_pi = 2.0 * dev::pi;
      const double div = 2.0 * fabs(E);
      x[i] = gx[i] * div;
      w[i] = gw[i] * div / two_pi;
      const double log_start = log(div);
      const double log_ext = log(1e6 * fabs(E) / div);
      const double xl = exp(log_start + log_ext * gx[i]);
      x[i + h] = xl;
      w[i + h] = (gw[i] * log_ext * xl) / two_pi;
    }

    __global__ void to_f32_kernel(int n, const double *a, float *b)
    {
      const int i = blockIdx.x * blockDim.x + threadIdx.x;
      if (i < n) b[i] = (float)a[i];
    }

    bool is_close_abs(double a, double b, double eps)
    {
      const double diff = std::fabs(a - b);
      return diff <= eps || diff <= std::fmax(std::fabs(a), std::fabs(b)) * eps;
    }

    // MatsubaraQuadrature::predict_size (matsubara.cu:12-21)
    int predict_size(double T, double E, int step, int vacuum_quad_size, int precision_factor)
    {
      const double rel = std::fabs(E) / std::fabs(T + 1e-16);
      if (is_close_abs(T, 0.0, 2.220446049250313e-16) || rel > 1e+2) return -vacuum_quad_size;
      const double E_max = (precision_factor + 100) * std::fabs(E);
      int size = 5 + (int)std::sqrt(4.0 * E_max / (M_PI * M_PI * std::fabs(T)));
      size = (int)std::ceil(size / (double)step) * step;
      return size;
    }
  } // namespace
} // namespace dfrg

using namespace dfrg;

extern "C" {

int dfrg_gl_generate(int order, double *d_x, double *d_w, void *stream)
{
  DFRG_TRY(require_init());
  DFRG_CHECK_ARG(order >= 1 && order <= 4096, "Gauss-Legendre order %d out of range [1,4096]", order);
  DFRG_CHECK_ARG(d_x && d_w, "output pointers are NULL");
  const int roots = (order + 1) / 2;
  const int block = 64;
  hipLaunchKernelGGL(gl_kernel, dim3((roots + block - 1) / block), dim3(block), 0, resolve_stream(stream), order,
                     d_x, d_w);
  DFRG_LAUNCHED("gl_kernel");
  return DFRG_OK;
}

int dfrg_matsubara_generate(double T, double E, int step, int min_size, int max_size, int vacuum_quad_size,
                            int precision_factor, double *d_x, double *d_w, int cap, int *size, double *T_rule,
                            void *stream)
{
  DFRG_TRY(require_init());
  DFRG_CHECK_ARG(d_x && d_w && size, "output pointers are NULL");
  DFRG_CHECK_ARG(step >= 1, "step must be positive");
  DFRG_CHECK_ARG(std::isfinite(T) && std::isfinite(E) && T >= 0.0, "invalid T=%g / E=%g", T, E);
  if (precision_factor < 0) precision_factor = 1;
  if (vacuum_quad_size <= 6) vacuum_quad_size = 6;
  hipStream_t s = resolve_stream(stream);
  int m = predict_size(T, E, step, vacuum_quad_size, precision_factor);
  if (m < 0) {
    m = -m;
    if (m % 2) ++m;
    DFRG_CHECK_ARG(m <= cap, "vacuum Matsubara rule of %d nodes exceeds capacity %d", m, cap);
    if (is_close_abs(E, 0.0, 2.220446049250313e-16)) E = 1.0;
    const int h = m / 2;
    double *g = nullptr;
    DFRG_HIP(hipMallocAsync((void **)&g, sizeof(double) * 2 * h, s));
    hipLaunchKernelGGL(gl_kernel, dim3((h + 1) / 2 / 64 + 1), dim3(64), 0, s, h, g, g + h);
    DFRG_LAUNCHED("gl_kernel");
    hipLaunchKernelGGL(matsubara_vacuum_kernel, dim3(h / 64 + 1), dim3(64), 0, s, h, E, g, g + h, d_x, d_w);
    DFRG_LAUNCHED("matsubara_vacuum_kernel");
    DFRG_HIP(hipFreeAsync(g, s));
    DFRG_HIP(hipStreamSynchronize(s));
    *size = m;
    if (T_rule) *T_rule = 0.0;
    return DFRG_OK;
  }
  m = std::max(min_size, std::min(max_size, m));
  DFRG_CHECK_ARG(m >= 1 && m <= 1024, "Matsubara rule size %d out of range [1,1024]", m);
  DFRG_CHECK_ARG(m <= cap, "Matsubara rule of %d nodes exceeds capacity %d", m, cap);
  double *scratch = nullptr;
  DFRG_HIP(hipMallocAsync((void **)&scratch, sizeof(double) * 2 * (size_t)m * m, s));
  const int block = ((m + 63) / 64) * 64;
  hipLaunchKernelGGL(matsubara_kernel, dim3(1), dim3(block), 3 * m * sizeof(double), s, m, T, d_x, d_w, scratch);
  DFRG_LAUNCHED("matsubara_kernel");
  DFRG_HIP(hipFreeAsync(scratch, s));
  DFRG_HIP(hipStreamSynchronize(s));
  *size = m;
  if (T_rule) *T_rule = T;
  return DFRG_OK;
}

} // extern "C"

// ---------------------------------------------------------------------------------------------
// provider cache (QuadratureProvider, quadrature_provider.hh:138-248, .cu:12-160)
// ---------------------------------------------------------------------------------------------
struct Rule {
  int size = 0;
  double T_rule = 0.0;
  double *d_x = nullptr, *d_w = nullptr;
  float *d_xf = nullptr, *d_wf = nullptr;
  std::vector<double> hx, hw;
  std::vector<float> hxf, hwf;
};

struct dfrg_qp {
  std::mutex mutex;
  int vacuum_quad_size = 48;
  int min_matsubara_size = 32;
  int add_matsubara_size = 0;
  std::map<int, Rule> gl;
  std::map<double, std::map<double, Rule>> mats;
};

namespace
{
  int finish_rule(Rule &r, hipStream_t s)
  {
    const int n = r.size;
    DFRG_HIP(hipMalloc((void **)&r.d_xf, sizeof(float) * n));
    DFRG_HIP(hipMalloc((void **)&r.d_wf, sizeof(float) * n));
    hipLaunchKernelGGL(to_f32_kernel, dim3(n / 64 + 1), dim3(64), 0, s, n, r.d_x, r.d_xf);
    hipLaunchKernelGGL(to_f32_kernel, dim3(n / 64 + 1), dim3(64), 0, s, n, r.d_w, r.d_wf);
    DFRG_LAUNCHED("to_f32_kernel");
    r.hx.resize(n);
    r.hw.resize(n);
    DFRG_HIP(hipMemcpyAsync(r.hx.data(), r.d_x, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    DFRG_HIP(hipMemcpyAsync(r.hw.data(), r.d_w, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    DFRG_HIP(hipStreamSynchronize(s));
    r.hxf.assign(r.hx.begin(), r.hx.end());
    r.hwf.assign(r.hw.begin(), r.hw.end());
    return DFRG_OK;
  }

  void free_rule(Rule &r)
  {
    hipFree(r.d_x);
    hipFree(r.d_w);
    hipFree(r.d_xf);
    hipFree(r.d_wf);
  }

  void hand_out(const Rule &r, int fp32, const void **d_x, const void **d_w, const void **h_x, const void **h_w)
  {
    if (d_x) *d_x = fp32 ? (const void *)r.d_xf : (const void *)r.d_x;
    if (d_w) *d_w = fp32 ? (const void *)r.d_wf : (const void *)r.d_w;
    if (h_x) *h_x = fp32 ? (const void *)r.hxf.data() : (const void *)r.hx.data();
    if (h_w) *h_w = fp32 ? (const void *)r.hwf.data() : (const void *)r.hw.data();
  }

  // find_T / find_E (quadrature_provider.cu:31-60): lower_bound, then the entry itself or its
  // predecessor when within 1e-6 * key.
  template <typename Map> typename Map::iterator find_close(Map &m, double key)
  {
    auto it = m.lower_bound(key);
    if (it != m.end() && is_close_abs(it->first, key, 1e-6 * key)) return it;
    if (it != m.begin()) {
      auto p = std::prev(it);
      if (is_close_abs(p->first, key, 1e-6 * key)) return p;
    }
    return m.end();
  }
} // namespace

extern "C" {

int dfrg_qp_create(dfrg_qp **qp, int vacuum_quad_size, int min_matsubara_size, int add_matsubara_size)
{
  DFRG_TRY(require_init());
  DFRG_CHECK_ARG(qp != nullptr, "qp is NULL");
  DFRG_CHECK_ARG(vacuum_quad_size > 0, "MatsubaraStorage: Vacuum quadrature size must be positive.");
  DFRG_CHECK_ARG(add_matsubara_size >= 0, "MatsubaraStorage: add_matsubara_size must be positive.");
  DFRG_CHECK_ARG(min_matsubara_size > 0, "MatsubaraStorage: min_matsubara_size must be positive.");
  dfrg_qp *p = new dfrg_qp;
  p->vacuum_quad_size = vacuum_quad_size;
  p->min_matsubara_size = min_matsubara_size;
  p->add_matsubara_size = add_matsubara_size;
  *qp = p;
  return DFRG_OK;
}

int dfrg_qp_destroy(dfrg_qp *qp)
{
  if (!qp) return DFRG_OK;
  DFRG_TRY(require_init());
  hipDeviceSynchronize();
  for (auto &kv : qp->gl) free_rule(kv.second);
  for (auto &t : qp->mats)
    for (auto &kv : t.second) free_rule(kv.second);
  delete qp;
  return DFRG_OK;
}

int dfrg_qp_gl(dfrg_qp *qp, int order, int fp32, const void **d_x, const void **d_w, const void **h_x,
               const void **h_w)
{
  DFRG_TRY(require_init());
  DFRG_CHECK_ARG(qp != nullptr, "qp is NULL");
  DFRG_CHECK_ARG(order >= 1 && order <= 4096, "Gauss-Legendre order %d out of range [1,4096]", order);
  std::lock_guard<std::mutex> lock(qp->mutex);
  auto it = qp->gl.find(order);
  if (it == qp->gl.end()) {
    Rule r;
    r.size = order;
    DFRG_HIP(hipMalloc((void **)&r.d_x, sizeof(double) * order));
    DFRG_HIP(hipMalloc((void **)&r.d_w, sizeof(double) * order));
    hipStream_t s = resolve_stream(nullptr);
    DFRG_TRY(dfrg_gl_generate(order, r.d_x, r.d_w, s));
    DFRG_TRY(finish_rule(r, s));
    it = qp->gl.emplace(order, std::move(r)).first;
  }
  hand_out(it->second, fp32, d_x, d_w, h_x, h_w);
  return DFRG_OK;
}

int dfrg_qp_matsubara(dfrg_qp *qp, double T, double E, int fp32, int *size, double *T_rule, const void **d_x,
                      const void **d_w, const void **h_x, const void **h_w)
{
  DFRG_TRY(require_init());
  DFRG_CHECK_ARG(qp != nullptr, "qp is NULL");
  DFRG_CHECK_ARG(std::isfinite(T) && T >= 0.0, "invalid temperature %g", T);
  std::lock_guard<std::mutex> lock(qp->mutex);
  auto tit = find_close(qp->mats, T);
  if (tit == qp->mats.end()) tit = qp->mats.emplace(T, std::map<double, Rule>()).first;
  auto &emap = tit->second;
  auto eit = find_close(emap, E);
  if (eit == emap.end()) {
    Rule r;
    const int cap = std::max(qp->vacuum_quad_size + 2, 130);
    DFRG_HIP(hipMalloc((void **)&r.d_x, sizeof(double) * cap));
    DFRG_HIP(hipMalloc((void **)&r.d_w, sizeof(double) * cap));
    hipStream_t s = resolve_stream(nullptr);
    // MatsubaraStorage::find_E_d: reinit(T_key, E, 2, min_matsubara_size, 128, vacuum_quad_size,
    // add_matsubara_size) (quadrature_provider.cu:52)
    DFRG_TRY(dfrg_matsubara_generate(tit->first, E, 2, qp->min_matsubara_size, 128, qp->vacuum_quad_size,
                                     qp->add_matsubara_size, r.d_x, r.d_w, cap, &r.size, &r.T_rule, s));
    DFRG_TRY(finish_rule(r, s));
    eit = emap.emplace(E, std::move(r)).first;
  }
  if (size) *size = eit->second.size;
  if (T_rule) *T_rule = eit->second.T_rule;
  hand_out(eit->second, fp32, d_x, d_w, h_x, h_w);
  return DFRG_OK;
}

} // extern "C"
