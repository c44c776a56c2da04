/*
 * oracle.c -- CPU restatement of DiFfRG v1.0.1's momentum-loop integrators.
 *
 * TEST INFRASTRUCTURE ONLY (reference values of the tests and the CPU baseline); see oracle.h.
 * Integrator loops follow the reference's TBB forms (element, weight and kernel
 * per node, plain FP64 sums).  Summation is made deterministic: per outer node
 * partial sums are added in index order, independent of the OpenMP thread count.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

int orc_num_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------------------------------ */
/* random inputs                                                                                */
/* ------------------------------------------------------------------------------------------ */

uint64_t orc_splitmix64(uint64_t *state)
{
  uint64_t z = (*state += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

void orc_uniform01(uint64_t seed, double *out, size_t n)
{
  uint64_t s = seed;
  for (size_t i = 0; i < n; ++i) out[i] = (double)(orc_splitmix64(&s) >> 11) * 0x1.0p-53;
}

/* ------------------------------------------------------------------------------------------ */
/* helpers (math.hh:45-55 powr, math.hh:104-121 S_d_prec)                                      */
/* ------------------------------------------------------------------------------------------ */

static inline double ipow(double x, int n)
{
  if (n == 0) return 1.0;
  if (n < 0) return 1.0 / ipow(x, -n);
  /* powr<n>(x) = x * powr<n-1>(x) (math.hh:45-55): same association */
  return n == 1 ? x : x * ipow(x, n - 1);
}

static double S_d(int d)
{
  switch (d) {
  case 1: return 2.0;
  case 2: return 2.0 * M_PI;
  case 3: return 4.0 * M_PI;
  case 4: return 2.0 * M_PI * M_PI;
  case 5: return 8.0 * M_PI * M_PI / 3.0;
  case 6: return M_PI * M_PI * M_PI;
  case 7: return 16.0 * M_PI * M_PI * M_PI / 15.0;
  }
  return NAN;
}

/* ------------------------------------------------------------------------------------------ */
/* Gauss-Legendre on [0,1]                                                                      */
/* The reference obtains these from GSL's gsl_integration_fixed_legendre on [0,1]               */
/* (quadrature.cu:12-61).  Restated as Newton iteration on the Legendre three-term recurrence   */
/* in long double, which reproduces the exact rule to double rounding; pinned against the      */
/* 30-digit tables GLQuadrature<N> (gauss_legendre.hh).                                        */
/* ------------------------------------------------------------------------------------------ */

static void legendre_eval(int n, long double z, long double *pn, long double *dpn)
{
  long double p0 = 1.0L, p1 = z;
  if (n == 0) { *pn = 1.0L; *dpn = 0.0L; return; }
  for (int j = 2; j <= n; ++j) {
    long double p2 = ((2.0L * j - 1.0L) * z * p1 - (j - 1.0L) * p0) / (long double)j;
    p0 = p1;
    p1 = p2;
  }
  *pn = p1;
  *dpn = (long double)n * (z * p1 - p0) / (z * z - 1.0L);
}

int orc_gl_nodes(int n, double *x, double *w)
{
  if (n <= 0) return -1;
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int i = 0; i < (n + 1) / 2; ++i) {
    /* i-th root counted from the top, Tricomi-type initial guess */
    long double z = cosl(pi * ((long double)i + 0.75L) / ((long double)n + 0.5L));
    long double pn, dpn;
    for (int it = 0; it < 100; ++it) {
      legendre_eval(n, z, &pn, &dpn);
      long double dz = pn / dpn;
      z -= dz;
      if (fabsl(dz) <= 1e-19L * (fabsl(z) > 1e-3L ? fabsl(z) : 1e-3L)) break;
    }
    legendre_eval(n, z, &pn, &dpn);
    long double wz = 2.0L / ((1.0L - z * z) * dpn * dpn);
    /* map t = (1 +- z)/2, ascending */
    x[n - 1 - i] = (double)((1.0L + z) / 2.0L);
    x[i] = (double)((1.0L - z) / 2.0L);
    w[n - 1 - i] = (double)(wz / 2.0L);
    w[i] = (double)(wz / 2.0L);
  }
  if (n % 2 == 1) x[n / 2] = 0.5;
  return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* Implicit QL (diagonalization.hh:34-145): eigenvalues of the symmetric tridiagonal matrix    */
/* (d, e) overwrite d, z is replaced by Q'z, eigenvalues sorted ascending with z permuted.     */
/* ------------------------------------------------------------------------------------------ */

static inline double sgn(double v) { return v >= 0.0 ? 1.0 : -1.0; }

int orc_imtqlx(int n, double *d, double *e, double *z)
{
  if (n == 1) return 0;
  const double eps = DBL_EPSILON;
  const int max_iter = 30;
  e[n - 1] = 0.0;
  int m = 0;
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    for (;;) {
      /* look for a small sub-diagonal element */
      for (m = l; m < n - 1; ++m)
        if (fabs(e[m]) <= eps * (fabs(d[m]) + fabs(d[m + 1]))) break;
      double p = d[l];
      if (m == l) break;
      if (iter >= max_iter) return -1;
      ++iter;
      double g = (d[l + 1] - p) / (2.0 * e[l]);
      double r = sqrt(g * g + 1.0);
      g = d[m] - p + e[l] / (g + fabs(r) * sgn(g));
      double s = 1.0, c = 1.0;
      p = 0.0;
      for (int i = m - 1; i >= l; --i) {
        double f = s * e[i];
        double bb = c * e[i];
        if (fabs(g) <= fabs(f)) {
          c = g / f;
          r = sqrt(c * c + 1.0);
          e[i + 1] = f * r;
          s = 1.0 / r;
          c *= s;
        } else {
          s = f / g;
          r = sqrt(s * s + 1.0);
          e[i + 1] = g * r;
          c = 1.0 / r;
          s *= c;
        }
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * bb;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - bb;
        double zf = z[i + 1];
        z[i + 1] = s * z[i] + c * zf;
        z[i] = c * z[i] - s * zf;
      }
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  /* selection sort, ascending.  The reference sorts over ii = 2..m with m the last
   * convergence index, which equals n at exit (diagonalization.hh:125). */
  for (int i = 0; i < n - 1; ++i) {
    int k = i;
    double p = d[i];
    for (int j = i + 1; j < n; ++j)
      if (d[j] < p) { k = j; p = d[j]; }
    if (k != i) {
      d[k] = d[i];
      d[i] = p;
      double t = z[i];
      z[i] = z[k];
      z[k] = t;
    }
  }
  return 0;
}

int orc_make_quadrature(int n, double *a, double *b, double mu0, double *x, double *w)
{
  /* quadrature.hh:27-45: off-diagonal = sqrt(b), z = e_1, eigen-decompose, w = mu0 z^2 */
  w[0] = 1.0;
  for (int i = 1; i < n; ++i) {
    b[i - 1] = sqrt(b[i - 1]);
    w[i] = 0.0;
  }
  if (orc_imtqlx(n, a, b, w) != 0) return -1;
  for (int i = 0; i < n; ++i) {
    x[i] = a[i];
    w[i] = mu0 * (w[i] * w[i]);
  }
  return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* Matsubara quadrature (Monien, Math. Comp. 79, 857 (2010)), matsubara.cu:12-161              */
/* ------------------------------------------------------------------------------------------ */

static int is_close_eps(double a, double b, double eps)
{
  double diff = fabs(a - b);
  if (diff <= eps) return 1;
  if (diff <= fmax(fabs(a), fabs(b)) * eps) return 1;
  return 0;
}

int orc_matsubara_predict_size(double T, double E, int step, int vacuum_quad_size, int precision_factor)
{
  const double rel = fabs(E) / fabs(T + 1e-16);
  if (is_close_eps(T, 0.0, DBL_EPSILON) || rel > 1e+2) return -vacuum_quad_size;
  const double E_max = (precision_factor + 100) * fabs(E);
  int size = 5 + (int)sqrt(4.0 * E_max / (M_PI * M_PI * fabs(T)));
  size = (int)ceil(size / (double)step) * step;
  return size;
}

static int matsubara_vacuum(int size, double E, double *x, double *w, int cap)
{
  /* reinit_0 (matsubara.cu:128-161): GL on [0, 2E] plus log-mapped GL on [2E, 1e6 E], / 2pi */
  if (size % 2 != 0) size++;
  if (size > cap) return -1;
  if (is_close_eps(E, 0.0, DBL_EPSILON)) E = 1.0;
  const int h = size / 2;
  double *gx = (double *)malloc(sizeof(double) * h), *gw = (double *)malloc(sizeof(double) * h);
  orc_gl_nodes(h, gx, gw);
  const long double div = 2.0L * fabsl((long double)E);
  for (int i = 0; i < h; ++i) {
    x[i] = (double)(gx[i] * div);
    w[i] = (double)(gw[i] * div / (2.0 * M_PI));
  }
  const long double extent = 1e6L * fabsl((long double)E);
  const long double log_start = logl(div);
  const long double log_ext = logl(extent / div);
  for (int i = 0; i < h; ++i) {
    x[i + h] = (double)expl(log_start + log_ext * gx[i]);
    w[i + h] = (double)((gw[i] * log_ext * x[i + h]) / (2.0 * M_PI));
  }
  free(gx);
  free(gw);
  return size;
}

int orc_matsubara(double T, double E, int step, int min_size, int max_size, int vacuum_quad_size,
                  int precision_factor, double *x, double *w, int cap, double *T_eff)
{
  if (precision_factor < 0) precision_factor = 1;
  if (vacuum_quad_size <= 6) vacuum_quad_size = 6;
  int size = orc_matsubara_predict_size(T, E, step, vacuum_quad_size, precision_factor);
  if (size < 0) {
    if (T_eff) *T_eff = 0.0;
    return matsubara_vacuum(-size, E, x, w, cap);
  }
  if (size > max_size) size = max_size;
  if (size < min_size) size = min_size;
  if (size > cap) return -1;
  if (T_eff) *T_eff = T;
  double *a = (double *)calloc(size, sizeof(double)), *b = (double *)calloc(size, sizeof(double));
  for (int j = 0; j < size; ++j) {
    const double j1 = j + 1;
    a[j] = 2.0 * ipow(M_PI, 2) / (4 * j + 1) / (4 * j + 5);
    b[j] = ipow(M_PI, 4) / ((4 * j1 - 1) * (4 * j1 + 3)) / ipow(4 * j1 + 1, 2);
  }
  a[0] = ipow(M_PI, 2) / 15.0;
  const double mu0 = ipow(M_PI, 2) / 6.0;
  int rc = orc_make_quadrature(size, a, b, mu0, x, w);
  free(a);
  free(b);
  if (rc != 0) return -1;
  for (int i = 0; i < size; ++i) {
    w[i] = T * w[i] / x[i];
    x[i] = 2.0 * M_PI * T / sqrt(x[i]);
  }
  return size;
}

/* ------------------------------------------------------------------------------------------ */
/* regulators (regulators.hh:24-51 Litim, 144-191 exponential)                                  */
/* ------------------------------------------------------------------------------------------ */

static inline double xpow(double v, double b) { return b == 1.0 ? v : (b == 2.0 ? v * v : pow(v, b)); }

double orc_exp_RB(double b, double k2, double q2) { return k2 * exp(-xpow(q2 / k2, b)); }
double orc_exp_RBdot(double b, double k2, double q2)
{
  const double xb = xpow(q2 / k2, b);
  return 2.0 * exp(-xb) * k2 * (1.0 + b * xb);
}
double orc_exp_dq2RB(double b, double k2, double q2)
{
  const double xbm1 = xpow(q2 / k2, b - 1.0);
  const double xb = xbm1 * (q2 / k2);
  return -b * xbm1 * exp(-xb);
}
double orc_litim_RB(double k2, double q2) { return (k2 - q2) * (double)(k2 > q2); }
double orc_litim_RBdot(double k2, double q2) { return 2.0 * k2 * (double)(k2 > q2); }

/* ------------------------------------------------------------------------------------------ */
/* coordinates and interpolation                                                                */
/* ------------------------------------------------------------------------------------------ */

double orc_logcoord_forward(const orc_logcoord *c, unsigned i)
{
  const double a = c->bias, b = (c->stop - c->start) / expm1(a), gem1inv = 1.0 / (c->n - 1.0);
  return b * expm1(a * i * gem1inv) + c->start;
}

double orc_logcoord_backward(const orc_logcoord *c, double y)
{
  const double a = c->bias, b = (c->stop - c->start) / expm1(a), gem1 = c->n - 1.0;
  return log1p((y - c->start) / b) * gem1 / a;
}

static inline double lerp_idx(const double *data, unsigned n, double idx)
{
  idx = fmax(0.0, fmin(idx, (double)(n - 1)));
  const double fl = floor(idx);
  return data[(unsigned)fl] * (1.0 - idx + fl) + data[(unsigned)ceil(idx)] * (idx - fl);
}

double orc_lerp_log(const double *data, const orc_logcoord *c, double y)
{
  return lerp_idx(data, c->n, orc_logcoord_backward(c, y));
}

double orc_lerp_lin(const double *data, unsigned n, double start, double stop, double y)
{
  const double a = (stop - start) / (n - 1.0);
  return lerp_idx(data, n, (y - start) / a);
}

/* ------------------------------------------------------------------------------------------ */
/* C0: radial threshold integrals (integrator_cpu.hh:51-71)                                      */
/* ------------------------------------------------------------------------------------------ */

static double radial_threshold(int d, double m2, const double *x, const double *w, int nx, double x_extent,
                               double k, double b)
{
  const double Sd = S_d(d);
  const double k2 = k * k;
  double value = 0.0;
  for (int j = 0; j < nx; ++j) {
    const double q = k * sqrt(x[j] * x_extent);
    const double int_element = Sd * (ipow(q, d - 2) / 2.0 * (k * k)) / ipow(2.0 * M_PI, d);
    const double weight = w[j] * x_extent;
    const double q2 = q * q;
    const double den = q2 + orc_exp_RB(b, k2, q2) + m2;
    const double f = orc_exp_RBdot(b, k2, q2) / (den * den);
    value += int_element * weight * f;
  }
  return value;
}

void orc_c0_threshold(int d, const double *m2, int n_int, const double *x, const double *w, int nx,
                      double x_extent, double k, double b, double *out)
{
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n_int; ++i) out[i] = radial_threshold(d, m2[i], x, w, nx, x_extent, k, b);
}

/* ------------------------------------------------------------------------------------------ */
/* C1/C3: two-angle self-energy (integrator_4D_2ang_cpu.hh:62-100)                               */
/* ------------------------------------------------------------------------------------------ */

static inline double se_G(const double *tab, const orc_logcoord *c, double k2, double b, double y2, double y)
{
  const double Z = orc_lerp_log(tab, c, y);
  return 1.0 / (Z * y2 + orc_exp_RB(b, k2, y2));
}

static double selfenergy_one(const orc_se_integral *I, double x_extent, double k, double b)
{
  const double k2 = k * k;
  const double p = I->p;
  double total = 0.0;
  for (int ix = 0; ix < I->nq; ++ix) {
    const double q = k * sqrt(I->xq[ix] * x_extent);
    double vq = 0.0;
    for (int i1 = 0; i1 < I->na; ++i1) {
      const double cos1 = 2.0 * (I->xa[i1] - 0.5);
      const double int_element = 2.0 * M_PI * (q * q * 0.5 * k2) * sqrt(1.0 - cos1 * cos1) / ipow(2.0 * M_PI, 4);
      for (int i2 = 0; i2 < I->na; ++i2) {
        /* cos2 = 2 (xa[i2] - 1/2) is an argument of the kernel; this integrand does not use it */
        const double weight = 2.0 * I->wa[i2] * 2.0 * I->wa[i1] * I->wq[ix] * x_extent;
        /* kernel(q, cos1, cos2, k, p, Z) */
        const double q2 = q * q;
        const double pq2 = p * p + q2 - 2.0 * p * q * cos1;
        const double Gq = se_G(I->tab, &I->coord, k2, b, q2, q);
        const double Gpq = se_G(I->tab, &I->coord, k2, b, pq2, sqrt(pq2));
        const double f = orc_exp_RBdot(b, k2, q2) * Gq * Gq * Gpq;
        vq += int_element * weight * f;
      }
    }
    total += vq;
  }
  return total;
}

void orc_c1_selfenergy(const orc_se_integral *ints, int n_int, double x_extent, double k, double b, double *out)
{
#pragma omp parallel for schedule(dynamic, 1)
  for (int i = 0; i < n_int; ++i) out[i] = selfenergy_one(&ints[i], x_extent, k, b);
}

/* ------------------------------------------------------------------------------------------ */
/* C2: finite-T q0 form (integrator_finiteTq0_cpu.hh:77-114)                                     */
/* ------------------------------------------------------------------------------------------ */

static inline double ft_G(double w, double s, double m, double k2, double b)
{
  return 1.0 / (w * w + s + orc_exp_RB(b, k2, s) + m);
}

static inline double ft_kernel(double q, double q0, double k2, double b, const orc_ft_integral *I)
{
  const double q2 = q * q;
  const double g1 = ft_G(q0, q2, I->m1, k2, b);
  const double g2 = ft_G(q0 + I->p0, q2 + I->pabs * I->pabs, I->m2, k2, b);
  return orc_exp_RBdot(b, k2, q2) * g1 * g1 * g2;
}

static double finiteT_one(int d, int radial_mode, const orc_ft_integral *I, const double *x, const double *w,
                          int nx, double extent, double k, double b)
{
  const double Sdm1 = S_d(d - 1);
  const double k2 = k * k;
  double total = 0.0;
  for (int ix = 0; ix < nx; ++ix) {
    double q, x_weight, int_element;
    if (radial_mode == 0) {
      q = k * sqrt(x[ix] * extent);
      x_weight = w[ix] * extent;
      int_element = Sdm1 * (ipow(q, d - 3) / 2.0 * k2) / ipow(2.0 * M_PI, d - 1);
    } else {
      q = extent * x[ix];
      x_weight = w[ix] * extent;
      int_element = Sdm1 * ipow(q, d - 2) / ipow(2.0 * M_PI, d - 1);
    }
    double vq = 0.0;
    for (int iy = 0; iy < I->nm; ++iy) {
      const double q0 = I->mx[iy];
      const double weight = x_weight * I->mw[iy];
      vq += int_element * weight * (ft_kernel(q, q0, k2, b, I) + ft_kernel(q, -q0, k2, b, I));
      if (I->T > 0 && iy == 0) vq += int_element * x_weight * I->T * ft_kernel(q, 0.0, k2, b, I);
    }
    total += vq;
  }
  return total;
}

void orc_c2_finiteT(int d, int radial_mode, const orc_ft_integral *ints, int n_int, const double *x,
                    const double *w, int nx, double extent, double k, double b, double *out)
{
#pragma omp parallel for schedule(dynamic, 4)
  for (int i = 0; i < n_int; ++i) out[i] = finiteT_one(d, radial_mode, &ints[i], x, w, nx, extent, k, b);
}

/* ------------------------------------------------------------------------------------------ */
/* known-answer integrands of the reference tests                                               */
/* ------------------------------------------------------------------------------------------ */

static inline double poly6(const double *c, double v)
{
  return c[0] + c[1] * v + c[2] * v * v + c[3] * v * v * v + c[4] * v * v * v * v + c[5] * v * v * v * v * v;
}
static inline double poly4(const double *c, double v) { return c[0] + c[1] * v + c[2] * v * v + c[3] * v * v * v; }

double orc_kat_radial(int d, const double *x, const double *w, int nx, double x_extent, double k,
                      const double *c6, double constant)
{
  const double Sd = S_d(d);
  double value = 0.0;
  for (int j = 0; j < nx; ++j) {
    const double q = k * sqrt(x[j] * x_extent);
    const double int_element = Sd * (ipow(q, d - 2) / 2.0 * (k * k)) / ipow(2.0 * M_PI, d);
    const double weight = w[j] * x_extent;
    value += int_element * weight * poly6(c6, q);
  }
  return constant + value;
}

double orc_kat_4d2ang(const double *xq, const double *wq, int nq, const double *xa, const double *wa, int na,
                      double x_extent, double k, const double *c6, const double *a4, const double *b4,
                      double constant)
{
  double value = 0.0;
  for (int ix = 0; ix < nq; ++ix) {
    const double q = k * sqrt(xq[ix] * x_extent);
    for (int i1 = 0; i1 < na; ++i1) {
      const double cos1 = 2.0 * (xa[i1] - 0.5);
      const double int_element = 2.0 * M_PI * (q * q * 0.5 * k * k) * sqrt(1.0 - cos1 * cos1) / ipow(2.0 * M_PI, 4);
      for (int i2 = 0; i2 < na; ++i2) {
        const double cos2 = 2.0 * (xa[i2] - 0.5);
        const double weight = 2.0 * wa[i2] * 2.0 * wa[i1] * wq[ix] * x_extent;
        value += int_element * weight * poly6(c6, q) * poly4(a4, cos1) * poly4(b4, cos2);
      }
    }
  }
  return constant + value;
}

double orc_kat_4d(const double *xq, const double *wq, int nq, const double *xa, const double *wa, int na,
                  double x_extent, double k, const double *c6, const double *a4, const double *b4,
                  const double *p4, double constant)
{
  /* integrator_4D_cpu.hh:64-100 with equal angular orders */
  double value = 0.0;
  for (int ix = 0; ix < nq; ++ix) {
    const double q = k * sqrt(xq[ix] * x_extent);
    for (int i1 = 0; i1 < na; ++i1) {
      const double cos1 = 2.0 * (xa[i1] - 0.5);
      const double int_element = (q * q * 0.5 * k * k) * sqrt(1.0 - cos1 * cos1) / ipow(2.0 * M_PI, 4);
      for (int i2 = 0; i2 < na; ++i2) {
        const double cos2 = 2.0 * (xa[i2] - 0.5);
        for (int i3 = 0; i3 < na; ++i3) {
          const double phi = 2.0 * M_PI * xa[i3];
          const double weight = 2.0 * M_PI * wa[i3] * 2.0 * wa[i2] * 2.0 * wa[i1] * wq[ix] * x_extent;
          value += int_element * weight * poly6(c6, q) * poly4(a4, cos1) * poly4(b4, cos2) * poly4(p4, phi);
        }
      }
    }
  }
  return constant + value;
}

double orc_kat_finiteT(int d, const double *x, const double *w, int nx, double x_extent, double k,
                       const double *mx, const double *mw, int nm, double T, const double *c6,
                       const double *d4, double constant)
{
  const double Sdm1 = S_d(d - 1);
  double value = 0.0;
  for (int ix = 0; ix < nx; ++ix) {
    const double q = k * sqrt(x[ix] * x_extent);
    const double x_weight = w[ix] * x_extent;
    const double int_element = Sdm1 * (ipow(q, d - 3) / 2.0 * (k * k)) / ipow(2.0 * M_PI, d - 1);
    for (int iy = 0; iy < nm; ++iy) {
      const double q0 = mx[iy];
      const double weight = x_weight * mw[iy];
      value += int_element * weight * (poly6(c6, q) / poly4(d4, q0) + poly6(c6, q) / poly4(d4, -q0));
      if (T > 0 && iy == 0) value += int_element * x_weight * T * poly6(c6, q) / poly4(d4, 0.0);
    }
  }
  return constant + value;
}

double orc_kat_angle(int d, const double *xq, const double *wq, int nq, const double *xa, const double *wa, int na,
                     double x_extent, double k, const double *c6, const double *a4, double constant)
{
  /* integrator_angle_cpu.hh:66-79: S_d q^{d-2}/2 k^2 (1/2) / (2pi)^d, weight 2 w_cos w_x x_extent */
  const double Sd = S_d(d);
  double value = 0.0;
  for (int ix = 0; ix < nq; ++ix) {
    const double q = k * sqrt(xq[ix] * x_extent);
    double part = 0.0;
    for (int iy = 0; iy < na; ++iy) {
      const double cos1 = 2.0 * (xa[iy] - 0.5);
      const double int_element = Sd * (ipow(q, d - 2) / 2.0 * (k * k)) * 0.5 / ipow(2.0 * M_PI, d);
      const double weight = 2.0 * wa[iy] * wq[ix] * x_extent;
      part += int_element * weight * poly6(c6, q) * poly4(a4, cos1);
    }
    value += part;
  }
  return constant + value;
}

static inline double poly4c(const double *c, double v) { return c[0] + v * (c[1] + v * (c[2] + v * c[3])); }

double orc_kat_cartesian(int D, const double *const *x, const double *const *w, const int *n, const double *lo,
                         const double *hi, const double *c, double constant)
{
  /* integrator_{1,2,3}D_cartesian_cpu.hh (e.g. 3D: lines 98-121): q_i = q_min_i + q_ext_i x_i,
   * weight prod_i q_ext_i w_i, factor (2pi)^-D; integrand prod_i cubic_i(q_i), 4 coefficients per axis */
  const int ny = D > 1 ? n[1] : 1, nz = D > 2 ? n[2] : 1;
  const double el = 1.0 / ipow(2.0 * M_PI, D);
  double value = 0.0;
  for (int ix = 0; ix < n[0]; ++ix) {
    const double ex = hi[0] - lo[0], qx = lo[0] + ex * x[0][ix], wx = ex * w[0][ix];
    double part = 0.0;
    for (int iy = 0; iy < ny; ++iy) {
      double qy = 0.0, wy = 1.0;
      if (D > 1) {
        const double ey = hi[1] - lo[1];
        qy = lo[1] + ey * x[1][iy];
        wy = ey * w[1][iy];
      }
      for (int iz = 0; iz < nz; ++iz) {
        double qz = 0.0, wz = 1.0;
        if (D > 2) {
          const double ez = hi[2] - lo[2];
          qz = lo[2] + ez * x[2][iz];
          wz = ez * w[2][iz];
        }
        double f = poly4c(c, qx);
        if (D > 1) f *= poly4c(c + 4, qy);
        if (D > 2) f *= poly4c(c + 8, qz);
        part += el * (wz * wy * wx) * f;
      }
    }
    value += part;
  }
  return constant + value;
}

double orc_kat_freq(int dim, int q0_axis, int px0, const double *xq, const double *wq, int nq, const double *xa,
                    const double *wa, int na, const double *xb, const double *wb, int nb, double x_extent, double ext,
                    double k, const double *c6, const double *a4, const double *p4, const double *z4, double constant)
{
  /* integrator_3D_cpu.hh:41-67 (dim 3, no q0 axis): (q/2 k^2)/(2pi)^3, weight 2pi w_phi 2 w_cos w_x x_extent;
   * integrator_2Dpq0_cpu.hh:77-106 (dim 2): (k^2/2) S_2 (1/2)/(2pi)^3, weight w_q0 q0_ext 2 w_cos w_x x_extent,
   *   f(q0) + f(-q0) with q0 = x q0_ext;
   * integrator_3Dpq0_cpu.hh:43-75 (dim 3): (q/2 k^2)/(2pi)^4, weight w_q0 q0_ext 2pi w_phi 2 w_cos w_x x_extent;
   * px0 (integrator_2Dpx0_cpu.hh:78-108, integrator_3Dpx0_cpu.hh:43-75): q0 = k x x0_ext, measure times k.
   * Integrand poly6(q) poly4(cos) [poly4(phi)] [poly4(q0)]. */
  const int nphi = dim == 3 ? na : 1, nz = q0_axis ? nb : 1;
  double value = 0.0;
  for (int ix = 0; ix < nq; ++ix) {
    const double q = k * sqrt(xq[ix] * x_extent);
    double int_element;
    if (dim == 2)
      int_element = (0.5 * k * k) * (2.0 * M_PI) * 0.5 / ipow(2.0 * M_PI, 3);
    else
      int_element = (q / 2.0 * k * k) / ipow(2.0 * M_PI, q0_axis ? 4 : 3);
    if (px0) int_element *= k;
    double part = 0.0;
    for (int ic = 0; ic < na; ++ic) {
      const double cos1 = 2.0 * (xa[ic] - 0.5);
      for (int ip = 0; ip < nphi; ++ip) {
        const double phi = 2.0 * M_PI * xa[ip];
        for (int iz = 0; iz < nz; ++iz) {
          double weight = 2.0 * wa[ic] * wq[ix] * x_extent;
          double f = poly6(c6, q) * poly4(a4, cos1);
          if (dim == 3) {
            weight *= 2.0 * M_PI * wa[ip];
            f *= poly4(p4, phi);
          }
          if (q0_axis) {
            const double q0 = (px0 ? k : 1.0) * xb[iz] * ext;
            weight *= wb[iz] * ext;
            f *= poly4(z4, q0) + poly4(z4, -q0);
          }
          part += int_element * weight * f;
        }
      }
    }
    value += part;
  }
  return constant + value;
}

double orc_kat_finiteTx0(int kind, int d, const double *xq, const double *wq, int nq, const double *xa,
                         const double *wa, int na, const double *x0, const double *w0, int n0, double x_extent,
                         double x0_extent, int summands, double T, double k, const double *c6, const double *a4,
                         const double *p4, const double *z4, double constant)
{
  /* integrator_finiteTx0_cpu.hh:64-109 (kind 0), integrator_angle_finiteTx0_cpu.hh:69-122 (kind 1),
   * integrator_4D_finiteTx0_cpu.hh:66-124 (kind 2, d = 4): sum over n < summands (summands = 0 at T = 0) with
   * T S_{d-1} q^{d-3}/2 k^2 [1/2] / (2pi)^{d-1}, integral over x0 nodes log-mapped from 2 pi summands T / k
   * (1e-3 at T = 0, shifted back) to x0_extent with S_{d-1} q^{d-3}/2 k^2 k [1/2] / (2pi)^d and weight
   * w log_ext q0 / k; kind 2 has no S_{d-1} and no 1/2 and explicit cos and phi weights.  Integrand
   * poly6(q) [poly4(cos)] [poly4(phi)] poly4(q0) + (q0 -> -q0). */
  const double z = T == 0.0 ? 1e-3 : 0.0;
  const int nm = T == 0.0 ? 0 : summands;
  const double start = (2.0 * nm * M_PI * T) / k;
  const double ls = log(start + z), le = log(x0_extent / (start + z));
  const double S = kind == 2 ? 1.0 : S_d(d - 1);
  const int nc = kind >= 1 ? na : 1, nphi = kind == 2 ? na : 1;
  double value = 0.0;
  for (int ix = 0; ix < nq; ++ix) {
    const double q = k * sqrt(xq[ix] * x_extent);
    const double base = S * (ipow(q, d - 3) / 2.0 * k * k) * (kind == 1 ? 0.5 : 1.0);
    double part = 0.0;
    for (int ic = 0; ic < nc; ++ic) {
      const double cos1 = 2.0 * (xa[ic] - 0.5);
      for (int ip = 0; ip < nphi; ++ip) {
        const double phi = 2.0 * M_PI * xa[ip];
        double weight = wq[ix] * x_extent;
        double g = poly6(c6, q);
        if (kind >= 1) {
          weight *= 2.0 * wa[ic];
          g *= poly4(a4, cos1);
        }
        if (kind == 2) {
          weight *= 2.0 * M_PI * wa[ip];
          g *= poly4(p4, phi);
        }
        for (int n = 0; n < nm; ++n) {
          const double q0 = 2.0 * M_PI * T * n;
          const double f = n == 0 ? poly4(z4, 0.0) : poly4(z4, q0) + poly4(z4, -q0);
          part += T * base / ipow(2.0 * M_PI, d - 1) * weight * g * f;
        }
        for (int j = 0; j < n0; ++j) {
          const double q0 = k * (exp(ls + le * x0[j]) - z);
          const double wj = w0[j] * le * q0 / k;
          part += base * k / ipow(2.0 * M_PI, d) * weight * wj * g * (poly4(z4, q0) + poly4(z4, -q0));
        }
      }
    }
    value += part;
  }
  return constant + value;
}

static inline double poly5(const double *c, double v)
{
  return c[0] + c[1] * v + c[2] * v * v + c[3] * v * v * v + c[4] * v * v * v * v;
}

double orc_kat_4d_finiteT(const double *xq, const double *wq, int nq, const double *xa, const double *wa, int na,
                          double x_extent, double k, const double *mx, const double *mw, int nm, double T,
                          const double *c6, const double *a4, const double *p4, const double *d5, double constant)
{
  /* integrator_4D_finiteTq0_cpu.hh:95-117: q/2 k^2 / (2pi)^3, weight 2pi w_phi 2 w_cos w_x x_extent,
   * sum_n w_n (f(q0_n) + f(-q0_n)) + T f(0) */
  double value = 0.0;
  for (int ix = 0; ix < nq; ++ix) {
    const double q = k * sqrt(xq[ix] * x_extent);
    const double int_element = (q / 2.0 * (k * k)) / ipow(2.0 * M_PI, 3);
    double part = 0.0;
    for (int iy = 0; iy < na; ++iy) {
      const double cos1 = 2.0 * (xa[iy] - 0.5);
      for (int iz = 0; iz < na; ++iz) {
        const double phi = 2.0 * M_PI * xa[iz];
        const double x_weight = 2.0 * M_PI * wa[iz] * 2.0 * wa[iy] * wq[ix] * x_extent;
        const double sp = poly6(c6, q) * poly4(a4, cos1) * poly4(p4, phi);
        for (int in = 0; in < nm; ++in) {
          const double q0 = mx[in];
          part += int_element * x_weight * mw[in] * (sp / poly5(d5, q0) + sp / poly5(d5, -q0));
          if (T > 0 && in == 0) part += int_element * x_weight * T * sp / poly5(d5, 0.0);
        }
      }
    }
    value += part;
  }
  return constant + value;
}

double orc_kat_angle_finiteT(int d, const double *xq, const double *wq, int nq, const double *xa, const double *wa,
                             int na, double x_extent, double k, const double *mx, const double *mw, int nm, double T,
                             const double *c6, const double *a4, const double *d4, double constant)
{
  /* integrator_angle_finiteTq0_cpu.hh:95-118: S_{d-1} q^{d-3}/2 k^2 (1/2) / (2pi)^{d-1},
   * weight 2 w_cos w_x x_extent, Matsubara sum as above */
  const double Sdm1 = S_d(d - 1);
  double value = 0.0;
  for (int ix = 0; ix < nq; ++ix) {
    const double q = k * sqrt(xq[ix] * x_extent);
    const double int_element = Sdm1 * (ipow(q, d - 3) / 2.0 * (k * k)) * 0.5 / ipow(2.0 * M_PI, d - 1);
    double part = 0.0;
    for (int iy = 0; iy < na; ++iy) {
      const double cos1 = 2.0 * (xa[iy] - 0.5);
      const double x_weight = 2.0 * wa[iy] * wq[ix] * x_extent;
      const double sp = poly6(c6, q) * poly4(a4, cos1);
      for (int in = 0; in < nm; ++in) {
        const double q0 = mx[in];
        part += int_element * x_weight * mw[in] * (sp / poly4(d4, q0) + sp / poly4(d4, -q0));
        if (T > 0 && in == 0) part += int_element * x_weight * T * sp / poly4(d4, 0.0);
      }
    }
    value += part;
  }
  return constant + value;
}

/* ------------------------------------------------------------------------------------------ */
/* LPA flow                                                                                     */
/* ------------------------------------------------------------------------------------------ */

/* S2(E) = T sum_n (w_n^2 + E^2)^-2 with e = exp(-E/T):
 *   coth(E/2T) = (1+e)/(1-e),  csch^2(E/2T) = 4e/(1-e)^2,  1-e = -expm1(-E/T). */
static inline double bose_S2(double E, double T)
{
  const double e = exp(-E / T);
  const double ome = -expm1(-E / T);
  const double coth = (1.0 + e) / ome;
  const double csch2 = 4.0 * e / (ome * ome);
  return coth / (4.0 * E * E * E) + csch2 / (8.0 * T * E * E);
}

double orc_lpa_threshold(double m2, double T, double k, const double *x, const double *w, int nx,
                         double x_extent, double b)
{
  /* radial d=3 element (integrator_cpu.hh:61-66) */
  const double k2 = k * k;
  const double S3 = S_d(3);
  double value = 0.0;
  for (int j = 0; j < nx; ++j) {
    const double q = k * sqrt(x[j] * x_extent);
    const double int_element = S3 * (q / 2.0 * k2) / ipow(2.0 * M_PI, 3);
    const double weight = w[j] * x_extent;
    const double q2 = q * q;
    const double E = sqrt(q2 + orc_exp_RB(b, k2, q2) + m2);
    value += int_element * weight * (orc_exp_RBdot(b, k2, q2) * bose_S2(E, T));
  }
  return value;
}

/* orc_lpa_threshold with the node-only factors computed once per right-hand side: a_j = element
 * x weight, r_j = RBdot, s_j = q^2 + RB (the same values and operation order, so bitwise the same). */
static double lpa_threshold_tab(double m2, double T, const double *a, const double *r, const double *s, int nx)
{
  double value = 0.0;
  for (int j = 0; j < nx; ++j) {
    const double E = sqrt(s[j] + m2);
    value += a[j] * (r[j] * bose_S2(E, T));
  }
  return value;
}

void orc_flow_rhs(const orc_flow_params *fp, const double *u, double k, const double *x, const double *w,
                  int nx, double x_extent, double b, double *dudt)
{
  const int n = fp->n_grid;
  const double h = fp->rho_max / (n - 1);
  double *tab = (double *)malloc(sizeof(double) * 3 * nx);
  double *ta = tab, *tr = tab + nx, *ts = tab + 2 * nx;
  {
    const double k2 = k * k, S3 = S_d(3);
    for (int j = 0; j < nx; ++j) {
      const double q = k * sqrt(x[j] * x_extent);
      const double q2 = q * q;
      ta[j] = (S3 * (q / 2.0 * k2) / ipow(2.0 * M_PI, 3)) * (w[j] * x_extent);
      tr[j] = orc_exp_RBdot(b, k2, q2);
      ts[j] = q2 + orc_exp_RB(b, k2, q2);
    }
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    double up, upp;
    if (i == 0) {
      up = (-3.0 * u[0] + 4.0 * u[1] - u[2]) / (2.0 * h);
      upp = (u[0] - 2.0 * u[1] + u[2]) / (h * h);
    } else if (i == n - 1) {
      up = (3.0 * u[n - 1] - 4.0 * u[n - 2] + u[n - 3]) / (2.0 * h);
      upp = (u[n - 1] - 2.0 * u[n - 2] + u[n - 3]) / (h * h);
    } else {
      up = (u[i + 1] - u[i - 1]) / (2.0 * h);
      upp = (u[i + 1] - 2.0 * u[i] + u[i - 1]) / (h * h);
    }
    const double rho = i * h;
    const double m2pi = u[i];
    const double m2sigma = u[i] + 2.0 * rho * up;
    const double Ipi = lpa_threshold_tab(m2pi, fp->T, ta, tr, ts, nx);
    const double Isig = lpa_threshold_tab(m2sigma, fp->T, ta, tr, ts, nx);
    dudt[i] = 0.5 * ((fp->N - 1.0) * up * Ipi + (3.0 * up + 2.0 * rho * upp) * Isig);
  }
  free(tab);
}

int orc_flow_lpa(const orc_flow_params *fp, const double *x, const double *w, int nx, double x_extent,
                 double b, double k_end, int n_steps, double *u_out)
{
  const int n = fp->n_grid;
  if (n < 4) return -1;
  const double h = fp->rho_max / (n - 1);
  double *u = u_out;
  double *k1 = (double *)malloc(sizeof(double) * n), *k2 = (double *)malloc(sizeof(double) * n);
  double *k3 = (double *)malloc(sizeof(double) * n), *k4 = (double *)malloc(sizeof(double) * n);
  double *tmp = (double *)malloc(sizeof(double) * n);
  for (int i = 0; i < n; ++i) u[i] = fp->m2_0 + fp->lambda / 2.0 * (i * h);
  const double dt = log(1.0 / k_end) / n_steps;
  int rc = 0;
  for (int s = 0; s < n_steps; ++s) {
    const double t = s * dt;
    const double ka = exp(-t), kb = exp(-(t + 0.5 * dt)), kc = exp(-(t + dt));
    orc_flow_rhs(fp, u, ka, x, w, nx, x_extent, b, k1);
    for (int i = 0; i < n; ++i) tmp[i] = u[i] + 0.5 * dt * k1[i];
    orc_flow_rhs(fp, tmp, kb, x, w, nx, x_extent, b, k2);
    for (int i = 0; i < n; ++i) tmp[i] = u[i] + 0.5 * dt * k2[i];
    orc_flow_rhs(fp, tmp, kb, x, w, nx, x_extent, b, k3);
    for (int i = 0; i < n; ++i) tmp[i] = u[i] + dt * k3[i];
    orc_flow_rhs(fp, tmp, kc, x, w, nx, x_extent, b, k4);
    int bad = 0;
    for (int i = 0; i < n; ++i) {
      u[i] = u[i] + dt / 6.0 * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]);
      if (!isfinite(u[i])) bad = 1;
    }
    if (bad) { rc = -1; break; }
  }
  free(k1); free(k2); free(k3); free(k4); free(tmp);
  return rc;
}

/* Many flows, one flow per thread (OpenMP over flows; the right-hand side of each flow then runs
 * on its thread alone).  u_out[f] receives flow f's n_grid values; rc[f] its return code. */
void orc_flow_lpa_many(const orc_flow_params *fps, int n_flows, const double *x, const double *w, int nx,
                       double x_extent, double b, double k_end, int n_steps, double **u_out, int *rc)
{
#pragma omp parallel for schedule(dynamic, 1)
  for (int f = 0; f < n_flows; ++f) rc[f] = orc_flow_lpa(&fps[f], x, w, nx, x_extent, b, k_end, n_steps, u_out[f]);
}
