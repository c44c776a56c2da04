/*
 * oracle.h -- CPU restatement of DiFfRG v1.0.1's momentum-loop integrator path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in diffrg_amd/ or include/ links, loads or
 * calls this code.  It is used by tests/ (as the parity reference), by the smoke
 * check (as its reference values) and by bench.py's cpu_baseline leg (as the timed CPU
 * baseline).  Every function cites the reference file:line it restates; paths
 * are relative to the DiFfRG source tree.
 *
 * Pinning (what checks each part, see tests/test_oracle_*.py):
 *   orc_gl_nodes          30-digit GLQuadrature<N> tables (gauss_legendre.hh), fixture
 *   orc_imtqlx            2x2/3x3 closed forms (tests/common/quadrature/test_diagonalization.cc),
 *                         Legendre Jacobi matrix vs the GL tables (test_quadrature.cc:10-33)
 *   orc_matsubara         closed-form Matsubara sums at 1e-13 and the T=0 limit
 *                         (test_matsubara.cc:11-59), brute-force sums
 *   regulators, coords,   closed forms (regulators.hh:144-191, coordinates.hh:141-196),
 *   lerp                  exact lerp on linear data (linear_interpolation_1D_cpu.cc:28-38)
 *   orc_int_*             polynomial / volume known answers of the reference tests
 *                         (integrator_cpu.cc, integrator_4D_cpu.cc, integrator_4D_2ang_cpu.cc,
 *                         integrator_finiteTq0_cpu.cc), brute force on tiny grids
 *   orc_flow_*            linear-in-rho exact solution and mpmath-free self-convergence
 */
#ifndef DFRG_ORACLE_H
#define DFRG_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- random inputs ---------------- */
/* SplitMix64 (Steele, Lea, Flood 2014); u = (z >> 11) * 2^-53. */
uint64_t orc_splitmix64(uint64_t *state);
void orc_uniform01(uint64_t seed, double *out, size_t n);

/* ---------------- quadrature ---------------- */
/* Gauss-Legendre nodes/weights on [0,1], ascending (quadrature.cu:12-61 uses GSL on [0,1]). */
int orc_gl_nodes(int n, double *x, double *w);
/* Implicit QL on a symmetric tridiagonal matrix (diagonalization.hh:34-145). 0 ok, -1 iteration limit. */
int orc_imtqlx(int n, double *d, double *e, double *z);
/* Golub-Welsch from a three-term recurrence (quadrature.hh:27-45). b holds squared off-diagonals. */
int orc_make_quadrature(int n, double *a, double *b, double mu0, double *x, double *w);
/* MatsubaraQuadrature::predict_size (matsubara.cu:12-21). Negative = T=0 rule of that size. */
int orc_matsubara_predict_size(double T, double E, int step, int vacuum_quad_size, int precision_factor);
/* MatsubaraQuadrature::reinit / reinit_0 (matsubara.cu:33-85, 128-161).  Writes at most cap nodes;
 * returns the rule size (or -1 on error).  *T_eff receives the rule's T (0 for the vacuum rule). */
int orc_matsubara(double T, double E, int step, int min_size, int max_size, int vacuum_quad_size,
                  int precision_factor, double *x, double *w, int cap, double *T_eff);

/* ---------------- regulators, coordinates, interpolation ---------------- */
double orc_exp_RB(double b, double k2, double q2);     /* regulators.hh:150-155 */
double orc_exp_RBdot(double b, double k2, double q2);  /* regulators.hh:163-169 */
double orc_exp_dq2RB(double b, double k2, double q2);  /* regulators.hh:156-162 */
double orc_litim_RB(double k2, double q2);             /* regulators.hh:24-28 */
double orc_litim_RBdot(double k2, double q2);          /* regulators.hh:30-34 */
/* LogarithmicCoordinates1D (coordinates.hh:141-196). */
typedef struct { double start, stop, bias; unsigned n; } orc_logcoord;
double orc_logcoord_forward(const orc_logcoord *c, unsigned i);
double orc_logcoord_backward(const orc_logcoord *c, double y);
/* LinearInterpolator1D::operator() with log coordinates (linear_interpolation_1D.hh:75-92). */
double orc_lerp_log(const double *data, const orc_logcoord *c, double y);
double orc_lerp_lin(const double *data, unsigned n, double start, double stop, double y);

/* ---------------- integrators with built-in integrands ---------------- */
/* Radial d-dim (integrator_cpu.hh:51-71) with the LPA threshold integrand
 *   f(q) = RBdot(k^2,q^2) / (q^2 + RB(k^2,q^2) + m2)^2, exponential regulator exponent b.
 * One integral per m2[i]; out[i]. */
void orc_c0_threshold(int d, const double *m2, int n_int, const double *x, const double *w, int nx,
                      double x_extent, double k, double b, double *out);

/* 4D two-angle (integrator_4D_2ang_cpu.hh:62-100) with the self-energy integrand
 *   f = RBdot(q) G(q)^2 G(|p-q|),  G(y) = 1 / (Z(y) y^2 + RB(k^2, y^2)),  |p-q|^2 = p^2+q^2-2pq cos1,
 * Z a per-integral table on log coordinates.  Integral i uses nodes (xq[i], nq[i]) and (xa[i], na[i])
 * for both angles, table tabs + tab_off[i] with coordinates coords[i] and momentum p[i]. */
typedef struct {
  const double *xq, *wq; int nq;
  const double *xa, *wa; int na;
  const double *tab; orc_logcoord coord;
  double p;
} orc_se_integral;
void orc_c1_selfenergy(const orc_se_integral *ints, int n_int, double x_extent, double k, double b, double *out);

/* Finite-T q0 form (integrator_finiteTq0_cpu.hh:77-114) with the two-propagator integrand
 *   f(q,q0) = RBdot(q^2) G(q0,q^2,m1)^2 G(q0+p0, q^2+|p|^2, m2),  G(w,s,m) = 1/(w^2+s+RB(k^2,s)+m).
 * radial_mode 0: reference mapping q = k sqrt(x x_ext), element S_{d-1} q^{d-3}/2 k^2/(2pi)^{d-1}, weight w x_ext.
 * radial_mode 1: linear |q| mapping q = qmax x, element S_{d-1} q^{d-2}/(2pi)^{d-1}, weight w qmax.
 * Matsubara rule per integral: (mx + mat_off[i], mw + mat_off[i], mat_n[i]) plus T[i] f(q,0). */
typedef struct {
  double T, p0, pabs, m1, m2;
  const double *mx, *mw; int nm;
} orc_ft_integral;
void orc_c2_finiteT(int d, int radial_mode, const orc_ft_integral *ints, int n_int, const double *x,
                    const double *w, int nx, double extent, double k, double b, double *out);

/* Known-answer integrands of the reference tests. */
/* radial: poly(q) = sum c_i q^i (i<6) + constant (tests/physics/integration/integrator_cpu.cc:15-32) */
double orc_kat_radial(int d, const double *x, const double *w, int nx, double x_extent, double k,
                      const double *c6, double constant);
/* 4D 2ang: poly(q) poly1(cos1) poly2(cos2) (integrator_4D_2ang_cpu.cc:14-36) */
double orc_kat_4d2ang(const double *xq, const double *wq, int nq, const double *xa, const double *wa, int na,
                      double x_extent, double k, const double *c6, const double *a4, const double *b4,
                      double constant);
/* 4D: poly(q) poly1(cos1) poly2(cos2) poly3(phi) (integrator_4D_cpu.cc:14-39) */
double orc_kat_4d(const double *xq, const double *wq, int nq, const double *xa, const double *wa, int na,
                  double x_extent, double k, const double *c6, const double *a4, const double *b4,
                  const double *p4, double constant);
/* finite-T q0: poly(q) / poly(q0) (integrator_finiteTq0_cpu.cc:14-34) */
double orc_kat_finiteT(int d, const double *x, const double *w, int nx, double x_extent, double k,
                       const double *mx, const double *mw, int nm, double T, const double *c6,
                       const double *d4, double constant);
/* angle: poly(q) poly1(cos) in d dims (integrator_angle_cpu.cc, integrator_angle_cpu.hh:60-80) */
double orc_kat_angle(int d, const double *xq, const double *wq, int nq, const double *xa, const double *wa, int na,
                     double x_extent, double k, const double *c6, const double *a4, double constant);
/* Cartesian box in D = 1..3 dims: prod_i cubic_i(q_i) over [lo_i, hi_i], factor (2pi)^-D
 * (integrator_{1,2,3}D_cartesian_cpu.cc, integrator_3D_cartesian_cpu.hh:98-121); x[i], w[i]: the
 * rule of axis i on [0,1] with n[i] nodes; c: 4 coefficients per axis */
double orc_kat_cartesian(int D, const double *const *x, const double *const *w, const int *n, const double *lo,
                         const double *hi, const double *c, double constant);
/* zero-temperature (q, cos[, phi][, q0]) forms: dim 3 without q0 axis = Integrator3D; dim 2/3 with a q0 axis =
 * Integrator{2,3}Dpq0 (px0: the axis in x0 = q0/k); integrand poly6(q) poly4(cos) [poly4(phi)]
 * [poly4(q0) + poly4(-q0)] (integrator_3D_cpu.cc, integrator_{2,3}D{pq0,px0}_cpu.cc) */
double orc_kat_freq(int dim, int q0_axis, int px0, const double *xq, const double *wq, int nq, const double *xa,
                    const double *wa, int na, const double *xb, const double *wb, int nb, double x_extent, double ext,
                    double k, const double *c6, const double *a4, const double *p4, const double *z4, double constant);
/* finite-T in x0 = q0/k: kind 0 IntegratorFiniteTx0<d>, 1 IntegratorAngleFiniteTx0<d>, 2 Integrator4DFiniteTx0;
 * x0, w0: the integral's rule on [0,1] (grid order + summands at T = 0); integrand poly6(q) [poly4(cos)]
 * [poly4(phi)] poly4(q0) (integrator_{,angle_,4D_}finiteTx0_cpu.cc) */
double orc_kat_finiteTx0(int kind, int d, const double *xq, const double *wq, int nq, const double *xa,
                         const double *wa, int na, const double *x0, const double *w0, int n0, double x_extent,
                         double x0_extent, int summands, double T, double k, const double *c6, const double *a4,
                         const double *p4, const double *z4, double constant);
/* 4D finite-T q0: poly(q) poly1(cos) poly2(phi) / poly5(q0) (integrator_4D_finiteTq0_cpu.cc:14-40,
 * integrator_4D_finiteTq0_cpu.hh:90-120); equal cos and phi orders */
double orc_kat_4d_finiteT(const double *xq, const double *wq, int nq, const double *xa, const double *wa, int na,
                          double x_extent, double k, const double *mx, const double *mw, int nm, double T,
                          const double *c6, const double *a4, const double *p4, const double *d5, double constant);
/* angle finite-T q0 in d-1 spatial dims: poly(q) poly1(cos) / poly4(q0)
 * (integrator_angle_finiteTq0_gpu.cu:14-34, integrator_angle_finiteTq0_cpu.hh:95-120) */
double orc_kat_angle_finiteT(int d, const double *xq, const double *wq, int nq, const double *xa, const double *wa,
                             int na, double x_extent, double k, const double *mx, const double *mw, int nm, double T,
                             const double *c6, const double *a4, const double *d4, double constant);

/* ---------------- LPA flow (uniform rho grid, explicit RK4) ---------------- */
/* Finite-T d=3 threshold integral with the closed-form bosonic Matsubara sum,
 *   I_T(m2) = int d^3q/(2pi)^3 RBdot(k^2,q^2) S2(E),  E^2 = q^2+RB+m2,
 *   S2(E) = T sum_n 1/(w_n^2+E^2)^2 = coth(E/2T)/(4E^3) + 1/(8 T E^2 sinh^2(E/2T)). */
double orc_lpa_threshold(double m2, double T, double k, const double *x, const double *w, int nx,
                         double x_extent, double b);
typedef struct {
  int n_grid; double rho_max; double T; double mu; double N; double m2_0; double lambda;
} orc_flow_params;
/* Integrate u = U'(rho) from k=Lambda=1 to k_end with n_steps fixed RK4 steps in t = ln(Lambda/k).
 * du/dt = 1/2 [ (N-1) u' I_T(u) + (3u' + 2 rho u'') I_T(u + 2 rho u') ].  u_out has n_grid entries.
 * Returns 0, or -1 if the flow became non-finite. */
int orc_flow_lpa(const orc_flow_params *fp, const double *x, const double *w, int nx, double x_extent,
                 double b, double k_end, int n_steps, double *u_out);
/* One RHS evaluation (exposed for tests). */
/* Many flows at once, OpenMP over the flows. */
void orc_flow_lpa_many(const orc_flow_params *fps, int n_flows, const double *x, const double *w, int nx,
                       double x_extent, double b, double k_end, int n_steps, double **u_out, int *rc);
void orc_flow_rhs(const orc_flow_params *fp, const double *u, double k, const double *x, const double *w,
                  int nx, double x_extent, double b, double *dudt);

/* threads used by the OpenMP loops (omp_get_max_threads) */
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
