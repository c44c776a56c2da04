/*
 * dfrg.h -- C ABI of libdfrg.so, the MI355X (gfx950) runtime under the DiFfRG-compatible
 * integrator headers in include/DiFfRG/.
 *
 * Conventions
 *   - Every function returns an int status: DFRG_OK (0) or a negative DFRG_ERR_* code.
 *     dfrg_last_error() returns a thread-local message for the last failure on the calling
 *     thread.  The C++ headers rethrow DFRG_ERR_INVALID as std::invalid_argument and every
 *     other code as std::runtime_error, which is what the reference throws on the same faults.
 *   - Pointers named d_* are device (HBM) pointers, h_* are host pointers.  `stream` is a
 *     hipStream_t passed as void* (NULL = the library's stream of the calling thread).
 *   - Batched entry points read inputs that are already resident in HBM and write results to
 *     HBM; they enqueue work and return without synchronising the stream.
 *   - Objects (dfrg_qp, dfrg_se_batch, dfrg_fs) are created and destroyed in pairs; the library
 *     owns their device memory.  A dfrg_qp must outlive every object created from it.
 *
 * Reference interfaces replaced (paths relative to the DiFfRG source tree):
 *   QuadratureProvider            DiFfRG/include/DiFfRG/common/quadrature/quadrature_provider.hh:138-248
 *   Quadrature<NT>::reinit        DiFfRG/src/DiFfRG/common/quadrature/quadrature.cu:12-61
 *   MatsubaraQuadrature<NT>       DiFfRG/src/DiFfRG/common/quadrature/matsubara.cu:12-161
 *   IntegratorGPU<d>::get         DiFfRG/include/DiFfRG/physics/integration/integrator_gpu.hh:34-60,99-106
 *   Integrator4D2AngGPU::get      DiFfRG/include/DiFfRG/physics/integration/integrator_4D_2ang_gpu.hh:20-45,138-197
 *   IntegratorFiniteTq0GPU::get   DiFfRG/include/DiFfRG/physics/integration_finiteT/integrator_finiteTq0_gpu.hh:20-44,142-179
 *   request_data / update_data    DiFfRG/include/DiFfRG/physics/utils.hh:31-147 (many requests, one sync)
 *   check_cuda                    DiFfRG/src/DiFfRG/common/cuda_prefix.cu:9-14
 */
#ifndef DFRG_H
#define DFRG_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DFRG_OK 0
#define DFRG_ERR_INVALID (-1)   /* invalid argument (std::invalid_argument in the C++ headers) */
#define DFRG_ERR_RUNTIME (-2)   /* runtime failure (std::runtime_error) */
#define DFRG_ERR_HIP (-3)       /* a HIP call failed; message names the call (check_cuda analogue) */
#define DFRG_ERR_NO_DEVICE (-4) /* no GPU visible ("No CUDA devices found!" in the reference) */
#define DFRG_ERR_NOT_INIT (-5)  /* dfrg_init was not called */

#define DFRG_FP64 0  /* plain: everything in FP64 */
#define DFRG_MIXED 1 /* integrand in FP32, nodes/weights/denominators/accumulation in FP64 */

/* ------------------------------------------------------------------------------------------ */
/* runtime                                                                                    */
/* ------------------------------------------------------------------------------------------ */
/* Select `device` (ordinal among visible GPUs) and initialise the library on it. Idempotent. */
int dfrg_init(int device);
/* Release every library resource on the current device. */
int dfrg_finalize(void);
const char *dfrg_last_error(void);
int dfrg_device_count(int *count);
/* name: buffer of len bytes; any output pointer may be NULL. */
int dfrg_device_props(char *name, int len, int *cu_count, int *clock_khz, unsigned long long *hbm_bytes);
/* Measured vector FMA rate of the device in FLOP/s (fp64 = 1: FP64 FMAs, 0: plain FP32 FMAs), from
 * independent FMA chains at eight waves per SIMD; the best of three ~1 ms launches after a warm-up.
 * No counterpart in the reference: it confirms on the device the published vector peak that the
 * integrands' VALU roofline is quoted against.  Synchronous; allocates and frees its own buffer. */
int dfrg_valu_peak(int fp64, double *flop_per_s);

int dfrg_malloc(void **d_ptr, size_t bytes);
int dfrg_free(void *d_ptr);
int dfrg_host_alloc(void **h_ptr, size_t bytes); /* pinned host memory */
int dfrg_host_free(void *h_ptr);
/* Copies are enqueued on `stream` (NULL: the calling thread's stream, below); synchronise that
 * stream before reusing or reading the host buffer. */
int dfrg_memcpy_h2d(void *d_dst, const void *h_src, size_t bytes, void *stream);
int dfrg_memcpy_d2h(void *h_dst, const void *d_src, size_t bytes, void *stream);
int dfrg_memcpy_d2d(void *d_dst, const void *d_src, size_t bytes, void *stream);
/* The calling thread's stream (created on first use; the reference uses a per-thread
 * default stream, setup_build_system.cmake:185-197 "--default-stream per-thread"). */
int dfrg_thread_stream(void **stream);
int dfrg_stream_create(void **stream);
int dfrg_stream_destroy(void *stream);
int dfrg_stream_sync(void *stream);
int dfrg_event_create(void **event);
int dfrg_event_destroy(void *event);
int dfrg_event_record(void *event, void *stream);
int dfrg_event_sync(void *event);
int dfrg_event_elapsed_ms(void *start, void *stop, float *ms);
/* Capture of enqueued work into a hipGraph, replayed with one launch.  No counterpart in the
 * reference: it takes the host's per-launch cost out of repeated small launches.  Between
 * dfrg_graph_begin and dfrg_graph_end the calling thread's work on `stream` (NULL: the thread's
 * stream) is recorded instead of run.  dfrg_threshold_batch, dfrg_ft_batch, dfrg_se_batch_run and
 * dfrg_se_batch_run_noalloc enqueue kernels only (no memsets, no synchronisation), so any
 * sequence of them can be captured; calls that synchronise (the flow server, table updates)
 * cannot. */
typedef struct dfrg_graph dfrg_graph;
int dfrg_graph_begin(void *stream);
int dfrg_graph_end(void *stream, dfrg_graph **graph);
int dfrg_graph_launch(dfrg_graph *graph, void *stream);
int dfrg_graph_destroy(dfrg_graph *graph);
/* Stream-ordered caching pool for per-call scratch (partials, counters, result slots) used by
 * the C++ integrator templates; replaces the per-integrator rmm pools
 * (integrator_4D_2ang_gpu.hh:86-90).  Memory returned by dfrg_pool_alloc is zeroed once at
 * allocation; a returned block may be reused by any later call. */
int dfrg_pool_alloc(void **d_ptr, size_t bytes);
int dfrg_pool_free(void *d_ptr);
int dfrg_pool_host_alloc(void **h_ptr, size_t bytes);
int dfrg_pool_host_free(void *h_ptr);

/* ------------------------------------------------------------------------------------------ */
/* quadrature (QuadratureProvider, quadrature_provider.hh:138-248)                             */
/* ------------------------------------------------------------------------------------------ */
/* Gauss-Legendre nodes/weights on [0,1], ascending, generated in FP64 on the GPU
 * (Newton on the Legendre recurrence, one root per lane).  Replaces GSL's
 * gsl_integration_fixed_legendre call (quadrature.cu:23-58). */
int dfrg_gl_generate(int order, double *d_x, double *d_w, void *stream);
/* Matsubara rule of MatsubaraQuadrature<double>::reinit(T, E, step, min_size, max_size,
 * vacuum_quad_size, precision_factor) (matsubara.cu:33-85, 128-161), generated on the GPU:
 * Monien's Jacobi matrix, eigenvalues by parallel Sturm bisection, weights from the first
 * eigenvector component of a twisted LDL^T/UDU^T factorisation.  Writes at most cap nodes;
 * *size receives the rule size and *T_rule the rule temperature (0 for the vacuum rule).
 * Synchronises `stream`. */
int dfrg_matsubara_generate(double T, double E, int step, int min_size, int max_size, int vacuum_quad_size,
                            int precision_factor, double *d_x, double *d_w, int cap, int *size, double *T_rule,
                            void *stream);

typedef struct dfrg_qp dfrg_qp;
/* QuadratureProvider() uses (48, 32, 0); QuadratureProvider(json) reads
 * /integration/{vacuum_quad_size=48, min_matsubara_size=16, add_matsubara_size=0}
 * (quadrature_provider.cu:213-227).  Invalid values give DFRG_ERR_INVALID like
 * MatsubaraStorage::set_* (quadrature_provider.cu:111-126). */
int dfrg_qp_create(dfrg_qp **qp, int vacuum_quad_size, int min_matsubara_size, int add_matsubara_size);
int dfrg_qp_destroy(dfrg_qp *qp);
/* Cached GL rule of `order`; fp32 != 0 returns FP32 copies.  Outputs may be NULL. */
int dfrg_qp_gl(dfrg_qp *qp, int order, int fp32, const void **d_x, const void **d_w, const void **h_x,
               const void **h_w);
/* Cached Matsubara rule for (T, E), keys matched to 1e-6 relative (quadrature_provider.cu:31-60). */
int dfrg_qp_matsubara(dfrg_qp *qp, double T, double E, int fp32, int *size, double *T_rule, const void **d_x,
                      const void **d_w, const void **h_x, const void **h_w);

/* ------------------------------------------------------------------------------------------ */
/* interpolation tables (LinearInterpolator1D<double, LogarithmicCoordinates1D<double>>,       */
/* linear_interpolation_1D.hh:22-113, coordinates.hh:141-196)                                   */
/* ------------------------------------------------------------------------------------------ */
/* n FP64 values on y(i) = b expm1(bias i/(n-1)) + start, b = (stop-start)/expm1(bias), held in
 * HBM.  Evaluation is the template's: idx = log1p((y-start)/b)(n-1)/bias clamped to [0, n-1]
 * (fmax/fmin: a NaN index gives the last value), linear between floor and ceil.  Errors:
 * n < 2, stop <= start or bias == 0 give DFRG_ERR_INVALID. */
typedef struct dfrg_table dfrg_table;
int dfrg_table_create(dfrg_table **t, int n, double start, double stop, double bias, const double *values);
/* copy n new host values (the reference's update(); enqueued on `stream`) */
int dfrg_table_update(dfrg_table *t, const double *values, void *stream);
/* device pointer to the n values (a non-owning view, as the reference's copied interpolator) */
int dfrg_table_view(dfrg_table *t, const double **d_values, int *n);
/* d_out[i] = table(d_y[i]) for n_points device values */
int dfrg_table_eval(dfrg_table *t, const double *d_y, int n_points, double *d_out, void *stream);
int dfrg_table_destroy(dfrg_table *t);

/* ------------------------------------------------------------------------------------------ */
/* batched integrals with the built-in integrands                                             */
/* ------------------------------------------------------------------------------------------ */
/* LPA threshold integrals, radial d-dim form of IntegratorGPU<d> (integrator_gpu.hh:34-60):
 *   out[i] = sum_j S_d q_j^{d-2}/2 k^2/(2pi)^d w_j x_ext RBdot(k^2,q_j^2)/(q_j^2+RB(k^2,q_j^2)+m2[i])^2,
 * q_j = k sqrt(x_j x_ext), exponential regulator with exponent 1.  One wave per integral. */
int dfrg_threshold_batch(int d, const double *d_m2, int n_int, const double *d_x, const double *d_w, int nx,
                         double x_extent, double k, int mode, double *d_out, void *stream);

/* Two-angle self-energy batch (Integrator4D2AngGPU form, integrator_4D_2ang_gpu.hh:20-45):
 *   I_i = 2pi/(2pi)^4 sum_{q,c1,c2} q^2/2 k^2 sqrt(1-c1^2) (2 w_c1)(2 w_c2)(w_q x_ext) f,
 *   f = RBdot(q) G(q)^2 G(|p_i - q|),  G(y) = 1/(Z_i(y) y^2 + RB(k^2,y^2)),
 * each integral with its own radial order nq[i], angular order na[i] (both angles), external
 * momentum p[i] and dressing table Z_i (tables[tab_off[i] .. + tab_n[i]]) on
 * LogarithmicCoordinates1D(tab_n[i], start[i], stop[i], bias[i]).  Radial orders 1..1024, angular
 * orders powers of two in [2, 64], tables of 2..256 values.  Integrals may share a table (the same
 * tab_off, then the same tab_n and coordinates); distinct tables must not overlap.  Host arrays;
 * copied to HBM once. */
typedef struct dfrg_se_batch dfrg_se_batch;
int dfrg_se_batch_create(dfrg_se_batch **b, dfrg_qp *qp, int n_int, const double *p, const int *nq,
                         const int *na, const int *tab_off, const int *tab_n, const double *start,
                         const double *stop, const double *bias, const double *tables, int tables_len,
                         double x_extent, double k);
int dfrg_se_batch_destroy(dfrg_se_batch *b);
/* Work shape: threads per block (64..1024, multiple of 64), target integrand evaluations per tile
 * and persistent grid size (0 = one block per tile).  Rebuilds the tile list. */
int dfrg_se_batch_set_shape(dfrg_se_batch *b, int block, int tile_evals, int grid);
int dfrg_se_batch_get_shape(dfrg_se_batch *b, int *block, int *tile_evals, int *grid, int *n_tiles);
/* Enqueue one evaluation of every integral: d_out[i] = I_i (n_int doubles in HBM). */
int dfrg_se_batch_run(dfrg_se_batch *b, int mode, double *d_out, void *stream);
/* Time shapes on this batch and keep the fastest (per mode): after ~4 ms of launches that let the
 * clocks settle, every shape with `reps` launches, then the four fastest again in three
 * interleaved passes of 4 reps launches (the best pass counts).  cache_path (may be NULL) is a
 * text file of earlier results keyed by device name and batch signature: a hit skips the
 * search; a miss appends the winner. */
int dfrg_se_batch_autotune(dfrg_se_batch *b, int mode, int reps, const char *cache_path, int *block,
                           int *tile_evals, int *grid, float *ms);
/* Capture `launches` back-to-back evaluations into a hipGraph (owned by the batch; recaptured
 * when mode/shape/out change) and launch it once on `stream`. */
int dfrg_se_batch_graph_launch(dfrg_se_batch *b, int mode, double *d_out, int launches, void *stream);
/* Evaluations (integrand calls) performed by one run of the batch. */
int dfrg_se_batch_evals(dfrg_se_batch *b, unsigned long long *evals);
/* Integrand form a run in `mode` uses for each integral (variant[0 .. n_int-1]): 0 FP64; 1 mixed;
 * 2 mixed for an integral whose table starts at >= 1e-30 and whose |p - q|^2 64 log2(e)/k^2
 * stays below 2^30, which rounds the exponent with one FMA and drops a guard (within the mixed
 * tolerance of form 1).  The choice depends on the integral alone, not on the batch. */
int dfrg_se_batch_variants(dfrg_se_batch *b, int mode, int *variant);
/* Radial unroll of the mixed path's paired evaluation loop at the current shape: 2 (two radial
 * nodes per step) when the launch puts at most 16 waves on a CU, else 1.  The plain path always
 * uses 1.  Results do not depend on it. */
int dfrg_se_batch_unroll(dfrg_se_batch *b, int *unroll);

/* dfrg_se_batch_run without any allocation.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The batch (inputs) must exist already and `stream` must not be NULL (the
 * NULL stream would create the calling thread's stream on first use).  Partial slots, ticket
 * counters and the work queue live in a static device array of DFRG_NOALLOC_SCRATCH_BYTES per GPU
 * (held until the process ends), so a batch may have at most DFRG_NOALLOC_MAX_INTEGRALS integrals
 * and DFRG_NOALLOC_MAX_TILES tiles at its current shape (dfrg_se_batch_get_shape); larger ones give
 * DFRG_ERR_INVALID and launch nothing.  Each call clears the counters and queue it uses with a fill
 * kernel before its launch.  Calls on one GPU share that GPU's array: calls that overlap on one
 * GPU (different streams, or different host threads) race and must be serialised by the caller.
 * Results are bitwise those of dfrg_se_batch_run. */
#define DFRG_NOALLOC_MAX_INTEGRALS 4096
#define DFRG_NOALLOC_MAX_TILES 16384
#define DFRG_NOALLOC_SCRATCH_BYTES (256 + 128 * DFRG_NOALLOC_MAX_INTEGRALS + 128 * DFRG_NOALLOC_MAX_TILES)
int dfrg_se_batch_run_noalloc(dfrg_se_batch *b, int mode, double *d_out, void *stream);

/* Finite-T two-propagator batch (IntegratorFiniteTq0GPU form, integrator_finiteTq0_gpu.hh:20-44):
 *   I_i = sum_{j,n} el(q_j) wq_j [ wm_n (f(q_j,wn) + f(q_j,-wn)) ] + T_i sum_j el(q_j) wq_j f(q_j,0),
 *   f(q,q0) = RBdot(q^2) G(q0,q^2,m1_i)^2 G(q0+p0_i, q^2+pabs_i^2, m2_i),  G(w,s,m) = 1/(w^2+s+RB(k^2,s)+m).
 * radial_mode 0: q = k sqrt(x extent), el = S_{d-1} q^{d-3}/2 k^2/(2pi)^{d-1}, wq = w extent
 *                (reference x = q^2/k^2 mapping, integrator_finiteTq0_gpu.hh:164-170);
 * radial_mode 1: q = extent x, el = S_{d-1} q^{d-2}/(2pi)^{d-1}, wq = w extent (linear |q|).
 * Matsubara rule of integral i: d_mx/d_mw[m_off[i] .. + m_n[i]] (positive nodes). */
int dfrg_ft_batch(int d, int radial_mode, int n_int, const double *d_T, const double *d_p0, const double *d_pabs,
                  const double *d_m1, const double *d_m2, const double *d_mx, const double *d_mw, const int *d_m_off,
                  const int *d_m_n, const double *d_x, const double *d_w, int nx, double extent, double k, int mode,
                  double *d_out, void *stream);

/* Threads per integral of dfrg_ft_batch (64, 128, 256 or 512; 0 = automatic: about two waves
 * per SIMD over the batch).  Process-wide; replaces the max_block_size constructor argument of
 * IntegratorFiniteTq0GPU (integrator_finiteTq0_gpu.hh:53-68) for the batched form. */
int dfrg_ft_set_block(int block);

/* ------------------------------------------------------------------------------------------ */
/* LPA flow server: continuous batching of O(N) LPA flows with an on-device RK4 stepper.       */
/* No reference counterpart: the reference runs LPA flows as FE systems under SUNDIALS IDA     */
/* (Examples/ONfiniteT/CG.cc:18-51) with one IntegratorTBB<3> call per quadrature point; its   */
/* threshold integral is the V.kernel form (Examples/ONfiniteT/flows/V/V.kernel:4-15).         */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
  int n_grid;       /* uniform rho grid points (>= 4) */
  double rho_max;   /* rho in [0, rho_max] */
  double T;         /* temperature (> 0) */
  double mu;        /* chemical potential (recorded; the bosonic O(N) kernel has no mu coupling) */
  double N;         /* O(N) */
  double m2_0;      /* u(rho) = m2_0 + lambda/2 rho at k = Lambda = 1 */
  double lambda;
} dfrg_flow_desc;

typedef struct dfrg_fs dfrg_fs;
/* hbm_budget: bytes of HBM the resident flows may hold (admission control); max_resident: slot
 * limit; nx: GL nodes of each threshold integral (x_extent 20, exponential regulator b=1);
 * n_steps RK4 steps from k = 1 to k_end; steps_per_launch: RK4 steps per graph launch. */
int dfrg_fs_create(dfrg_fs **fs, dfrg_qp *qp, size_t hbm_budget, int max_resident, int nx, int n_steps,
                   double k_end, int steps_per_launch);
int dfrg_fs_destroy(dfrg_fs *fs);
/* Queue a flow; *id receives its handle.  Timestamped on arrival. */
int dfrg_fs_submit(dfrg_fs *fs, const dfrg_flow_desc *desc, long long *id);
/* Admit queued flows that fit, advance every resident flow by one launch of steps, retire the
 * finished ones.  *resident / *queued / *done (may be NULL) report the state afterwards. */
int dfrg_fs_progress(dfrg_fs *fs, int *resident, int *queued, int *done);
/* state: 0 queued, 1 resident, 2 finished, 3 failed (non-finite state; slot freed). */
int dfrg_fs_poll(dfrg_fs *fs, long long id, int *state);
/* Final u(rho) of a finished flow (n_grid doubles) and its latency in seconds (arrival to
 * completion).  Releases the host copy. */
int dfrg_fs_result(dfrg_fs *fs, long long id, double *u_out, int n, double *latency_s);
/* Aggregate statistics of retired flows: count, p50 and p99 latency (s), busy seconds. */
int dfrg_fs_stats(dfrg_fs *fs, int *finished, double *p50_s, double *p99_s, double *busy_s);
/* Bytes of HBM one flow of n_grid points occupies while resident. */
size_t dfrg_fs_flow_bytes(int n_grid);
/* Graph launches so far, their summed device time in seconds (HIP events around each launch on
 * the server's stream) and the grid-point RK4 steps they advanced (each resident flow counts the
 * steps it had left, up to steps_per_launch).  No reference counterpart (measurement). */
int dfrg_fs_timing(dfrg_fs *fs, long long *launches, double *device_s, long long *point_steps);

#ifdef __cplusplus
}
#endif
#endif
