// tables.hip -- interpolation tables behind the C ABI: the data of a
// LinearInterpolator1D<double, LogarithmicCoordinates1D<double>> (linear_interpolation_1D.hh:22-113,
// coordinates.hh:141-196) held in HBM, for callers that bind the library through an FFI and cannot
// instantiate the C++ templates.  A table is n FP64 values on the coordinates
//   y(i) = b expm1(bias i/(n-1)) + start,  b = (stop - start)/expm1(bias),
// and evaluates with the reference's index: idx = log1p((y - start)/b) (n-1)/bias, clamped to [0, n-1]
// with fmax/fmin (a NaN index gives the last value), linear between floor and ceil.
#include "internal.hh"

#include <cmath>

using namespace dfrg;

struct dfrg_table {
  int n = 0;
  double start = 0, stop = 0, bias = 0, inv_b = 0, scale = 0;
  double *d_values = nullptr;
};

namespace
{
  __global__ void table_eval_kernel(const double *__restrict__ v, int n, double c, double inv_b, double scale,
                                    const double *__restrict__ y, int n_points, double *__restrict__ out)
  {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_points) return;
    double idx = log1p((y[i] - c) * inv_b) * scale;
    idx = fmax(0.0, fmin(idx, (double)(n - 1)));
    const double fl = floor(idx);
    out[i] = v[(int)fl] * (1.0 - idx + fl) + v[(int)ceil(idx)] * (idx - fl);
  }
} // namespace

extern "C" {

int dfrg_table_create(dfrg_table **t, int n, double start, double stop, double bias, const double *values)
{
  DFRG_TRY(require_init());
  DFRG_CHECK_ARG(t != nullptr && values != nullptr, "NULL argument");
  DFRG_CHECK_ARG(n >= 2, "a table needs at least 2 values (got %d)", n);
  DFRG_CHECK_ARG(std::isfinite(start) && std::isfinite(stop) && stop > start, "invalid range [%g, %g]", start, stop);
  DFRG_CHECK_ARG(std::isfinite(bias) && bias != 0.0, "invalid bias %g", bias);
  dfrg_table *r = new dfrg_table;
  r->n = n;
  r->start = start;
  r->stop = stop;
  r->bias = bias;
  r->inv_b = std::expm1(bias) / (stop - start);
  r->scale = (n - 1.0) / bias;
  if (hipMalloc((void **)&r->d_values, sizeof(double) * n) != hipSuccess) {
    delete r;
    set_error("dfrg_table_create: out of device memory (%d values)", n);
    return DFRG_ERR_HIP;
  }
  if (hipMemcpy(r->d_values, values, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess) {
    hipFree(r->d_values);
    delete r;
    set_error("dfrg_table_create: copy failed");
    return DFRG_ERR_HIP;
  }
  *t = r;
  return DFRG_OK;
}

int dfrg_table_update(dfrg_table *t, const double *values, void *stream)
{
  DFRG_TRY(require_init());
  DFRG_CHECK_ARG(t != nullptr && values != nullptr, "NULL argument");
  DFRG_HIP(hipMemcpyAsync(t->d_values, values, sizeof(double) * t->n, hipMemcpyHostToDevice,
                          resolve_stream(stream)));
  return DFRG_OK;
}

int dfrg_table_view(dfrg_table *t, const double **d_values, int *n)
{
  DFRG_CHECK_ARG(t != nullptr, "NULL table");
  if (d_values) *d_values = t->d_values;
  if (n) *n = t->n;
  return DFRG_OK;
}

int dfrg_table_eval(dfrg_table *t, const double *d_y, int n_points, double *d_out, void *stream)
{
  DFRG_TRY(require_init());
  DFRG_CHECK_ARG(t != nullptr, "NULL table");
  DFRG_CHECK_ARG(n_points >= 0, "n_points must be >= 0");
  if (n_points == 0) return DFRG_OK;
  DFRG_CHECK_ARG(d_y && d_out, "NULL device pointer");
  const int block = 256;
  hipLaunchKernelGGL(table_eval_kernel, dim3((n_points + block - 1) / block), dim3(block), 0, resolve_stream(stream),
                     t->d_values, t->n, t->start, t->inv_b, t->scale, d_y, n_points, d_out);
  DFRG_LAUNCHED("table_eval_kernel");
  return DFRG_OK;
}

int dfrg_table_destroy(dfrg_table *t)
{
  if (!t) return DFRG_OK;
  hipFree(t->d_values);
  delete t;
  return DFRG_OK;
}

} // extern "C"
