// peak.hip -- FMA-chain probe of the vector ALU's FMA rate on the current device, the measured
// counterpart of the published FP64/FP32 vector peaks that the integrands' roofline is quoted
// against.  Every lane runs kChains independent FMA chains a <- a b + c (b < 1, so a converges to
// c / (1 - b) and stays normal), eight waves per SIMD, so neither dependency latency nor occupancy
// limits the issue rate; FLOP = 2 per FMA per lane.  FP32 chains are plain (not packed) FMAs, as
// the library's own FP32 code is built.
#include "internal.hh"

#include <algorithm>

using namespace dfrg;

namespace
{
  constexpr int kChains = 8;
  constexpr int kBlock = 256;

  template <typename F> __global__ __launch_bounds__(kBlock) void fma_chain_kernel(F b, F c, int iters, F *out)
  {
    F a[kChains];
#pragma unroll
    for (int i = 0; i < kChains; ++i) a[i] = (F)(threadIdx.x + i) * (F)1e-3;
#pragma unroll 8
    for (int it = 0; it < iters; ++it) {
#pragma unroll
      for (int i = 0; i < kChains; ++i) a[i] = fma(a[i], b, c);
    }
    F s = 0;
#pragma unroll
    for (int i = 0; i < kChains; ++i) s += a[i];
    out[(size_t)blockIdx.x * kBlock + threadIdx.x] = s; // keeps the chains live
  }

  template <typename F> hipError_t launch_chain(int grid, int iters, void *out, hipStream_t s)
  {
    hipLaunchKernelGGL(fma_chain_kernel<F>, dim3(grid), dim3(kBlock), 0, s, (F)0.999999, (F)1e-6, iters,
                       (F *)out);
    return hipGetLastError();
  }
} // namespace

int dfrg_valu_peak(int fp64, double *flop_per_s)
{
  DFRG_TRY(require_init());
  DFRG_CHECK_ARG(flop_per_s != nullptr, "NULL output");
  DFRG_CHECK_ARG(fp64 == 0 || fp64 == 1, "fp64 must be 0 or 1");
  const int cus = context().cu_count > 0 ? context().cu_count : 256;
  const int grid = 8 * cus;  // 8 workgroups of 4 waves per CU: 8 waves per SIMD
  const int iters = 8192;    // about 1 ms per launch at the published rate
  void *out = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float best = 0.0f;
  hipError_t e = hipMalloc(&out, sizeof(double) * (size_t)grid * kBlock);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  for (int rep = 0; rep < 4 && e == hipSuccess; ++rep) { // launch 0 warms the clocks up
    e = hipEventRecord(e0, s);
    if (e == hipSuccess) e = fp64 ? launch_chain<double>(grid, iters, out, s) : launch_chain<float>(grid, iters, out, s);
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess && rep > 0) best = best == 0.0f ? ms : std::min(best, ms);
  }
  if (e1) hipEventDestroy(e1);
  if (e0) hipEventDestroy(e0);
  if (s) hipStreamDestroy(s);
  if (out) hipFree(out);
  if (e != hipSuccess) {
    set_error("FMA-chain probe failed: %s", hipGetErrorString(e));
    return DFRG_ERR_HIP;
  }
  const double flop = 2.0 * kChains * (double)iters * (double)grid * kBlock;
  *flop_per_s = flop / (best * 1e-3);
  return DFRG_OK;
}
