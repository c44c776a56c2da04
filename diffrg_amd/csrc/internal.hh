// internal.hh -- shared host-side plumbing of libdfrg.so: error state, HIP status checks,
// device context.  Not installed; the public ABI is include/dfrg.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <string>

#include "dfrg.h"

namespace dfrg
{
  // thread-local message of the last failure (dfrg_last_error)
  void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
  void clear_error();

  // Library-wide state of the selected device.
  struct Context {
    bool initialised = false;
    int device = -1;
    int cu_count = 0;
    int clock_khz = 0;
    unsigned long long hbm_bytes = 0;
    char name[256] = {0};
    std::mutex mutex;
  };
  Context &context();

  // Returns DFRG_ERR_NOT_INIT (with message) unless dfrg_init succeeded; makes the library
  // device current on the calling thread.
  int require_init();

  // Stream of the calling thread when `stream` is NULL.
  hipStream_t resolve_stream(void *stream);

  // Device address of the current device's static scratch (DFRG_NOALLOC_SCRATCH_BYTES), resolved
  // in dfrg_init; nullptr before it.
  void *noalloc_scratch();
} // namespace dfrg

#define DFRG_HIP(call)                                                                                         \
  do {                                                                                                         \
    hipError_t dfrg_e_ = (call);                                                                               \
    if (dfrg_e_ != hipSuccess) {                                                                               \
      ::dfrg::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(dfrg_e_), __FILE__, __LINE__);      \
      return DFRG_ERR_HIP;                                                                                     \
    }                                                                                                          \
  } while (0)

#define DFRG_CHECK_ARG(cond, ...)                                                                              \
  do {                                                                                                         \
    if (!(cond)) {                                                                                             \
      ::dfrg::set_error(__VA_ARGS__);                                                                          \
      return DFRG_ERR_INVALID;                                                                                 \
    }                                                                                                          \
  } while (0)

#define DFRG_TRY(expr)                                                                                         \
  do {                                                                                                         \
    int dfrg_rc_ = (expr);                                                                                     \
    if (dfrg_rc_ != DFRG_OK) return dfrg_rc_;                                                                  \
  } while (0)

// Launch-time check of a kernel just enqueued (check_cuda analogue, cuda_prefix.cu:9-14).
#define DFRG_LAUNCHED(name)                                                                                    \
  do {                                                                                                         \
    hipError_t dfrg_e_ = hipGetLastError();                                                                    \
    if (dfrg_e_ != hipSuccess) {                                                                               \
      ::dfrg::set_error("launch of %s failed: %s", name, hipGetErrorString(dfrg_e_));                         \
      return DFRG_ERR_HIP;                                                                                     \
    }                                                                                                          \
  } while (0)
