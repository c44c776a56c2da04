// dfrg/check.hh -- the C ABI of libdfrg.so and its error translation for the C++ headers.
// check_hip() is the analogue of the reference's check_cuda() (DiFfRG/src/DiFfRG/common/
// cuda_prefix.cu:9-14): it throws std::runtime_error on a pending device error.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "dfrg.h"

namespace dfrg
{
  // DFRG_ERR_INVALID -> std::invalid_argument, every other failure -> std::runtime_error
  inline void check(int rc, const char *what)
  {
    if (rc == DFRG_OK) return;
    const std::string msg = std::string(what) + ": " + dfrg_last_error();
    if (rc == DFRG_ERR_INVALID) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
  }

  inline void check_hip(const std::string &prefix = "")
  {
    const hipError_t err = hipPeekAtLastError();
    if (err != hipSuccess) throw std::runtime_error(prefix + " HIP error: " + hipGetErrorString(err));
  }

  // Selects device 0 on first use; throws when no GPU is visible.
  inline void ensure_device()
  {
    const int rc = dfrg_init(0);
    if (rc == DFRG_ERR_NO_DEVICE) throw std::runtime_error("No HIP devices found!");
    check(rc, "dfrg_init");
  }
} // namespace dfrg
