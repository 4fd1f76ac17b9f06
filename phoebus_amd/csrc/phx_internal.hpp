// phx_internal.hpp -- the library context (opaque phx_ctx of include/phx.h).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "phx.h"

namespace phx {
// partial minima of the CFL reduction; bounds its grid (grid-stride loop beyond)
constexpr int kDtPartials = 4096;
}  // namespace phx

struct phx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;  // default stream of the context
  double *dt_partial = nullptr;  // [kDtPartials]
  std::string last_error;

  hipStream_t stream_of(phx_stream s) const {
    return s ? reinterpret_cast<hipStream_t>(s) : stream;
  }
  int fail(int code, const std::string &msg) {
    last_error = msg;
    return code;
  }
  int check_launch() {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(PHX_ERR_DEVICE, hipGetErrorString(e));
    return PHX_OK;
  }
};
