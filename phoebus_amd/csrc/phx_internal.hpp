// phx_internal.hpp -- the library context (opaque phx_ctx of include/phx.h).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "phx.h"

#include "phx_device.hpp"

namespace phx {

inline BlockGeom make_block_geom(const phx_block *b) {
  BlockGeom g;
  g.ndim = b->ndim;
  g.ng = b->ng;
  int n[3];
  for (int d = 0; d < 3; ++d) {
    const bool act = d < b->ndim;
    n[d] = b->nx[d] + (act ? 2 * b->ng : 0);
    g.istart[d] = act ? b->ng : 0;
    g.is[d] = g.istart[d];
    g.ie[d] = g.istart[d] + b->nx[d] - 1;
    g.xmin[d] = b->xmin[d];
    g.dx[d] = b->dx[d];
  }
  g.ni = n[0];
  g.nj = n[1];
  g.nk = n[2];
  g.ncell = (long)n[0] * n[1] * n[2];
  return g;
}

// flux launches (phx_flux.hip): direction d of fluid.cpp:753-770 with the fused
// reconstruction; writes flux[d][m][cells] and fsig[d][cells]
void launch_fluxes(int d, int recon, int riemann, int mhd, int geom, const DevParams &dp,
                   const BlockGeom &bg, const double *prim, double *flux, double *fsig,
                   hipStream_t s);

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
