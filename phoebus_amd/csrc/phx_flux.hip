// phx_flux.hip -- fused reconstruction + Riemann flux kernels (fluid.cpp:676-770) for gfx950.
//
// The reference reconstructs every cell once per direction into ql/qr arrays and a second
// kernel reads them back.  Here each cell is reconstructed once per direction in
// registers and the face states never reach memory:
//
//  * x (unit stride): one lane per cell along a row; primitives are staged through LDS
//    once per variable and the cell's upper-edge state is published in LDS, so lane l forms
//    face i from the upper edge of cell i-1 (lane l-1) and the lower edge of cell i (its
//    own).  Blocks overlap by 2H+1 cells.
//  * y, z (strided): one 64-lane wave of a few columns marches along the sweep with an LDS
//    ring of primitive rows, passing upper edges to the next row by cross-lane shuffles;
//    or LDS tiles of a few consecutive i by a stretch of the own direction, overlapping by
//    2H+1 rows.  The autotuner picks the kind and shape per block size.
//
// Results are bit-identical to reconstructing per face: the same arithmetic runs on the
// same operands.
//
// This file builds the sweeps with the exact arithmetic and holds what both arithmetics
// share: the sweep shapes and the dispatch by phx_fluid_params::flux_arith.
#include <hip/hip_runtime.h>

#include "phx.h"
#include "phx_device.hpp"
#include "phx_internal.hpp"

#define PHX_FLUX_NS flux_exact
#define PHX_FLUX_DISPATCH
#include "phx_flux_sweeps.hpp"

namespace phx {

// ---- dispatch --------------------------------------------------------------------------
TileShape resolve_tile_shape(int d, int recon, int mhd, const BlockGeom &bg, TileShape want) {
  const int H = recon == PHX_RECON_CONSTANT ? 0 : recon == PHX_RECON_LINEAR ? 1 : 2;
  const int NP = mhd ? 10 : 7;
  if (want.ts == 0) {
    // the default: marching waves of 8 columns x 8 rows in segments of at least 24 faces
    // when that fills the chip with at least 1024 waves; a small block's march would be a
    // few waves stepping serially, so it keeps 8-wide tiles of up to 64 rows
    const int nf = bg.ie[d] - bg.is[d] + 3;
    const int nseg = nf / 24 < 1 ? 1 : nf / 24 > 8 ? 8 : nf / 24;
    const TileShape m =
        resolve_tile_shape(d, recon, mhd, bg, TileShape{want.ti > 0 ? want.ti : 8, -nseg});
    if (flux_march_groups(bg, d, m.ti, -m.ts) >= 1024) return m;
    want.ts = 64;
  }
  if (want.ts < 0) {  // marching: ti a power of two up to 64, -ts segments of the faces
    const int want_ti = want.ti > 0 ? want.ti : 8;
    int ti = 1;
    while (ti * 2 <= want_ti && ti < kMarchLanes) ti *= 2;
    while (ti > 1 && march_lds_bytes(NP, H, ti) > kMarchMaxLdsBytes) ti /= 2;
    const int nf = bg.ie[d] - bg.is[d] + 3;
    return TileShape{ti, -(-want.ts < nf ? -want.ts : nf)};
  }
  // own-direction cells the face range needs: s-2-H .. e+1+H
  const int need = bg.ie[d] - bg.is[d] + 4 + 2 * H;
  const int cap = need <= kTileMaxRows ? need : kTileMaxRows;
  TileShape t = want;
  if (t.ti <= 0) t.ti = 8;
  if (t.ts == 0) t.ts = 64;
  if (t.ts > cap) t.ts = cap;
  if (t.ts < 2 * H + 2) t.ts = 2 * H + 2 <= cap ? 2 * H + 2 : cap;
  if (t.ti < 1) t.ti = 1;
  while (t.ti > 1 && (t.ti * t.ts > kTileMaxThreads ||
                      (long)NP * t.ti * t.ts * 8 > kTileMaxLdsBytes))
    t.ti /= 2;
  while (t.ti * t.ts > kTileMaxThreads || (long)NP * t.ti * t.ts * 8 > kTileMaxLdsBytes)
    t.ts /= 2;
  return t;
}

long pack_groups_flux(int d, int recon, const BlockGeom &bg, TileShape shape) {
  const int H = recon == PHX_RECON_CONSTANT ? 0 : recon == PHX_RECON_LINEAR ? 1 : 2;
  if (d == 0) return flux_x_groups(bg, H);
  if (shape.ts < 0) return flux_march_groups(bg, d, shape.ti, -shape.ts);
  return flux_tile_groups(bg, d, shape.ti, shape.ts, H);
}

void launch_fluxes(int d, int recon, int riemann, int mhd, int geom, const DevParams &dp,
                   const BlockGeom &bg, const PrimPtrs &prim, const FluxPtrs &flux, double *fsig,
                   hipStream_t s, TileShape sh, bool fast) {
  if (fast)
    flux_fast::launch_fluxes(d, recon, riemann, mhd, geom, dp, bg, prim, flux, fsig, s, sh);
  else
    flux_exact::launch_fluxes(d, recon, riemann, mhd, geom, dp, bg, prim, flux, fsig, s, sh);
}

void launch_pack_flux(int d, int recon, int riemann, int geom, const DevParams &dp,
                      const PackView &v, long groups, TileShape sh, hipStream_t s, bool fast) {
  if (fast)
    flux_fast::launch_pack_flux(d, recon, riemann, geom, dp, v, groups, sh, s);
  else
    flux_exact::launch_pack_flux(d, recon, riemann, geom, dp, v, groups, sh, s);
}

}  // namespace phx
