// phx_flux_shapes.hpp -- launch geometry of the flux sweeps (phx_flux_sweeps.hpp): the
// workgroup counts and LDS sizes of the x sweep, the y/z tiles and the y/z marching waves.
#pragma once

#include "phx_device.hpp"

namespace phx {

template <int RECON>
struct ReconWidth {
  static constexpr int H = RECON == PHX_RECON_CONSTANT ? 0 : RECON == PHX_RECON_LINEAR ? 1 : 2;
};

// ---- x: one lane per cell of full rows, stencils and upper edges staged in LDS ----------
// A block of XB lanes covers XB consecutive cells of the rows it spans; lanes H+1 ..
// XB-H-1 emit the face at their cell's lower edge, so consecutive blocks overlap by 2H+1
// cells.  Each primitive is read from HBM once per lane into LDS, reconstructed from LDS,
// and the cell's upper edge is published in LDS for the next lane.
constexpr int XB = 256;

// number of XB-lane groups the x sweep of block b takes
__host__ __device__ inline long flux_x_groups(const BlockGeom &b, int H) {
  const int dj = b.ndim > 1, dk = b.ndim > 2;
  const long total = (long)b.ni * (b.ie[1] - b.is[1] + 1 + 2 * dj) * (b.ie[2] - b.is[2] + 1 + 2 * dk);
  const long stride = XB - 2 * H - 1;
  return (total - 2 * H - 1 + stride - 1) / stride;
}

// ---- y, z: LDS tiles of ti_n (i) x ts_n (own direction) cells ----------------------------
// Rows of ti_n consecutive i are read into LDS; each lane reconstructs its cell along the
// own direction from LDS and publishes the upper edge for the lane one row up.  Lanes in
// own-rows H+1 .. ts_n-H-1 emit faces; tiles along the own direction overlap by 2H+1
// cells.  The shape trades redundant halo work (short ts_n) against partial cache lines
// (narrow ti_n: a row segment of ti_n doubles uses ti_n/16 of a 128-byte line).  Tiles
// run when the autotuner (phx_autotune) finds one faster than the marching waves below
// (the default) or when a caller fixes one.  Blocks are renumbered XCD by XCD so
// neighbouring i chunks share an L2.
constexpr int kTileMaxThreads = 512;
constexpr int kTileMaxRows = 256;
constexpr int kTileMaxLdsBytes = 64 * 1024;  // two blocks per CU

// number of ti_n x ts_n tiles the sweep along DIR of block b takes
__host__ __device__ inline long flux_tile_groups(const BlockGeom &b, int dir, int ti_n,
                                                 int ts_n, int H) {
  const int T = dir == 1 ? 2 : 1;
  const int S = ts_n - 2 * H - 1;
  const long nic = (b.ie[0] - b.is[0] + 3 + ti_n - 1) / ti_n;
  const long nft = (b.ie[dir] - b.is[dir] + 3 + S - 1) / S;
  const long ntr = b.ie[T] - b.is[T] + 1 + 2 * (b.ndim > T);
  return nic * nft * ntr;
}

// ---- y, z marching: one wave per workgroup --------------------------------------------
// A wave holds ti columns (consecutive in the flattened face range of (i, transverse
// index), i fastest) by MR = 64 / ti rows along the sweep, and marches over one of nseg
// segments of its columns' faces, MR cells per step.  The primitives of the rows a step
// reconstructs (its MR cells and H on either side) sit in an LDS ring of MR + 2H rows, so a
// step reads only its MR new rows from memory.  Row r takes the upper edge of row r - 1 by
// a cross-lane shuffle and row 0 the previous step's last row from LDS: no cell is
// reconstructed twice except the one before each segment (the tiles above redo 2H + 1 of
// every ts rows).
constexpr int kMarchLanes = 64;
constexpr int kMarchMaxLdsBytes = 16 * 1024;

// LDS bytes of a marching workgroup: the ring [NP][MR + 2H][ti] and the carry row [NP][ti]
__host__ __device__ inline long march_lds_bytes(int np, int h, int ti) {
  return 8L * np * ti * (kMarchLanes / ti + 2 * h + 1);
}

// number of marching workgroups of the DIR sweep of block b
__host__ __device__ inline long flux_march_groups(const BlockGeom &b, int dir, int ti,
                                                  int nseg) {
  const int T = dir == 1 ? 2 : 1;
  const long ncol = (long)(b.ie[0] - b.is[0] + 3) * (b.ie[T] - b.is[T] + 1 + 2 * (b.ndim > T));
  return (ncol + ti - 1) / ti * nseg;
}

}  // namespace phx
