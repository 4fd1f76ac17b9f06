// phx_cells.hpp -- per-cell bodies that several kernels share: the Lorentz factor, the
// robust con2prim of one cell (fluid.cpp:478-517) and p2c of one cell (fluid.cpp:359-449).
#pragma once

#include "phx_device.hpp"

namespace phx {

// cell t of the entire block (i fastest); false beyond it
__device__ __forceinline__ bool entire_cell(const BlockGeom &b, long t, int idx[3]) {
  if (t >= b.ncell) return false;
  idx[0] = (int)(t % b.ni);
  idx[1] = (int)((t / b.ni) % b.nj);
  idx[2] = (int)(t / ((long)b.ni * b.nj));
  return true;
}

__device__ __forceinline__ bool is_interior(const BlockGeom &b, const int idx[3]) {
  return idx[0] >= b.is[0] && idx[0] <= b.ie[0] && idx[1] >= b.is[1] && idx[1] <= b.ie[1] &&
         idx[2] >= b.is[2] && idx[2] <= b.ie[2];
}

// cell t of the ghost layer face-adjacent to the interior of a 3-D block (the cells the c2p
// fixup reads besides the interior): the x1 faces' layers, then x2's, then x3's, each
// inner side first; false beyond them
__device__ __forceinline__ long face_layer_cells(const BlockGeom &b) {
  const long n0 = b.ie[0] - b.is[0] + 1, n1 = b.ie[1] - b.is[1] + 1, n2 = b.ie[2] - b.is[2] + 1;
  return 2 * (n1 * n2 + n0 * n2 + n0 * n1);
}
__device__ __forceinline__ bool face_layer_cell(const BlockGeom &b, long t, int idx[3]) {
  const int n0 = b.ie[0] - b.is[0] + 1, n1 = b.ie[1] - b.is[1] + 1, n2 = b.ie[2] - b.is[2] + 1;
  // face d spans the other two dimensions (e0 fastest, e1); one select per index keeps
  // idx in registers
  const long f0 = (long)n1 * n2, f1 = (long)n0 * n2, f2 = (long)n0 * n1;
  const int d = t < 2 * f0 ? 0 : t < 2 * (f0 + f1) ? 1 : 2;
  t -= d == 0 ? 0 : d == 1 ? 2 * f0 : 2 * (f0 + f1);
  const long face = d == 0 ? f0 : d == 1 ? f1 : f2;
  const int ne0 = d == 0 ? n1 : n0;
  const int side = t >= face;
  const long r = t - side * face;
  const int a0 = (int)(r % ne0), a1 = (int)(r / ne0);
  const int lo = d == 0 ? b.is[0] : d == 1 ? b.is[1] : b.is[2];
  const int hi = d == 0 ? b.ie[0] : d == 1 ? b.ie[1] : b.ie[2];
  const int own = side ? hi + 1 : lo - 1;
  if (t >= 2 * face) return false;  // beyond the layer: idx untouched
  idx[0] = d == 0 ? own : a0 + b.is[0];
  idx[1] = d == 1 ? own : (d == 0 ? a0 : a1) + b.is[1];
  idx[2] = d == 2 ? own : a1 + b.is[2];
  return true;
}

// ghost cell t of a block (every cell outside the interior box): the ghost planes of x3,
// then for interior k the ghost rows of x2, then for interior (j, k) the ghost cells of x1
__device__ __forceinline__ long ghost_cells(const BlockGeom &b) {
  return b.ncell - (long)(b.ie[0] - b.is[0] + 1) * (b.ie[1] - b.is[1] + 1) * (b.ie[2] - b.is[2] + 1);
}
__device__ __forceinline__ bool ghost_cell(const BlockGeom &b, long t, int idx[3]) {
  const int n0 = b.ie[0] - b.is[0] + 1, n1 = b.ie[1] - b.is[1] + 1, n2 = b.ie[2] - b.is[2] + 1;
  const int g0 = b.ni - n0, g1 = b.nj - n1, g2 = b.nk - n2;  // ghost counts per dimension
  const long A = (long)g2 * b.nj * b.ni, Bc = (long)n2 * g1 * b.ni, Cc = (long)n2 * n1 * g0;
  int i, j, k;
  if (t < A) {
    const long pl = (long)b.nj * b.ni;
    const int q = (int)(t / pl);
    k = q < b.is[2] ? q : b.ie[2] + 1 + (q - b.is[2]);
    j = (int)((t % pl) / b.ni);
    i = (int)(t % b.ni);
  } else if (t < A + Bc) {
    const long u = t - A, row = (long)g1 * b.ni;
    k = b.is[2] + (int)(u / row);
    const int q = (int)((u % row) / b.ni);
    j = q < b.is[1] ? q : b.ie[1] + 1 + (q - b.is[1]);
    i = (int)(u % b.ni);
  } else if (t < A + Bc + Cc) {
    const long u = t - A - Bc, pl = (long)n1 * g0;
    k = b.is[2] + (int)(u / pl);
    j = b.is[1] + (int)((u % pl) / g0);
    const int q = (int)(u % g0);
    i = q < b.is[0] ? q : b.ie[0] + 1 + (q - b.is[0]);
  } else {
    return false;
  }
  idx[0] = i;
  idx[1] = j;
  idx[2] = k;
  return true;
}

// work-list kinds of the pack cell kernels: 4 interior, 5 entire blocks, 6 face layer,
// 7 ghost cells
// (a compile-time kind: a runtime choice between the three keeps idx in scratch memory)
template <int KIND>
__device__ __forceinline__ bool cell_of_kind(const BlockGeom &b, long t, int idx[3]) {
  if constexpr (KIND == 5) return entire_cell(b, t, idx);
  if constexpr (KIND == 6) return face_layer_cell(b, t, idx);
  if constexpr (KIND == 7) return ghost_cell(b, t, idx);
  const int ni = b.ie[0] - b.is[0] + 1, nj = b.ie[1] - b.is[1] + 1, nk = b.ie[2] - b.is[2] + 1;
  if (t >= (long)ni * nj * nk) return false;
  idx[0] = (int)(t % ni) + b.is[0];
  idx[1] = (int)((t / ni) % nj) + b.is[1];
  idx[2] = (int)(t / ((long)ni * nj)) + b.is[2];
  return true;
}

// relativity_utils.hpp:54-59 (W from the primitive W v^i)
template <class M>
__device__ __forceinline__ double lorentz_factor(const M &g, const double vp[3]) {
  double vsq = 0.0;
#pragma unroll
  for (int ii = 0; ii < 3; ++ii)
#pragma unroll
    for (int jj = 0; jj < 3; ++jj)
      if (M::nz_cov(ii + 1, jj + 1)) vsq += g.gcov(ii + 1, jj + 1) * vp[ii] * vp[jj];
  return sqrt(1. + vsq);
}

// singularity-eos IdealGas BulkModulusFromDensityInternalEnergy (unit scales 1)
__device__ __forceinline__ double eos_bmod_rho_sie(const DevParams &p, double rho, double sie) {
  return (p.gm1 + 1.0) * p.gm1 * rho * sie;
}

// con2prim of cell c = (i, j, k) of block b, in place on cons / prim / temp / mu / csig;
// writes the fail flag and the iteration count
template <bool MHD, int GEOM, bool CACHED = false, class CP, class PP>
__device__ __forceinline__ void con2prim_cell(const DevParams &p, const BlockGeom &b, int i,
                                              int j, int k, const CP &cons, const PP &prim,
                                              double *__restrict__ temp,
                                              double *__restrict__ mu,
                                              double *__restrict__ csig,
                                              double *__restrict__ fail,
                                              int32_t *__restrict__ iters) {
  using L = Layout<MHD>;
  const long c = ((long)k * b.nj + j) * b.ni + i;
  const long n = b.ncell;
  double cs[L::ncons], sig[3];
#pragma unroll
  for (int v = 0; v < L::ncons; ++v) cs[v] = cons[v][c];
  // zones rejected up front by the NaN test keep every field (con2prim_robust.hpp:303-304)
  if (isnan(cs[0]) + isnan(cs[1]) + isnan(cs[4]) > 0) {
    fail[c] = 0.0;
    if (iters) iters[c] = 0;
    return;
  }
  const double X[3] = {xc(b, 0, i), xc(b, 1, j), xc(b, 2, k)};
  int it = 0;
  bool solved;
  const double st = con2prim_fields<MHD, GEOM, CACHED>(p, b, (long)j * b.ni + i, X, cs, prim,
                                                       temp, mu, cons, c, sig, it, solved);
#pragma unroll
  for (int d = 0; d < 3; ++d)
    if (solved && d < b.ndim) csig[d * n + c] = sig[d];
  fail[c] = st;
  if (iters) iters[c] = it;
}

// p2c of cell (i, j, k): cons (with B) and, when csig is given, the cell signal speeds
template <bool MHD, int GEOM, bool CACHED = false, class PP, class CP>
__device__ __forceinline__ void prim_to_con_cell(const DevParams &p, const BlockGeom &b, int i,
                                                 int j, int k, const PP &prim, const CP &cons,
                                                 double *__restrict__ csig) {
  using M = typename MetricOf<GEOM>::type;
  using L = Layout<MHD>;
  const long c = ((long)k * b.nj + j) * b.ni + i;
  const long n = b.ncell;
  const M g =
      metric_at<GEOM, CACHED>(p, b, 0, (long)j * b.ni + i, xc(b, 0, i), xc(b, 1, j), xc(b, 2, k));
  double pr[L::nprim];
#pragma unroll
  for (int v = 0; v < L::nprim; ++v) pr[v] = prim[v][c];
  const double vp[3] = {pr[1], pr[2], pr[3]};
  const double bp[3] = {MHD ? pr[5] : 0.0, MHD ? pr[6] : 0.0, MHD ? pr[7] : 0.0};
  double D, S[3], bc[3], tau, sig[3];
  p2c<MHD>(pr[0], vp, bp, pr[4], pr[L::prs], pr[L::gm1], g, D, S, bc, tau, sig);
  cons[0][c] = D;
  cons[1][c] = S[0];
  cons[2][c] = S[1];
  cons[3][c] = S[2];
  cons[4][c] = tau;
  if constexpr (MHD) {
    cons[5][c] = bc[0];
    cons[6][c] = bc[1];
    cons[7][c] = bc[2];
  }
  if (csig)
    for (int d = 0; d < b.ndim; ++d) csig[d * n + c] = sig[d];
}

}  // namespace phx
