// phx_kernels.hip -- HIP kernels of the fluid stage for gfx950 and the C ABI (include/phx.h).
//
// Zone-per-lane kernels over SoA fields with the cell index i fastest, so that a
// wavefront's 64 lanes read 64 consecutive doubles of each variable.  Reconstruction is
// fused into the Riemann kernel (face states never reach HBM); the CFL reduction is a
// wavefront shuffle reduction followed by one small finishing kernel.  All FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cfloat>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

#include "phx.h"
#include "phx_device.hpp"
#include "phx_internal.hpp"
#include "phx_cells.hpp"

namespace phx {

// ---- FluxCT (fluid.cpp:775-841) --------------------------------------------------------
// corner EMFs of a 3-D block: thread t of (nx+1)(ny+1)(nz+1)
template <int NC, class FP>
__device__ __forceinline__ void ct_emf_3d_body(const BlockGeom &b, long t, const FP &flux,
                                               double *__restrict__ emf) {
  const int ni = b.ie[0] - b.is[0] + 2, nj = b.ie[1] - b.is[1] + 2, nk = b.ie[2] - b.is[2] + 2;
  if (t >= (long)ni * nj * nk) return;
  const int i = (int)(t % ni) + b.is[0];
  const int j = (int)((t / ni) % nj) + b.is[1];
  const int k = (int)(t / ((long)ni * nj)) + b.is[2];
  const long c = ((long)k * b.nj + j) * b.ni + i;
  const long sj = b.ni, sk = (long)b.ni * b.nj, n = b.ncell;
  auto F = [&](int d, int comp, long cc) { return flux[d * 8 + 5 + comp][cc]; };
  emf[c] = 0.25 * (F(1, 2, c) + F(1, 2, c - sk) - F(2, 1, c) - F(2, 1, c - sj));
  emf[n + c] = -0.25 * (F(0, 2, c) + F(0, 2, c - sk) - F(2, 0, c) - F(2, 0, c - 1));
  emf[2 * n + c] = 0.25 * (F(0, 1, c) + F(0, 1, c - sj) - F(1, 0, c) - F(1, 0, c - 1));
}

template <int NC>
__global__ __launch_bounds__(256) void k_ct_emf_3d(BlockGeom b, FluxPtrs flux,
                                                   double *__restrict__ emf) {
  ct_emf_3d_body<NC>(b, (long)blockIdx.x * blockDim.x + threadIdx.x, flux, emf);
}

// B fluxes from the corner EMFs of a 3-D block: thread t of (nx+1)(ny+1)(nz+1)
template <int NC, class FP>
__device__ __forceinline__ void ct_flux_3d_body(const BlockGeom &b, long t, const FP &flux,
                                                const double *__restrict__ emf) {
  const int ni = b.ie[0] - b.is[0] + 2, nj = b.ie[1] - b.is[1] + 2, nk = b.ie[2] - b.is[2] + 2;
  if (t >= (long)ni * nj * nk) return;
  const int i = (int)(t % ni) + b.is[0];
  const int j = (int)((t / ni) % nj) + b.is[1];
  const int k = (int)(t / ((long)ni * nj)) + b.is[2];
  const long c = ((long)k * b.nj + j) * b.ni + i;
  const long sj = b.ni, sk = (long)b.ni * b.nj, n = b.ncell;
  const double *e0 = emf, *e1 = emf + n, *e2 = emf + 2 * n;
  auto F = [&](int d, int comp) -> double & { return flux[d * 8 + 5 + comp][c]; };
  F(0, 0) = 0.0;
  F(0, 1) = 0.5 * (e2[c] + e2[c + sj]);
  F(0, 2) = -0.5 * (e1[c] + e1[c + sk]);
  F(1, 0) = -0.5 * (e2[c] + e2[c + 1]);
  F(1, 1) = 0.0;
  F(1, 2) = 0.5 * (e0[c] + e0[c + sk]);
  F(2, 0) = 0.5 * (e1[c] + e1[c + 1]);
  F(2, 1) = -0.5 * (e0[c] + e0[c + sj]);
  F(2, 2) = 0.0;
}

template <int NC>
__global__ __launch_bounds__(256) void k_ct_flux_3d(BlockGeom b, FluxPtrs flux,
                                                    const double *__restrict__ emf) {
  ct_flux_3d_body<NC>(b, (long)blockIdx.x * blockDim.x + threadIdx.x, flux, emf);
}

template <int NC>
__global__ __launch_bounds__(256) void k_ct_emf_2d(BlockGeom b, FluxPtrs flux,
                                                   double *__restrict__ emf) {
  const int ni = b.ie[0] - b.is[0] + 2, nj = b.ie[1] - b.is[1] + 2, nk = b.ie[2] - b.is[2] + 1;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)ni * nj * nk) return;
  const int i = (int)(t % ni) + b.is[0];
  const int j = (int)((t / ni) % nj) + b.is[1];
  const int k = (int)(t / ((long)ni * nj)) + b.is[2];
  const long c = ((long)k * b.nj + j) * b.ni + i;
  const long sj = b.ni;
  auto F = [&](int d, int comp, long cc) { return flux[d * 8 + 5 + comp][cc]; };
  emf[c] = 0.25 * (F(0, 1, c) + F(0, 1, c - sj) - F(1, 0, c) - F(1, 0, c - 1));
}

template <int NC>
__global__ __launch_bounds__(256) void k_ct_flux_2d(BlockGeom b, FluxPtrs flux,
                                                    const double *__restrict__ emf) {
  const int ni = b.ie[0] - b.is[0] + 2, nj = b.ie[1] - b.is[1] + 2, nk = b.ie[2] - b.is[2] + 1;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)ni * nj * nk) return;
  const int i = (int)(t % ni) + b.is[0];
  const int j = (int)((t / ni) % nj) + b.is[1];
  const int k = (int)(t / ((long)ni * nj)) + b.is[2];
  const long c = ((long)k * b.nj + j) * b.ni + i;
  const long sj = b.ni;
  flux[0 * 8 + 5][c] = 0.0;
  flux[0 * 8 + 6][c] = 0.5 * (emf[c] + emf[c + sj]);
  flux[1 * 8 + 5][c] = -0.5 * (emf[c] + emf[c + 1]);
  flux[1 * 8 + 6][c] = 0.0;
}

// ---- flux divergence + sources + RK average/update (phoebus_driver.cpp:560-576) -------
template <int NC>
__global__ __launch_bounds__(256) void k_update(BlockGeom b, FluxPtrs flux, ConsPtrs src,
                                                ConsPtrs u_stage, ConsPtrs u_base, double beta,
                                                double bdt, ConsPtrs u_out) {
  const int ni = b.ie[0] - b.is[0] + 1, nj = b.ie[1] - b.is[1] + 1, nk = b.ie[2] - b.is[2] + 1;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)ni * nj * nk) return;
  const int i = (int)(t % ni) + b.is[0];
  const int j = (int)((t / ni) % nj) + b.is[1];
  const int k = (int)(t / ((long)ni * nj)) + b.is[2];
  const long c = ((long)k * b.nj + j) * b.ni + i;
  const double a1 = b.dx[1] * b.dx[2], a2 = b.dx[0] * b.dx[2], a3 = b.dx[0] * b.dx[1];
  const double vol = b.dx[0] * b.dx[1] * b.dx[2];
#pragma unroll
  for (int l = 0; l < NC; ++l) {
    const double *F1 = flux[l];
    double du = a1 * F1[c + 1] - a1 * F1[c];
    if (b.ndim >= 2) {
      const double *F2 = flux[8 + l];
      du += a2 * F2[c + b.ni] - a2 * F2[c];
    }
    if (b.ndim == 3) {
      const double *F3 = flux[16 + l];
      du += a3 * F3[c + (long)b.ni * b.nj] - a3 * F3[c];
    }
    double dudt = -du / vol;
    if (src[l] != nullptr && l >= 1 && l <= 4) dudt = dudt + src[l][c];
    const double us = beta * u_stage[l][c] + (1 - beta) * u_base[l][c];
    u_stage[l][c] = us;
    u_out[l][c] = us + bdt * dudt;
  }
}

// ---- con2prim over a cell range (fluid.cpp:478-517) ------------------------------------
template <bool MHD, int GEOM>
__global__ __launch_bounds__(256) void k_con2prim(DevParams p, BlockGeom b, Range3 rg,
                                                  ConsPtrs cons, PrimPtrs prim,
                                                  double *__restrict__ temp,
                                                  double *__restrict__ mu,
                                                  double *__restrict__ csig,
                                                  double *__restrict__ fail,
                                                  int32_t *__restrict__ iters) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)rg.n[0] * rg.n[1] * rg.n[2]) return;
  const int i = (int)(t % rg.n[0]) + rg.lo[0];
  const int j = (int)((t / rg.n[0]) % rg.n[1]) + rg.lo[1];
  const int k = (int)(t / ((long)rg.n[0] * rg.n[1])) + rg.lo[2];
  con2prim_cell<MHD, GEOM>(p, b, i, j, k, cons, prim, temp, mu, csig, fail, iters);
}

// ---- prim -> cons over a cell range (fluid.cpp:359-449) --------------------------------
template <bool MHD, int GEOM>
__global__ __launch_bounds__(256) void k_prim_to_con(DevParams p, BlockGeom b, Range3 rg,
                                                     PrimPtrs prim, ConsPtrs cons,
                                                     double *__restrict__ csig) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)rg.n[0] * rg.n[1] * rg.n[2]) return;
  const int i = (int)(t % rg.n[0]) + rg.lo[0];
  const int j = (int)((t / rg.n[0]) % rg.n[1]) + rg.lo[1];
  const int k = (int)(t / ((long)rg.n[0] * rg.n[1])) + rg.lo[2];
  prim_to_con_cell<MHD, GEOM>(p, b, i, j, k, prim, cons, csig);
}

// ---- CFL timestep (fluid.cpp:899-944): per-block minimum, then one finishing block ----
__global__ __launch_bounds__(256) void k_dt_partial(BlockGeom b, const double *__restrict__ csig,
                                                    const double *__restrict__ fsig,
                                                    double *__restrict__ partial) {
  const int ni = b.ie[0] - b.is[0] + 1, nj = b.ie[1] - b.is[1] + 1, nk = b.ie[2] - b.is[2] + 1;
  const long total = (long)ni * nj * nk;
  const long n = b.ncell;
  double lmin = DBL_MAX;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long)gridDim.x * blockDim.x) {
    const int i = (int)(t % ni) + b.is[0];
    const int j = (int)((t / ni) % nj) + b.is[1];
    const int k = (int)(t / ((long)ni * nj)) + b.is[2];
    const long c = ((long)k * b.nj + j) * b.ni + i;
    double ldt = 0.0;
    for (int d = 0; d < b.ndim; ++d) {
      double s = csig[d * n + c];
      if (fsig) {
        const long off = d == 0 ? 1 : d == 1 ? (long)b.ni : (long)b.ni * b.nj;
        s = smax(s, smax(fsig[d * n + c], fsig[d * n + c + off]));
      }
      ldt += s / b.dx[d];
    }
    lmin = smin(lmin, 1.0 / (ldt + 1.e-50));
  }
  lmin = block_min(lmin);
  if (threadIdx.x == 0) partial[blockIdx.x] = lmin;
}

__global__ __launch_bounds__(1024) void k_dt_finish(const double *__restrict__ partial, int np,
                                                    double cfl, double *__restrict__ dt) {
  double v = DBL_MAX;
  for (int t = threadIdx.x; t < np; t += blockDim.x) v = smin(v, partial[t]);
  v = block_min(v);
  if (threadIdx.x == 0) dt[0] = cfl * v;
}

// ---- geometric sources for curved metrics (fluid.cpp:566-657, tmunu.hpp) ----------------
struct KSDerivs {
  double dg[4][4][4];
  double da[4];
  double gi[4][4];
};

// inverse spacetime metric of spherical Kerr-Schild (spherical_kerr_schild.hpp:86-102)
__device__ void ks_gcon4(double a, double X1, double X2, double gi[4][4]) {
  const double a2 = a * a;
  const double r = fabs(X1), cth = cos(X2), sth = sin(X2);
  const double sth2 = sth * sth, rho2 = r * r + a2 * cth * cth;
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) gi[x][y] = 0.0;
  gi[0][1] = gi[1][0] = ratio(2 * r, rho2);
  gi[0][0] = -1 - gi[0][1];
  gi[1][1] = ratio(-2 * r + rho2 + a2 * sth2, rho2);
  gi[1][3] = gi[3][1] = ratio(a, rho2);
  gi[2][2] = ratio(1, rho2);
  gi[3][3] = ratio(1, rho2 * sth2);
}

__device__ void ks_derivs(double a, double X1, double X2, KSDerivs &o) {
  const double a2 = a * a;
  const double r = fabs(X1), th = X2, cth = cos(th), sth = sin(th);
  const double sth2 = sth * sth, rho2 = r * r + a2 * cth * cth;
  const double r2 = r * r, r3 = r2 * r, r5 = r2 * r3, a4 = a2 * a2;
  const double rho4 = rho2 * rho2, cth2 = cth * cth;
  const double s2th = sin(2 * th), c2th = cos(2 * th);
  for (int x = 0; x < 4; ++x)
    for (int y = 0; y < 4; ++y) {
      o.gi[x][y] = 0.0;
      for (int z = 0; z < 4; ++z) o.dg[x][y][z] = 0.0;
    }
  o.dg[0][0][1] = 2 * ratio(rho2 - 2 * r2, rho4);
  o.dg[0][1][1] = o.dg[1][0][1] = o.dg[0][0][1];
  o.dg[0][3][1] = o.dg[3][0][1] = ratio(2 * a * (r2 - a2 * cth2) * sth2, rho4);
  o.dg[1][1][1] = o.dg[0][0][1];
  o.dg[1][3][1] = o.dg[3][1][1] = o.dg[0][3][1];
  o.dg[2][2][1] = 2 * r;
  o.dg[3][3][1] = ratio(2 * sth2 * (r5 + a4 * r * cth2 * cth2 - a2 * r2 * sth2 + cth2 * (2 * a2 * r3 + a4 * sth2)), rho4);
  o.dg[0][0][2] = ratio(4 * r * (rho2 - r2) * sth, rho4 * cth);
  o.dg[0][1][2] = o.dg[1][0][2] = o.dg[0][0][2];
  o.dg[0][3][2] = o.dg[3][0][2] = -ratio(2 * a * r * (a2 + r2) * s2th, rho4);
  o.dg[1][1][2] = o.dg[0][0][2];
  o.dg[1][3][2] = o.dg[3][1][2] = -ratio(2 * a * cth * (r3 * (2 + r) + a4 * cth2 * cth2 + a2 * r * (2 + r + r * c2th)) * sth, rho4);
  o.dg[2][2][2] = -a2 * s2th;
  o.dg[3][3][2] = (a2 + r * (r - 2) + ratio(2 * r * (a2 + r2) * (a2 + r2), rho4)) * s2th;
  o.da[0] = o.da[3] = 0.0;
  o.da[1] = ratio(r, rho2) - ratio(1 + r, 2 * r + rho2);
  o.da[2] = -ratio(2 * a2 * r * cth * sth, 2 * r * rho2 + rho4);
  ks_gcon4(a, X1, X2, o.gi);
}

// FMKS derivatives by central differences (geometry_utils.hpp:94-187): step
// max(|X_d|, 1) eps^(1/3) per direction, g_{mu nu} and alpha at the same two points; the
// inverse spacetime metric from the Kerr-Schild one and the Jacobian inverse
// (modified_system.hpp:96-112), in the reference's summation order.
__device__ void fmks_derivs(const DevParams &p, double X1, double X2, double X3,
                            KSDerivs &o) {
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) {
      o.gi[x][y] = 0.0;
#pragma unroll
      for (int z = 0; z < 4; ++z) o.dg[x][y][z] = 0.0;
    }
  o.da[0] = o.da[3] = 0.0;
  const double X[3] = {X1, X2, X3};
  // the FMKS metric does not depend on X3 (FMKSMetric ignores it), so the X3 differences
  // are exactly zero: d < 3 gives the same numbers without evaluating those two points
#pragma unroll
  for (int d = 1; d < 3; ++d) {
    double L[3] = {X1, X2, X3}, R[3] = {X1, X2, X3};
    const double dx = smax(fabs(X[d - 1]), 1.0) * p.fd_eps;
    L[d - 1] -= dx;
    R[d - 1] += dx;
    const FMKSMetric gr(p, R[0], R[1], R[2]);
    const FMKSMetric gl(p, L[0], L[1], L[2]);
#pragma unroll
    for (int mu = 0; mu < 4; ++mu)
#pragma unroll
      for (int nu = 0; nu < 4; ++nu)
        o.dg[mu][nu][d] = ratio(gr.gcov(mu, nu) - gl.gcov(mu, nu), 2 * dx);
    o.da[d] = ratio(gr.alpha() - gl.alpha(), dx * (gl.alpha() + gr.alpha()));
  }
  const MgrMap m(p.mgr, X1, X2);
  KSDerivs k;
  ks_gcon4(p.a, m.r, m.th, k.gi);
  double SJ[4][4] = {};
  SJ[0][0] = 1.0;
  SJ[1][1] = ratio(1., m.r1);
  SJ[2][1] = -ratio(m.t1, m.r1 * m.t2);
  SJ[2][2] = ratio(1., m.t2);
  SJ[3][3] = 1.0;
#pragma unroll
  for (int mu = 0; mu < 4; ++mu)
#pragma unroll
    for (int nu = 0; nu < 4; ++nu) {
      double acc = 0.0;
#pragma unroll
      for (int lam = 0; lam < 4; ++lam)
#pragma unroll
        for (int kap = 0; kap < 4; ++kap) acc += k.gi[lam][kap] * SJ[mu][lam] * SJ[nu][kap];
      o.gi[mu][nu] = acc;
    }
}

template <int GEOM>
__device__ __forceinline__ void geom_derivs(const DevParams &p, double X1, double X2,
                                            double X3, KSDerivs &o) {
  if constexpr (GEOM == PHX_GEOM_FMKS)
    fmks_derivs(p, X1, X2, X3, o);
  else
    ks_derivs(p.a, X1, X2, o);
}

// the (mu, nu) pair e of a symmetric 4x4 in the cached plane's order (upper triangle, row
// by row)
__host__ __device__ constexpr int sym_a(int e) { return e < 4 ? 0 : e < 7 ? 1 : e < 9 ? 2 : 3; }
__host__ __device__ constexpr int sym_b(int e) {
  return e < 4 ? e : e < 7 ? e - 3 : e < 9 ? e - 5 : 3;
}

// geom_derivs at cell-plane point q of block b: from the cached plane when there is one
// (CACHED: always)
template <int GEOM, bool CACHED = false>
__device__ __forceinline__ void geom_derivs_at(const DevParams &p, const BlockGeom &b, long q,
                                               double X1, double X2, double X3, KSDerivs &o) {
  if constexpr (GEOM == PHX_GEOM_FMKS) {
    if (CACHED || b.geo) {
      const long s = (long)b.ni * b.nj;
      const double *t = b.geo + q;
      for (int x = 0; x < 4; ++x)
        for (int y = 0; y < 4; ++y)
          for (int z = 0; z < 4; ++z) o.dg[x][y][z] = 0.0;
#pragma unroll
      for (int d = 1; d < 3; ++d)
#pragma unroll
        for (int e = 0; e < 10; ++e) {
          const double v = t[(kGeoMetric + 10 * (d - 1) + e) * s];
          o.dg[sym_a(e)][sym_b(e)][d] = v;
          o.dg[sym_b(e)][sym_a(e)][d] = v;
        }
      o.da[0] = o.da[3] = 0.0;
      o.da[1] = t[(kGeoMetric + 20) * s];
      o.da[2] = t[(kGeoMetric + 21) * s];
#pragma unroll
      for (int e = 0; e < 10; ++e) {
        const double v = t[(kGeoMetric + 22 + e) * s];
        o.gi[sym_a(e)][sym_b(e)] = v;
        o.gi[sym_b(e)][sym_a(e)] = v;
      }
      return;
    }
  }
  if constexpr (!(CACHED && GEOM == PHX_GEOM_FMKS)) geom_derivs<GEOM>(p, X1, X2, X3, o);
}

// fills the cached plane of block b (cells of the entire (i, j) range): the FMKS metric at
// centres and x1 / x2 faces, and the centres' derivatives and inverse 4-metric
__global__ __launch_bounds__(256) void k_geo_fill(DevParams p, BlockGeom b, double *geo) {
  const long s = (long)b.ni * b.nj;
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= s) return;
  const int i = (int)(q % b.ni), j = (int)(q / b.ni);
  const double X1 = xc(b, 0, i), X2 = xc(b, 1, j), X3 = xc(b, 2, b.is[2]);
  FMKSMetric(p, X1, X2, X3).store(geo + q, s);
  FMKSMetric(p, xf(b, 0, i), X2, X3).store(geo + kGeoCentre * s + q, s);
  FMKSMetric(p, X1, xf(b, 1, j), X3).store(geo + (kGeoCentre + kGeoMetric) * s + q, s);
  KSDerivs o;
  fmks_derivs(p, X1, X2, X3, o);
  for (int d = 1; d < 3; ++d)
    for (int e = 0; e < 10; ++e)
      geo[(kGeoMetric + 10 * (d - 1) + e) * s + q] = o.dg[sym_a(e)][sym_b(e)][d];
  geo[(kGeoMetric + 20) * s + q] = o.da[1];
  geo[(kGeoMetric + 21) * s + q] = o.da[2];
  for (int e = 0; e < 10; ++e) geo[(kGeoMetric + 22 + e) * s + q] = o.gi[sym_a(e)][sym_b(e)];
}

void launch_geo_fill(const DevParams &dp, const BlockGeom &bg, double *geo, hipStream_t s) {
  const long n = (long)bg.ni * bg.nj;
  hipLaunchKernelGGL(k_geo_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dp, bg, geo);
}

// geometric source terms of interior cell t of block b
template <bool MHD, int GEOM, bool CACHED = false, class PP, class SP>
__device__ __forceinline__ void sources_body(const DevParams &p, const BlockGeom &b, long t,
                                             const PP &prim, const SP &src) {
  using L = Layout<MHD>;
  using M = typename MetricOf<GEOM>::type;
  const int ni = b.ie[0] - b.is[0] + 1, nj = b.ie[1] - b.is[1] + 1, nk = b.ie[2] - b.is[2] + 1;
  if (t >= (long)ni * nj * nk) return;
  const int i = (int)(t % ni) + b.is[0];
  const int j = (int)((t / ni) % nj) + b.is[1];
  const int k = (int)(t / ((long)ni * nj)) + b.is[2];
  const long c = ((long)k * b.nj + j) * b.ni + i;
  const double X1 = xc(b, 0, i), X2 = xc(b, 1, j), X3 = xc(b, 2, k);
  const long q = (long)j * b.ni + i;
  const M g = metric_at<GEOM, CACHED>(p, b, 0, q, X1, X2, X3);
  KSDerivs dv;
  geom_derivs_at<GEOM, CACHED>(p, b, q, X1, X2, X3, dv);
  double pr[L::nprim];
  for (int v = 0; v < L::nprim; ++v) pr[v] = prim[v][c];
  // tmunu.hpp:94-126
  double gam[3][3];
  for (int x = 0; x < 3; ++x)
    for (int y = 0; y < 3; ++y) gam[x][y] = g.gcov(x + 1, y + 1);
  const double vp[3] = {pr[1], pr[2], pr[3]};
  const double bp[3] = {MHD ? pr[5] : 0.0, MHD ? pr[6] : 0.0, MHD ? pr[7] : 0.0};
  double vsq = 0.;
  for (int ii = 0; ii < 3; ++ii)
    for (int jj = 0; jj < 3; ++jj) vsq += gam[ii][jj] * vp[ii] * vp[jj];
  const double W = sqrt(1. + vsq);
  double Bdotv = 0.0, Bsq = 0.0;
  for (int ii = 0; ii < 3; ++ii)
    for (int jj = 0; jj < 3; ++jj) {
      Bdotv += bp[ii] * (vp[jj] / W) * gam[ii][jj];
      Bsq += bp[ii] * bp[jj] * gam[ii][jj];
    }
  const double iW = ratio(1., W);
  const double alpha = g.alpha();
  double u[4], bb[4];
  u[0] = ratio(W, fabs(alpha));
  bb[0] = u[0] * Bdotv;
  for (int l = 1; l < 4; ++l) {
    u[l] = vp[l - 1] - u[0] * g.beta(l - 1);
    bb[l] = iW * (bp[l - 1] + alpha * bb[0] * u[l]);
  }
  const double bsq = Bsq * iW * iW + Bdotv * Bdotv;
  const double A = pr[0] + pr[4] + pr[L::prs] + bsq;
  const double B = pr[L::prs] + 0.5 * bsq;
  double T[4][4];
#pragma unroll
  for (int mu = 0; mu < 4; ++mu)
#pragma unroll
    for (int nu = 0; nu < 4; ++nu) T[mu][nu] = A * u[mu] * u[nu] + B * dv.gi[mu][nu] - bb[mu] * bb[nu];
  // fluid.cpp:604-640 with T symmetric and the metric independent of X0 and X3 (both
  // spherical Kerr-Schild and FMKS): only d_1 and d_2 of g are non-zero.  The momentum
  // source T^{mn}(d_m g_{nl} - Gamma_{lnm}) equals T^{mn} d_l g_{mn} / 2 (the reference's
  // form, restated in the oracle, differs by rounding), and the energy source's
  // T^{mn} Gamma_{rmn} equals sum_{m, n} T^{mn} d_n g_{rm} - T^{mn} d_r g_{mn} / 2.
  double P[3] = {0.0, 0.0, 0.0};  // P_c = T^{mn} d_c g_{mn}
#pragma unroll
  for (int cd = 1; cd < 3; ++cd)
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int nn = 0; nn < 4; ++nn) P[cd] += T[m][nn] * dv.dg[m][nn][cd];
  const double gdet = g.detg();
  src[1][c] = gdet * (0.5 * P[1]);
  src[2][c] = gdet * (0.5 * P[2]);
  src[3][c] = 0.0;
  const double ia2 = ratio(1, alpha * alpha);
  const double gcon0[4] = {-ia2, ia2 * g.beta(0), ia2 * g.beta(1), ia2 * g.beta(2)};
  double TGam = 0.0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    double Q = 0.0;  // sum_{m, n} T^{mn} d_n g_{rm}
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int cd = 1; cd < 3; ++cd) Q += T[m][cd] * dv.dg[r][m][cd];
    const double half = (r == 1 || r == 2) ? 0.5 * P[r] : 0.0;
    TGam += gcon0[r] * (Q - half);
  }
  const double Ta = T[1][0] * dv.da[1] + T[2][0] * dv.da[2];
  src[4][c] = gdet * alpha * (Ta - TGam);
}

template <bool MHD, int GEOM>
__global__ __launch_bounds__(128) void k_sources(DevParams p, BlockGeom b, PrimPtrs prim,
                                                 ConsPtrs src) {
  sources_body<MHD, GEOM>(p, b, (long)blockIdx.x * blockDim.x + threadIdx.x, prim, src);
}

template <int NC>
__global__ void k_zero_sources(BlockGeom b, ConsPtrs src) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= b.ncell) return;
  for (int l = 1; l <= 4; ++l) src[l][t] = 0.0;
}

// ---- packs: CT and sources over a work list of 3-D blocks -------------------------------
template <int NC>
__global__ __launch_bounds__(256) void k_pack_ct_emf(PackView v) {
  const long g = xcd_remap(blockIdx.x, gridDim.x);
  const int b = pack_find(v.start, v.n, g);
  const PackEntry &E = v.e[b];
  ct_emf_3d_body<NC>(E.bg, (g - v.start[b]) * 256 + threadIdx.x, contig(E.flux, E.bg.ncell),
                     E.emf);
}

template <int GEOM>
__global__ __launch_bounds__(128) void k_pack_sources(DevParams p, PackView v) {
  const long g = xcd_remap(blockIdx.x, gridDim.x);
  const int b = pack_find(v.start, v.n, g);
  const PackEntry &E = v.e[b];
  // two 128-thread halves per 256-cell group
  const Contig prim = contig(E.prim, E.bg.ncell), src = contig(E.src, E.bg.ncell);
  sources_body<true, GEOM, GEOM == PHX_GEOM_FMKS>(p, E.bg, (g - v.start[b]) * 256 + threadIdx.x,
                                                 prim, src);
  sources_body<true, GEOM, GEOM == PHX_GEOM_FMKS>(p, E.bg,
                                                 (g - v.start[b]) * 256 + 128 + threadIdx.x,
                                                 prim, src);
}

// ---- batched con2prim / prim2con over independent states --------------------------------
template <bool MHD, int GEOM>
__global__ __launch_bounds__(256) void k_con2prim_batch(DevParams p, long n,
                                                        const double *__restrict__ coords,
                                                        double *__restrict__ cons,
                                                        double *__restrict__ prim,
                                                        double *__restrict__ mu,
                                                        double *__restrict__ fail,
                                                        int32_t *__restrict__ iters) {
  using M = typename MetricOf<GEOM>::type;
  using L = Layout<MHD>;
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  double X[3] = {0.0, 0.0, 0.0};
  if (coords) {
    X[0] = coords[s];
    X[1] = coords[n + s];
  }
  const M g(p, X[0], X[1], X[2]);
  double cs[L::ncons], pr[L::nprim], sig[3];
#pragma unroll
  for (int v = 0; v < L::ncons; ++v) cs[v] = cons[v * n + s];
#pragma unroll
  for (int v = 0; v < L::nprim; ++v) pr[v] = 0.0;
  double tt = 0.0, m = mu[s];
  int it = 0;
  const double st = con2prim_zone<MHD>(p, g, X, cs, pr, tt, m, sig, it);
#pragma unroll
  for (int v = 0; v < 5; ++v) cons[v * n + s] = cs[v];
#pragma unroll
  for (int v = 0; v < L::nprim; ++v) prim[v * n + s] = pr[v];
  mu[s] = m;
  fail[s] = st;
  if (iters) iters[s] = it;
}

template <bool MHD, int GEOM>
__global__ __launch_bounds__(256) void k_prim_to_con_batch(DevParams p, long n,
                                                           const double *__restrict__ coords,
                                                           const double *__restrict__ prim,
                                                           double *__restrict__ cons) {
  using M = typename MetricOf<GEOM>::type;
  using L = Layout<MHD>;
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const M g(p, coords ? coords[s] : 0.0, coords ? coords[n + s] : 0.0, 0.0);
  double pr[L::nprim];
#pragma unroll
  for (int v = 0; v < L::nprim; ++v) pr[v] = prim[v * n + s];
  const double vp[3] = {pr[1], pr[2], pr[3]};
  const double bp[3] = {MHD ? pr[5] : 0.0, MHD ? pr[6] : 0.0, MHD ? pr[7] : 0.0};
  double D, S[3], bc[3], tau, sig[3];
  p2c<MHD>(pr[0], vp, bp, pr[4], pr[L::prs], pr[L::gm1], g, D, S, bc, tau, sig);
  cons[s] = D;
  cons[n + s] = S[0];
  cons[2 * n + s] = S[1];
  cons[3 * n + s] = S[2];
  cons[4 * n + s] = tau;
  if constexpr (MHD) {
    cons[5 * n + s] = bc[0];
    cons[6 * n + s] = bc[1];
    cons[7 * n + s] = bc[2];
  }
}

// ================================ host side ==============================================
static inline unsigned nblocks(long work, int bs) { return (unsigned)((work + bs - 1) / bs); }

DevParams make_dev_params(const phx_fluid_params *p) {
  DevParams d;
  d.gm1 = p->gamma - 1.0;
  d.cv = p->cv;
  d.tol = p->c2p_tol;
  d.fsf = p->c2p_floor_scale_fac;
  d.max_iter = p->c2p_max_iter;
  d.fail_on_floors = p->c2p_fail_on_floors;
  d.fail_on_ceilings = p->c2p_fail_on_ceilings;
  d.a = p->geom.a;
  d.bnd.floor_type = p->bounds.floor_type;
  d.bnd.r0 = p->bounds.rho0;
  d.bnd.s0 = p->bounds.sie0;
  d.bnd.ralpha = p->bounds.rho_exp;
  d.bnd.salpha = p->bounds.sie_exp;
  d.bnd.rho_min_eos = p->bounds.rho_min_eos;
  d.bnd.sie_min_eos = p->bounds.sie_min_eos;
  d.bnd.gam_max = p->bounds.gam_max;
  d.bnd.sie_max = p->bounds.sie_max;
  const phx_mgr &m = p->geom.mgr;
  d.mgr.derefine_poles = m.derefine_poles;
  d.mgr.h = m.h;
  d.mgr.xt = m.xt;
  d.mgr.alpha = m.alpha;
  d.mgr.x0 = m.x0;
  d.mgr.smooth = m.smooth;
  d.mgr.hexp_br = m.hexp_br;
  d.mgr.hexp_nsq = m.hexp_nsq;
  d.mgr.hexp_csq = m.hexp_csq;
  // mckinney_gammie_ryan.hpp:192-194 (GetNorm_), on the host in the same operation order
  d.mgr.norm = 0.5 * M_PI * 1. / (1. + 1. / (m.alpha + 1.) * 1. / std::pow(m.xt, m.alpha));
  d.fd_eps = std::pow(DBL_EPSILON, 1.0 / 3.0);  // geometry_utils.hpp:153
  return d;
}

int check_params(phx_ctx *ctx, const phx_fluid_params *p) {
  if (!p) return ctx->fail(PHX_ERR_INVALID, "null parameters");
  if (p->mhd != 0 && p->mhd != 1) return ctx->fail(PHX_ERR_INVALID, "mhd must be 0 or 1");
  if (p->recon < PHX_RECON_CONSTANT || p->recon > PHX_RECON_MP5)
    return ctx->fail(PHX_ERR_INVALID, "invalid reconstruction");
  if (p->riemann != PHX_RIEMANN_LLF && p->riemann != PHX_RIEMANN_HLL)
    return ctx->fail(PHX_ERR_INVALID, "invalid Riemann solver");
  if (p->geom.kind != PHX_GEOM_MINKOWSKI && p->geom.kind != PHX_GEOM_SPHERICAL_KS &&
      p->geom.kind != PHX_GEOM_FMKS)
    return ctx->fail(PHX_ERR_UNSUPPORTED, "unsupported geometry");
  if (p->geom.kind == PHX_GEOM_FMKS &&
      !(p->geom.mgr.xt > 0. && p->geom.mgr.alpha > -1. && p->geom.mgr.hexp_nsq >= 1.))
    return ctx->fail(PHX_ERR_INVALID, "invalid FMKS transformation parameters");
  if (!(p->c2p_floor_scale_fac <= 1. && p->c2p_floor_scale_fac > 0.))
    return ctx->fail(PHX_ERR_INVALID, "c2p_floor_scale_fac must be in (0, 1]");
  if (p->bounds.floor_type < 0 || p->bounds.floor_type > PHX_FLOOR_R_RHO_SIE)
    return ctx->fail(PHX_ERR_INVALID, "invalid floor type");
  if (p->c2p_max_iter < 0) return ctx->fail(PHX_ERR_INVALID, "c2p_max_iter < 0");
  if (p->flux_arith != PHX_FLUX_ARITH_EXACT && p->flux_arith != PHX_FLUX_ARITH_FAST)
    return ctx->fail(PHX_ERR_INVALID, "invalid flux arithmetic");
  return PHX_OK;
}

int check_block(phx_ctx *ctx, const phx_block *b, int recon) {
  if (!b) return ctx->fail(PHX_ERR_INVALID, "null block");
  if (b->ndim < 1 || b->ndim > 3) return ctx->fail(PHX_ERR_INVALID, "ndim must be 1..3");
  for (int d = 0; d < 3; ++d) {
    if (b->nx[d] < 1) return ctx->fail(PHX_ERR_INVALID, "nx must be >= 1");
    if (d >= b->ndim && b->nx[d] != 1)
      return ctx->fail(PHX_ERR_INVALID, "nx must be 1 in inactive dimensions");
  }
  const int need = recon >= PHX_RECON_WENO5Z ? 4 : 2;
  if (b->ng < need) return ctx->fail(PHX_ERR_INVALID, "not enough ghost zones for the reconstruction");
  long cells = 1;
  for (int d = 0; d < 3; ++d) cells *= b->nx[d] + (d < b->ndim ? 2 * b->ng : 0);
  if (cells > (1L << 31)) return ctx->fail(PHX_ERR_INVALID, "block too large");
  return PHX_OK;
}

long pack_groups_cells(const BlockGeom &bg, bool corners) {
  const long c = corners ? 1 : 0;
  const long n = (bg.ie[0] - bg.is[0] + 1 + c) * (bg.ie[1] - bg.is[1] + 1 + c) *
                 (bg.ie[2] - bg.is[2] + 1 + c);
  return (n + 255) / 256;
}

void launch_pack_sources(int geom, const DevParams &dp, const PackView &v, long groups,
                         hipStream_t s) {
  if (groups <= 0) return;
  with_geom(geom, [&](auto G) {
    constexpr int GEOM = decltype(G)::value;
    if constexpr (GEOM != PHX_GEOM_MINKOWSKI)
      hipLaunchKernelGGL((k_pack_sources<GEOM>), dim3((unsigned)groups), dim3(128), 0, s, dp, v);
  });
}

void launch_pack_ct_emf(const PackView &v, long groups, hipStream_t s) {
  if (groups <= 0) return;
  hipLaunchKernelGGL(k_pack_ct_emf<8>, dim3((unsigned)groups), dim3(256), 0, s, v);
}

void launch_ct_emf(const BlockGeom &bg, const FluxPtrs &flux, double *emf, hipStream_t s) {
  const int nx0 = bg.ie[0] - bg.is[0] + 1, nx1 = bg.ie[1] - bg.is[1] + 1;
  const int nx2 = bg.ie[2] - bg.is[2] + 1;
  if (bg.ndim == 3) {
    const long n = (long)(nx0 + 1) * (nx1 + 1) * (nx2 + 1);
    hipLaunchKernelGGL(k_ct_emf_3d<8>, dim3(nblocks(n, 256)), dim3(256), 0, s, bg, flux, emf);
  } else {
    const long n = (long)(nx0 + 1) * (nx1 + 1) * nx2;
    hipLaunchKernelGGL(k_ct_emf_2d<8>, dim3(nblocks(n, 256)), dim3(256), 0, s, bg, flux, emf);
  }
}

void launch_flux_ct(const BlockGeom &bg, const FluxPtrs &flux, double *emf, hipStream_t s) {
  const int nx0 = bg.ie[0] - bg.is[0] + 1, nx1 = bg.ie[1] - bg.is[1] + 1;
  const int nx2 = bg.ie[2] - bg.is[2] + 1;
  if (bg.ndim == 3) {
    const long n = (long)(nx0 + 1) * (nx1 + 1) * (nx2 + 1);
    hipLaunchKernelGGL(k_ct_emf_3d<8>, dim3(nblocks(n, 256)), dim3(256), 0, s, bg, flux, emf);
    hipLaunchKernelGGL(k_ct_flux_3d<8>, dim3(nblocks(n, 256)), dim3(256), 0, s, bg, flux, emf);
  } else {
    const long n = (long)(nx0 + 1) * (nx1 + 1) * nx2;
    hipLaunchKernelGGL(k_ct_emf_2d<8>, dim3(nblocks(n, 256)), dim3(256), 0, s, bg, flux, emf);
    hipLaunchKernelGGL(k_ct_flux_2d<8>, dim3(nblocks(n, 256)), dim3(256), 0, s, bg, flux, emf);
  }
}

void launch_sources(int geom, int mhd, const DevParams &dp, const BlockGeom &bg,
                    const PrimPtrs &prim, const ConsPtrs &src, hipStream_t s) {
  if (geom == PHX_GEOM_MINKOWSKI) {
    if (mhd)
      hipLaunchKernelGGL(k_zero_sources<8>, dim3(nblocks(bg.ncell, 256)), dim3(256), 0, s, bg, src);
    else
      hipLaunchKernelGGL(k_zero_sources<5>, dim3(nblocks(bg.ncell, 256)), dim3(256), 0, s, bg, src);
    return;
  }
  const long n = (long)(bg.ie[0] - bg.is[0] + 1) * (bg.ie[1] - bg.is[1] + 1) * (bg.ie[2] - bg.is[2] + 1);
  with_geom(geom, [&](auto G) {
    constexpr int GEOM = decltype(G)::value;
    if constexpr (GEOM != PHX_GEOM_MINKOWSKI)
      with_mhd(mhd, [&](auto H) {
        hipLaunchKernelGGL((k_sources<decltype(H)::value, GEOM>), dim3(nblocks(n, 128)), dim3(128),
                           0, s, dp, bg, prim, src);
      });
  });
}

void launch_dt_finish(const double *partial, int np, double cfl, double *dt, hipStream_t s) {
  hipLaunchKernelGGL(k_dt_finish, dim3(1), dim3(1024), 0, s, partial, np, cfl, dt);
}

int make_range(phx_ctx *ctx, const phx_block *b, const int lo[3], const int hi[3], Range3 &rg) {
  const BlockGeom bg = make_block_geom(b);
  const int n[3] = {bg.ni, bg.nj, bg.nk};
  for (int d = 0; d < 3; ++d) {
    const int l = lo ? lo[d] : 0, h = hi ? hi[d] : n[d] - 1;
    if (l < 0 || h >= n[d] || h < l) return ctx->fail(PHX_ERR_INVALID, "cell range outside the block");
    rg.lo[d] = l;
    rg.n[d] = h - l + 1;
  }
  return PHX_OK;
}

}  // namespace phx

using namespace phx;

// ================================ C ABI ==================================================
extern "C" {

int phx_calculate_fluxes(phx_ctx *ctx, const phx_fluid_params *p, const phx_block *b,
                         const phx_block_fields *f, phx_stream stream) {
  if (!ctx) return PHX_ERR_INVALID;
  int rc = check_params(ctx, p);
  if (rc) return rc;
  if ((rc = check_block(ctx, b, p->recon))) return rc;
  if (!f) return ctx->fail(PHX_ERR_INVALID, "null field");
  const BlockGeom bg = make_block_geom(b);
  const int np = p->mhd ? 10 : 7, nc = p->mhd ? 8 : 5;
  const PrimPtrs prim = prim_ptrs(f, bg.ncell, np);
  const FluxPtrs flux = flux_ptrs(f, bg.ncell, nc);
  if (!complete(prim, np) || !complete_flux(flux, b->ndim, nc) || !f->face_sig)
    return ctx->fail(PHX_ERR_INVALID, "null field");
  const DevParams dp = make_dev_params(p);
  hipStream_t s = ctx->stream_of(stream);
  for (int d = 0; d < b->ndim; ++d)
    launch_fluxes(d, p->recon, p->riemann, p->mhd, p->geom.kind, dp, bg, prim, flux,
                  f->face_sig, s, ctx->tile_for(p, b, d), flux_fast_of(p));
  return ctx->check_launch();
}

int phx_flux_ct(phx_ctx *ctx, const phx_fluid_params *p, const phx_block *b,
                const phx_block_fields *f, phx_stream stream) {
  if (!ctx) return PHX_ERR_INVALID;
  int rc = check_params(ctx, p);
  if (rc) return rc;
  if ((rc = check_block(ctx, b, PHX_RECON_LINEAR))) return rc;  // CT reads 2 ghost layers
  if (!p->mhd || b->ndim == 1) return PHX_OK;
  if (!f || !f->emf) return ctx->fail(PHX_ERR_INVALID, "null field");
  const BlockGeom bg = make_block_geom(b);
  const FluxPtrs flux = flux_ptrs(f, bg.ncell, 8);
  if (!complete_flux(flux, b->ndim, 8)) return ctx->fail(PHX_ERR_INVALID, "null field");
  launch_flux_ct(bg, flux, f->emf, ctx->stream_of(stream));
  return ctx->check_launch();
}

int phx_fluid_source_terms(phx_ctx *ctx, const phx_fluid_params *p, const phx_block *b,
                           const phx_block_fields *f, phx_stream stream) {
  if (!ctx) return PHX_ERR_INVALID;
  int rc = check_params(ctx, p);
  if (rc) return rc;
  if ((rc = check_block(ctx, b, PHX_RECON_CONSTANT))) return rc;
  if (!f) return ctx->fail(PHX_ERR_INVALID, "null field");
  const BlockGeom bg = make_block_geom(b);
  const int np = p->mhd ? 10 : 7, nc = p->mhd ? 8 : 5;
  const PrimPtrs prim = prim_ptrs(f, bg.ncell, np);
  const ConsPtrs src = src_ptrs(f, bg.ncell, nc);
  if (!complete(prim, np) || !complete(src, nc)) return ctx->fail(PHX_ERR_INVALID, "null field");
  launch_sources(p->geom.kind, p->mhd, make_dev_params(p), bg, prim, src, ctx->stream_of(stream));
  return ctx->check_launch();
}

int phx_flux_divergence_update(phx_ctx *ctx, const phx_fluid_params *p, const phx_block *b,
                               const double *flux, const double *src, double *u_stage,
                               const double *u_base, double beta, double dt, double *u_out,
                               phx_stream stream) {
  if (!ctx) return PHX_ERR_INVALID;
  int rc = check_params(ctx, p);
  if (rc) return rc;
  if ((rc = check_block(ctx, b, PHX_RECON_CONSTANT))) return rc;
  if (!flux || !u_stage || !u_base || !u_out) return ctx->fail(PHX_ERR_INVALID, "null field");
  const BlockGeom bg = make_block_geom(b);
  hipStream_t s = ctx->stream_of(stream);
  const long n = (long)b->nx[0] * b->nx[1] * b->nx[2];
  const double bdt = beta * dt;
  const int nc = p->mhd ? 8 : 5;
  const long m = bg.ncell;
  const FluxPtrs F = FluxPtrs::flux(flux, m, nc);
  const ConsPtrs S = ConsPtrs::contiguous(src, m, nc), Us = ConsPtrs::contiguous(u_stage, m, nc);
  const ConsPtrs Ub = ConsPtrs::contiguous(u_base, m, nc), Uo = ConsPtrs::contiguous(u_out, m, nc);
  if (p->mhd)
    hipLaunchKernelGGL(k_update<8>, dim3(nblocks(n, 256)), dim3(256), 0, s, bg, F, S, Us, Ub, beta, bdt, Uo);
  else
    hipLaunchKernelGGL(k_update<5>, dim3(nblocks(n, 256)), dim3(256), 0, s, bg, F, S, Us, Ub, beta, bdt, Uo);
  return ctx->check_launch();
}


int phx_con2prim(phx_ctx *ctx, const phx_fluid_params *p, const phx_block *b,
                 const phx_block_fields *f, const int lo[3], const int hi[3], int32_t *iters,
                 phx_stream stream) {
  if (!ctx) return PHX_ERR_INVALID;
  int rc = check_params(ctx, p);
  if (rc) return rc;
  if ((rc = check_block(ctx, b, PHX_RECON_CONSTANT))) return rc;
  if (!f) return ctx->fail(PHX_ERR_INVALID, "null field");
  const BlockGeom bg = make_block_geom(b);
  const int np = p->mhd ? 10 : 7, nc = p->mhd ? 8 : 5;
  const PrimPtrs prim = prim_ptrs(f, bg.ncell, np);
  const ConsPtrs cons = cons_ptrs(f, bg.ncell, nc);
  if (!complete(cons, nc) || !complete(prim, np) || !f->c2p_mu || !f->cell_sig || !f->fail)
    return ctx->fail(PHX_ERR_INVALID, "null field");
  Range3 rg;
  if ((rc = make_range(ctx, b, lo, hi, rg))) return rc;
  const DevParams dp = make_dev_params(p);
  hipStream_t s = ctx->stream_of(stream);
  const long n = (long)rg.n[0] * rg.n[1] * rg.n[2];
  const dim3 grid(nblocks(n, 256)), blk(256);
#define C2P_ARGS dp, bg, rg, cons, prim, f->temp, f->c2p_mu, f->cell_sig, f->fail, iters
  with_geom(p->geom.kind, [&](auto G) {
    with_mhd(p->mhd, [&](auto H) {
      hipLaunchKernelGGL((k_con2prim<decltype(H)::value, decltype(G)::value>), grid, blk, 0, s, C2P_ARGS);
    });
  });
#undef C2P_ARGS
  return ctx->check_launch();
}

int phx_prim_to_con(phx_ctx *ctx, const phx_fluid_params *p, const phx_block *b,
                    const phx_block_fields *f, const int lo[3], const int hi[3],
                    phx_stream stream) {
  if (!ctx) return PHX_ERR_INVALID;
  int rc = check_params(ctx, p);
  if (rc) return rc;
  if ((rc = check_block(ctx, b, PHX_RECON_CONSTANT))) return rc;
  if (!f) return ctx->fail(PHX_ERR_INVALID, "null field");
  const BlockGeom bg = make_block_geom(b);
  const int np = p->mhd ? 10 : 7, nc = p->mhd ? 8 : 5;
  const PrimPtrs prim = prim_ptrs(f, bg.ncell, np);
  const ConsPtrs cons = cons_ptrs(f, bg.ncell, nc);
  if (!complete(cons, nc) || !complete(prim, np)) return ctx->fail(PHX_ERR_INVALID, "null field");
  Range3 rg;
  if ((rc = make_range(ctx, b, lo, hi, rg))) return rc;
  const DevParams dp = make_dev_params(p);
  hipStream_t s = ctx->stream_of(stream);
  const long n = (long)rg.n[0] * rg.n[1] * rg.n[2];
  const dim3 grid(nblocks(n, 256)), blk(256);
  with_geom(p->geom.kind, [&](auto G) {
    with_mhd(p->mhd, [&](auto H) {
      hipLaunchKernelGGL((k_prim_to_con<decltype(H)::value, decltype(G)::value>), grid, blk, 0, s, dp, bg, rg, prim, cons, f->cell_sig);
    });
  });
  return ctx->check_launch();
}

int phx_estimate_dt(phx_ctx *ctx, const phx_fluid_params *p, const phx_block *b,
                    const phx_block_fields *f, int use_face_speeds, double cfl,
                    double *dt_out, phx_stream stream) {
  if (!ctx) return PHX_ERR_INVALID;
  int rc = check_params(ctx, p);
  if (rc) return rc;
  if ((rc = check_block(ctx, b, PHX_RECON_CONSTANT))) return rc;
  if (!f || !f->cell_sig || !dt_out || (use_face_speeds && !f->face_sig))
    return ctx->fail(PHX_ERR_INVALID, "null field");
  const BlockGeom bg = make_block_geom(b);
  hipStream_t s = ctx->stream_of(stream);
  const long n = (long)b->nx[0] * b->nx[1] * b->nx[2];
  const int np = (int)std::min<long>(nblocks(n, 256), kDtPartials);
  hipLaunchKernelGGL(k_dt_partial, dim3(np), dim3(256), 0, s, bg, f->cell_sig,
                     use_face_speeds ? f->face_sig : nullptr, ctx->dt_partial);
  launch_dt_finish(ctx->dt_partial, np, cfl, dt_out, s);
  return ctx->check_launch();
}

int phx_con2prim_batch(phx_ctx *ctx, const phx_fluid_params *p, int64_t n,
                       const double *coords, double *cons, double *prim, double *mu,
                       double *fail, int32_t *iters, phx_stream stream) {
  if (!ctx) return PHX_ERR_INVALID;
  int rc = check_params(ctx, p);
  if (rc) return rc;
  if (n < 0) return ctx->fail(PHX_ERR_INVALID, "negative batch size");
  if (n == 0) return PHX_OK;
  if (!cons || !prim || !mu || !fail) return ctx->fail(PHX_ERR_INVALID, "null field");
  if (p->geom.kind != PHX_GEOM_MINKOWSKI && !coords)
    return ctx->fail(PHX_ERR_INVALID, "coordinates required for curved geometry");
  const DevParams dp = make_dev_params(p);
  hipStream_t s = ctx->stream_of(stream);
  const dim3 grid(nblocks(n, 256)), blk(256);
  with_geom(p->geom.kind, [&](auto G) {
    with_mhd(p->mhd, [&](auto H) {
      hipLaunchKernelGGL((k_con2prim_batch<decltype(H)::value, decltype(G)::value>), grid, blk, 0, s, dp, (long)n, coords, cons, prim, mu, fail, iters);
    });
  });
  return ctx->check_launch();
}

int phx_prim_to_con_batch(phx_ctx *ctx, const phx_fluid_params *p, int64_t n,
                          const double *coords, const double *prim, double *cons,
                          phx_stream stream) {
  if (!ctx) return PHX_ERR_INVALID;
  int rc = check_params(ctx, p);
  if (rc) return rc;
  if (n < 0) return ctx->fail(PHX_ERR_INVALID, "negative batch size");
  if (n == 0) return PHX_OK;
  if (!cons || !prim) return ctx->fail(PHX_ERR_INVALID, "null field");
  if (p->geom.kind != PHX_GEOM_MINKOWSKI && !coords)
    return ctx->fail(PHX_ERR_INVALID, "coordinates required for curved geometry");
  const DevParams dp = make_dev_params(p);
  hipStream_t s = ctx->stream_of(stream);
  const dim3 grid(nblocks(n, 256)), blk(256);
  with_geom(p->geom.kind, [&](auto G) {
    with_mhd(p->mhd, [&](auto H) {
      hipLaunchKernelGGL((k_prim_to_con_batch<decltype(H)::value, decltype(G)::value>), grid, blk, 0, s, dp, (long)n, coords, prim, cons);
    });
  });
  return ctx->check_launch();
}

}  // extern "C"
