// phx_device.hpp -- device-side physics of the fluid stage for gfx950 (HIP, FP64).
//
// Point functions used by the kernels in phx_kernels.hip.  Each follows the reference
// function cited beside it in operation order (the library is built with
// -ffp-contract=off so that +,-,*,/,sqrt round exactly as the reference's would), but
// is written for the GPU: metrics are compile-time policies whose structural zeros are
// removed at compile time, nothing is heap- or pack-indexed, and every per-zone state
// lives in registers.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "phx.h"

namespace phx {

// ---- parameters as the kernels see them ------------------------------------------------
struct DevBounds {
  int floor_type;
  double r0, s0, ralpha, salpha, rho_min_eos, sie_min_eos, gam_max, sie_max;
};

struct DevParams {
  double gm1, cv;          // ideal gas: Gamma - 1 (computed on the host), C_v
  double tol;              // c2p bracket half-width tolerance
  double fsf;              // c2p_floor_scale_fac
  int max_iter;
  int fail_on_floors, fail_on_ceilings;
  double a;                // Kerr-Schild spin
  DevBounds bnd;
  // McKinney-Gammie-Ryan transformation (FMKS); norm = 0.5 pi / (1 + xt^-alpha/(alpha+1))
  struct Mgr {
    int derefine_poles;
    double h, xt, alpha, x0, smooth, hexp_br, hexp_nsq, hexp_csq, norm;
  } mgr;
  double fd_eps;  // eps^(1/3), the finite-difference scale of the FMKS derivatives
};

// ---- block geometry as the kernels see it ---------------------------------------------
struct BlockGeom {
  int ni, nj, nk;
  int ndim, ng;
  int is[3], ie[3];  // interior index range per dimension (i, j, k)
  long ncell;
  double xmin[3], dx[3];
  int istart[3];     // parthenon's istart: ng in active dimensions, 0 otherwise
  // geometry cached per (X1, X2) plane for metrics that do not depend on X3 (FMKS), or
  // null: [kGeoCentre][nj][ni] at cell centres (metric, derivatives, inverse 4-metric),
  // then [kGeoMetric][nj][ni] at x1 faces and at x2 faces
  const double *geo;
};

// ---- per-component field pointers ------------------------------------------------------
// A multi-component field as the kernels see it: component v's [nk][nj][ni] array starts
// at p[v].  Both the library's contiguous fields ([nv][cells]) and a host's separately
// allocated variables (parthenon allocates every variable and every direction of its flux
// on its own; phx_block_vars) become this on the host, and the kernels index p[v][c], so
// neither layout is copied.  Flux tables hold flux[d][m] at p[d * 8 + m].
template <int N>
struct Comp {
  double *p[N];
  __host__ __device__ __forceinline__ double *operator[](int v) const { return p[v]; }
  // component v of a contiguous field of nv components at base (null beyond nv)
  static __host__ __device__ __forceinline__ Comp contiguous(const double *base, long ncell,
                                                            int nv = N) {
    Comp c;
#pragma unroll
    for (int v = 0; v < N; ++v)
      c.p[v] = base && v < nv ? const_cast<double *>(base) + (long)v * ncell : nullptr;
    return c;
  }
  // flux[d][m] of a contiguous [ndim][nc][cells] flux field into the d * 8 + m slots
  static __host__ __device__ __forceinline__ Comp flux(const double *base, long ncell, int nc) {
    static_assert(N == 24, "flux tables have 3 x 8 slots");
    Comp c;
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int m = 0; m < 8; ++m)
        c.p[d * 8 + m] =
            base && m < nc ? const_cast<double *>(base) + ((long)d * nc + m) * ncell : nullptr;
    return c;
  }
};
using PrimPtrs = Comp<10>;
using ConsPtrs = Comp<8>;
using FluxPtrs = Comp<24>;
// The same view of a field the library laid out contiguously ([nv][cells] at base; a flux
// [3][8][cells] is one too, slot d * 8 + m): component v at base + v * n, computed where it
// is used.  The kernels over packs and meshes, whose blocks are always contiguous, take this
// form (their code is then what the contiguous layout compiles to); the per-block entry
// points take Comp tables.
struct Contig {
  double *base;
  long n;
  __device__ __forceinline__ double *operator[](int v) const { return base + (long)v * n; }
};
__device__ __forceinline__ Contig contig(const double *base, long n) {
  return Contig{const_cast<double *>(base), n};
}

// The hardware hands consecutive workgroups of a launch to the 8 XCDs in turn, and each
// XCD has its own L2.  Renumbering so that each XCD receives a contiguous range of
// logical blocks keeps neighbouring tiles (which share cache lines at their edges and
// along i) in one L2.  A bijection on [0, n) for any n.
__device__ __forceinline__ unsigned xcd_remap(unsigned bx, unsigned n) {
  constexpr unsigned kXcd = 8;
  const unsigned per = n / kXcd, rem = n % kXcd;
  const unsigned x = bx % kXcd, idx = bx / kXcd;
  return x < rem ? x * (per + 1) + idx : rem * (per + 1) + (x - rem) * per + idx;
}

__device__ __forceinline__ double xc(const BlockGeom &b, int d, int idx) {
  return b.xmin[d] + (idx - b.istart[d] + 0.5) * b.dx[d];
}
__device__ __forceinline__ double xf(const BlockGeom &b, int d, int idx) {
  return b.xmin[d] + (idx - b.istart[d]) * b.dx[d];
}

// ---- robust.hpp:10-40 -------------------------------------------------------------------
constexpr double kSmall = 10.0 * DBL_MIN;
constexpr double kEps = 10.0 * DBL_EPSILON;
__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double ratio(double a, double b) {
  const double sgn = b >= 0 ? 1.0 : -1.0;
  return a / (b + sgn * kSmall);
}
__device__ __forceinline__ double make_positive(double v) { return smax(v, kEps); }
__device__ __forceinline__ double make_bounded(double v, double lo, double hi) {
  return smin(smax(v, lo + kEps), hi * (1.0 - kEps));
}

// ---- arithmetic of the face states and fluxes ------------------------------------------
// The reconstruction, prim_to_flux and the Riemann solvers divide and take square roots
// through these.  By default they are IEEE division and sqrt, which round exactly as the
// reference.  A translation unit that defines PHX_FLUX_FAST as 1 (the sweeps of the fast
// flux arithmetic, phx_flux_fast.hip, also built with contraction on) gets reciprocals and
// reciprocal square roots from the hardware estimates (v_rcp_f64, v_rsq_f64, 24 bits)
// refined by one Newton-Raphson step (Markstein, IBM J. Res. Develop. 34 (1990) 111-119),
// with the quotient formed as a times the reciprocal: within 2.2e-15 (reciprocals) and
// 4.3e-15 (square roots) relative of the IEEE results, about 10-20 ulp, for finite
// positive denominators and radicands, which is what these code paths divide by (the WENO
// weights are bounded below by their 1e-100 epsilon, Lorentz factors by 1, signal-speed
// radicands by make_positive's EPS), in a third of the instructions.  A zero or infinite
// denominator gives NaN where IEEE division gives +-inf.
#ifndef PHX_FLUX_FAST
#define PHX_FLUX_FAST 0
#endif
constexpr bool kFluxFast = PHX_FLUX_FAST != 0;

__device__ __forceinline__ double frcp(double b) {
  if constexpr (kFluxFast) {
    // r <- r + r (1 - b r) squares the relative error of the estimate: 2^-24.4 for
    // v_rcp_f64, 2.2e-15 after one step (measured over 4 M log-uniform b in 1e-300..1e300)
    const double r = __builtin_amdgcn_rcp(b);
    return fma(fma(-b, r, 1.0), r, r);
  } else {
    return 1.0 / b;
  }
}
__device__ __forceinline__ double fdiv(double a, double b) {
  if constexpr (kFluxFast)
    return a * frcp(b);
  else
    return a / b;
}
// 1 / sqrt(x): y <- y + y (1 - x y^2) / 2 (v_rsq_f64: 2^-24.2; one step: 4.3e-15)
__device__ __forceinline__ double frsqrt(double x) {
  if constexpr (kFluxFast) {
    const double y = __builtin_amdgcn_rsq(x);
    return fma(0.5 * y, fma(-x * y, y, 1.0), y);
  } else {
    return 1.0 / sqrt(x);
  }
}
// sqrt(x): s = x y from the estimate y of 1/sqrt(x), then s <- s + y (x - s^2) / 2
__device__ __forceinline__ double fsqrt(double x) {
  if constexpr (kFluxFast) {
    const double y = __builtin_amdgcn_rsq(x);
    const double s = x * y;
    return fma(0.5 * y, fma(-s, s, x), s);
  } else {
    return sqrt(x);
  }
}
// robust::ratio with the flux arithmetic
__device__ __forceinline__ double fratio(double a, double b) {
  const double sgn = b >= 0 ? 1.0 : -1.0;
  return fdiv(a, b + sgn * kSmall);
}

// ---- reconstruction.hpp:101-244 -------------------------------------------------------
__device__ __forceinline__ double mc_limiter(double dm, double dp, double alpha) {
  const double dc = (dm * dp > 0.0) * 0.5 * (dm + dp);
  return copysign(smin(fabs(dc), alpha * smin(fabs(dm), fabs(dp))), dc);
}

__device__ __forceinline__ void recon_plm(double qm, double q0, double qp, double &ql,
                                          double &qr) {
  double dq = qp - q0;
  dq = mc_limiter(q0 - qm, dq, 1.5);
  ql = q0 + 0.5 * dq;
  qr = q0 - 0.5 * dq;
}

// One WENO-Z edge: weights from smoothness indicators (b_near, b_mid, b_far), candidate
// stencils evaluated from the far side toward the centre.
__device__ __forceinline__ void recon_weno5z(double q0, double q1, double q2, double q3,
                                             double q4, double &ql, double &qr) {
  constexpr double c00 = 1.0 / 3.0, c01 = -7.0 / 6.0, c02 = 11.0 / 6.0;
  constexpr double c10 = -1.0 / 6.0, c11 = 5.0 / 6.0, c12 = 1.0 / 3.0;
  constexpr double c20 = 1.0 / 3.0, c21 = 5.0 / 6.0, c22 = -1.0 / 6.0;
  constexpr double g0 = 0.1, g1 = 0.6, g2 = 0.3;
  constexpr double eps = 1e-100;
  constexpr double tt = 13.0 / 3.0;
  double a = q0 - 2 * q1 + q2;
  double b = q0 - 4.0 * q1 + 3.0 * q2;
  double beta0 = tt * a * a + b * b + eps;
  a = q1 - 2.0 * q2 + q3;
  b = q3 - q1;
  double beta1 = tt * a * a + b * b + eps;
  a = q2 - 2.0 * q3 + q4;
  b = q4 - 4.0 * q3 + 3.0 * q2;
  double beta2 = tt * a * a + b * b + eps;
  const double tau5 = fabs(beta2 - beta0);
  beta0 = fdiv(beta0 + tau5, beta0);
  beta1 = fdiv(beta1 + tau5, beta1);
  beta2 = fdiv(beta2 + tau5, beta2);

  double w0 = g0 * beta0 + eps, w1 = g1 * beta1 + eps, w2 = g2 * beta2 + eps;
  double wsum = frcp(w0 + w1 + w2);
  double l = w0 * (c00 * q0 + c01 * q1 + c02 * q2);
  l += w1 * (c10 * q1 + c11 * q2 + c12 * q3);
  l += w2 * (c20 * q2 + c21 * q3 + c22 * q4);
  l *= wsum;
  const double al = fdiv(3.0 * wsum * w0 * w1 * w2, g2 * w0 * w1 + g1 * w0 * w2 + g0 * w1 * w2) + eps;

  w0 = g0 * beta2 + eps;
  w1 = g1 * beta1 + eps;
  w2 = g2 * beta0 + eps;
  wsum = frcp(w0 + w1 + w2);
  double r = w0 * (c00 * q4 + c01 * q3 + c02 * q2);
  r += w1 * (c10 * q3 + c11 * q2 + c12 * q1);
  r += w2 * (c20 * q2 + c21 * q1 + c22 * q0);
  r *= wsum;
  const double ar = fdiv(3.0 * wsum * w0 * w1 * w2, g2 * w0 * w1 + g1 * w0 * w2 + g0 * w1 * w2) + eps;

  double dq = q3 - q2;
  dq = mc_limiter(q2 - q1, dq, 2.0);
  const double alin = fdiv(2.0 * al * ar, al + ar);
  ql = alin * l + (1.0 - alin) * (q2 + 0.5 * dq);
  qr = alin * r + (1.0 - alin) * (q2 - 0.5 * dq);
}

__device__ __forceinline__ double minmod2(double a, double b) {
  return a * b > 0.0 ? (fabs(a) < fabs(b) ? a : b) : 0.0;
}

__device__ __forceinline__ double mp5_edge(double fm2, double fm1, double f0, double fp1,
                                           double fp2) {
  constexpr double alpha = 4.0, epsm = 1.e-12;
  double f = 2.0 * fm2 - 13.0 * fm1 + 47.0 * f0 + 27.0 * fp1 - 3.0 * fp2;
  f = kFluxFast ? f * (1.0 / 60.0) : f / 60.0;
  const double fmp = f0 + minmod2(fp1 - f0, alpha * (f0 - fm1));
  if ((f - f0) * (f - fmp) <= epsm) return f;
  const double d2m = fm2 + f0 - 2.0 * fm1;
  const double d2 = fm1 + fp1 - 2.0 * f0;
  const double d2p = f0 + fp2 - 2.0 * fp1;
  const double dmmp = minmod2(minmod2(4.0 * d2 - d2p, 4.0 * d2p - d2), minmod2(d2, d2p));
  const double dmmm = minmod2(minmod2(4.0 * d2m - d2, 4.0 * d2 - d2m), minmod2(d2, d2m));
  const double ful = f0 + alpha * (f0 - fm1);
  const double fav = 0.5 * (f0 + fp1);
  const double fmd = fav - 0.5 * dmmp;
  const double flc = 0.5 * (3.0 * f0 - fm1) + 4.0 / 3.0 * dmmm;
  const double fmin_ = fmax(fmin(fmin(f0, fp1), fmd), fmin(fmin(f0, ful), flc));
  const double fmax_ = fmin(fmax(fmax(f0, fp1), fmd), fmax(fmax(f0, ful), flc));
  return f + minmod2(fmin_ - f, fmax_ - f);
}

// upper edge (ql) and lower edge (qr) of the centre cell of a 5-point stencil
template <int RECON>
__device__ __forceinline__ void reconstruct(double q0, double q1, double q2, double q3,
                                            double q4, double &ql, double &qr) {
  if constexpr (RECON == PHX_RECON_CONSTANT) {
    ql = q2;
    qr = q2;
  } else if constexpr (RECON == PHX_RECON_LINEAR) {
    recon_plm(q1, q2, q3, ql, qr);
  } else if constexpr (RECON == PHX_RECON_WENO5Z) {
    recon_weno5z(q0, q1, q2, q3, q4, ql, qr);
  } else {
    ql = mp5_edge(q0, q1, q2, q3, q4);
    qr = mp5_edge(q4, q3, q2, q1, q0);
  }
}

// edges of the centre cell of q[0 .. 2H] (H = 0, 1, 2 for constant, linear, WENO5/MP5)
template <int RECON>
__device__ __forceinline__ void reconstruct_edges(const double *q, double &ql, double &qr) {
  if constexpr (RECON == PHX_RECON_CONSTANT) {
    ql = q[0];
    qr = q[0];
  } else if constexpr (RECON == PHX_RECON_LINEAR) {
    recon_plm(q[0], q[1], q[2], ql, qr);
  } else {
    reconstruct<RECON>(q[0], q[1], q[2], q[3], q[4], ql, qr);
  }
}

constexpr double kPi = 3.14159265358979323846;  // M_PI

// ---- metrics -----------------------------------------------------------------------------
// A metric policy evaluates alpha, beta^i, g_{mu nu}, gamma^{ij}, sqrt(gamma), sqrt(-g) at
// one point.  nz_cov / nz_con / nz_beta mark entries that can be nonzero; loops below
// skip the others at compile time.

struct MinkowskiMetric {  // minkowski.hpp:35-115
  static constexpr __host__ __device__ bool nz_cov(int m, int n) { return m == n; }
  static constexpr __host__ __device__ bool nz_con(int m, int n) { return m == n; }
  static constexpr __host__ __device__ bool nz_beta(int) { return false; }
  __device__ MinkowskiMetric(const DevParams &, double, double, double) {}
  __device__ double alpha() const { return 1.0; }
  __device__ double beta(int) const { return 0.0; }
  __device__ double gcov(int m, int n) const { return m == n ? (m == 0 ? -1.0 : 1.0) : 0.0; }
  __device__ double gcon(int m, int n) const { return m == n ? 1.0 : 0.0; }
  __device__ double detgamma() const { return 1.0; }
  __device__ double detg() const { return 1.0; }
};

struct SphericalKSMetric {  // spherical_kerr_schild.hpp:42-243, X = (r, theta, phi)
  static constexpr __host__ __device__ bool nz_cov(int m, int n) {
    return (m == n) || (m + n == 1) || (m + n == 3 && m != 1 && m != 2) ||
           (m + n == 4 && m != 2);
  }
  static constexpr __host__ __device__ bool nz_con(int m, int n) {
    return (m == n) || (m + n == 2 && m != 1);
  }
  static constexpr __host__ __device__ bool nz_beta(int i) { return i == 0; }
  double alpha_, beta0_, g00_, g01_, g03_, g11_, g13_, g22_, g33_;
  double c00_, c02_, c11_, c22_, detgamma_, detg_;
  __device__ SphericalKSMetric(const DevParams &p, double X1, double X2, double)
      : SphericalKSMetric(p.a, X1, X2) {}
  __device__ SphericalKSMetric(double a, double X1, double X2) {
    const double a2 = a * a;
    const double r = fabs(X1), cth = cos(X2), sth = sin(X2);
    const double sth2 = sth * sth, rho2 = r * r + a2 * cth * cth;
    alpha_ = sqrt(ratio(rho2, 2 * r + rho2));
    beta0_ = ratio(2 * r, 2 * r + rho2);
    g01_ = ratio(2 * r, rho2);
    g00_ = -1 + g01_;
    g03_ = -ratio(2 * a * r * sth2, rho2);
    g11_ = 1 + g01_;
    g13_ = -a * (1 + g01_) * sth2;
    g22_ = rho2;
    g33_ = (rho2 + a2 * (1 + g01_) * sth2) * sth2;
    c00_ = ratio(rho2, 2 * r + rho2) + ratio(a2 * sth2, rho2);
    c02_ = ratio(a, rho2);
    c11_ = ratio(1, rho2);
    c22_ = ratio(1, rho2 * sth2);
    detgamma_ = sqrt(fabs(rho2 * (2 * r + rho2) * sth2));
    detg_ = sqrt(fabs(rho2 * rho2 * sth2));
  }
  __device__ double alpha() const { return alpha_; }
  __device__ double beta(int i) const { return i == 0 ? beta0_ : 0.0; }
  __device__ double gcov(int m, int n) const {
    const int lo = m < n ? m : n, hi = m < n ? n : m;
    if (lo == 0) return hi == 0 ? g00_ : hi == 1 ? g01_ : hi == 3 ? g03_ : 0.0;
    if (lo == 1) return hi == 1 ? g11_ : hi == 3 ? g13_ : 0.0;
    if (lo == 2) return hi == 2 ? g22_ : 0.0;
    return g33_;
  }
  __device__ double gcon(int m, int n) const {
    const int lo = m < n ? m : n, hi = m < n ? n : m;
    if (lo == 0) return hi == 0 ? c00_ : hi == 2 ? c02_ : 0.0;
    if (lo == 1) return hi == 1 ? c11_ : 0.0;
    return hi == 2 ? c22_ : 0.0;
  }
  __device__ double detgamma() const { return detgamma_; }
  __device__ double detg() const { return detg_; }
};

// McKinney-Gammie-Ryan map X -> (r, theta, phi) and its Jacobian
// (mckinney_gammie_ryan.hpp:66-108, 143-197): J = [[r1, 0, 0], [t1, t2, 0], [0, 0, 1]]
// with r1 = dr/dX1, t1 = dtheta/dX1, t2 = dtheta/dX2.
struct MgrMap {
  double r, th, r1, t1, t2;
  __device__ MgrMap(const DevParams::Mgr &m, double X1, double X2) {
    const double dr = X1 - m.hexp_br;
    const double on = dr > 0. ? 1. : 0.;
    r = exp(X1 + on * m.hexp_csq * pow(dr, m.hexp_nsq));
    r1 = exp(X1 + on * m.hexp_csq * pow(dr, m.hexp_nsq)) *
         (1. + on * m.hexp_csq * m.hexp_nsq * pow(dr, m.hexp_nsq - 1.));
    const double thG = kPi * X2 + ((1. - m.h) / 2.) * sin(2. * kPi * X2);
    const double dthGdX2 = kPi + kPi * (1 - m.h) * cos(2 * kPi * X2);
    if (m.derefine_poles) {
      const double y = 2. * X2 - 1.;
      const double py = pow(y / m.xt, m.alpha);
      const double thJ = m.norm * y * (1. + py / (m.alpha + 1.)) + 0.5 * kPi;
      const double e = exp(m.smooth * (m.x0 - X1));
      th = thG + e * (thJ - thG);
      t1 = -m.smooth * (thJ - thG) * e;
      t2 = dthGdX2 + e * (m.norm * (1 + py) * 2. - dthGdX2);
    } else {
      th = thG;
      t1 = 0.0;
      t2 = dthGdX2;
    }
    constexpr double kSingSmall = 1.e-20;  // keep theta off the poles
    if (fabs(th) < kSingSmall) th = th >= 0. ? kSingSmall : -kSingSmall;
    if (fabs(kPi - th) < kSingSmall) th = th >= kPi ? kPi + kSingSmall : kPi - kSingSmall;
  }
};

// FMKS (modified_system.hpp:63-168): spherical Kerr-Schild at (r, theta) contracted with
// the MGR Jacobian, term by term in the reference's summation order (zero terms elided).
struct FMKSMetric {
  static constexpr __host__ __device__ bool nz_cov(int m, int n) {
    return !((m + n == 2 && m != 1) || m + n == 5);  // g_{02} = g_{23} = 0
  }
  static constexpr __host__ __device__ bool nz_con(int, int) { return true; }
  static constexpr __host__ __device__ bool nz_beta(int i) { return i < 2; }
  double alpha_, b0_, b1_, g00_, g01_, g03_, g11_, g12_, g13_, g22_, g33_;
  double c00_, c01_, c02_, c11_, c12_, c22_, detgamma_, detg_;
  // Jcon = J^-1: [[1/r1, 0, 0], [-t1/(r1 t2), 1/t2, 0], [0, 0, 1]]
  double j00_, j10_, j11_;
  __device__ FMKSMetric(const DevParams &p, double X1, double X2, double) {
    const MgrMap m(p.mgr, X1, X2);
    const SphericalKSMetric k(p.a, m.r, m.th);
    j00_ = ratio(1., m.r1);
    j10_ = -ratio(m.t1, m.r1 * m.t2);
    j11_ = ratio(1., m.t2);
    alpha_ = k.alpha_;
    b0_ = j00_ * k.beta0_;
    b1_ = j10_ * k.beta0_;
    g00_ = k.g00_;
    g01_ = k.g01_ * m.r1;
    g03_ = k.g03_;
    g11_ = k.g11_ * m.r1 * m.r1 + k.g22_ * m.t1 * m.t1;
    g12_ = k.g22_ * m.t1 * m.t2;
    g13_ = k.g13_ * m.r1;
    g22_ = k.g22_ * m.t2 * m.t2;
    g33_ = k.g33_;
    c00_ = k.c00_ * j00_ * j00_;
    c01_ = k.c00_ * j00_ * j10_;
    c02_ = k.c02_ * j00_;
    c11_ = k.c00_ * j10_ * j10_ + k.c11_ * j11_ * j11_;
    c12_ = k.c02_ * j10_;
    c22_ = k.c22_;
    const double detJ = m.r1 * m.t2;  // linear_algebra.hpp:50-54 on a triangular J
    detgamma_ = fabs(k.detgamma_ * detJ);
    detg_ = fabs(k.detg_ * detJ);
  }
  __device__ double alpha() const { return alpha_; }
  __device__ double beta(int i) const { return i == 0 ? b0_ : i == 1 ? b1_ : 0.0; }
  __device__ double gcov(int m, int n) const {
    const int lo = m < n ? m : n, hi = m < n ? n : m;
    if (lo == 0) return hi == 0 ? g00_ : hi == 1 ? g01_ : hi == 3 ? g03_ : 0.0;
    if (lo == 1) return hi == 1 ? g11_ : hi == 2 ? g12_ : g13_;
    if (lo == 2) return hi == 2 ? g22_ : 0.0;
    return g33_;
  }
  __device__ double gcon(int m, int n) const {
    const int lo = m < n ? m : n, hi = m < n ? n : m;
    if (lo == 0) return hi == 0 ? c00_ : hi == 1 ? c01_ : c02_;
    if (lo == 1) return hi == 1 ? c11_ : c12_;
    return c22_;
  }
  __device__ double detgamma() const { return detgamma_; }
  __device__ double detg() const { return detg_; }
  // the metric as 19 numbers t[q * s], q = 0..18 (a cached geometry plane)
  __device__ FMKSMetric() {}
  __device__ void store(double *t, long s) const {
    const double v[19] = {alpha_, b0_, b1_, g00_, g01_, g03_, g11_, g12_, g13_, g22_,
                          g33_, c00_, c01_, c02_, c11_, c12_, c22_, detgamma_, detg_};
#pragma unroll
    for (int q = 0; q < 19; ++q) t[q * s] = v[q];
  }
  static __device__ FMKSMetric load(const double *__restrict__ t, long s) {
    FMKSMetric g;
    g.alpha_ = t[0]; g.b0_ = t[s]; g.b1_ = t[2 * s]; g.g00_ = t[3 * s]; g.g01_ = t[4 * s];
    g.g03_ = t[5 * s]; g.g11_ = t[6 * s]; g.g12_ = t[7 * s]; g.g13_ = t[8 * s];
    g.g22_ = t[9 * s]; g.g33_ = t[10 * s]; g.c00_ = t[11 * s]; g.c01_ = t[12 * s];
    g.c02_ = t[13 * s]; g.c11_ = t[14 * s]; g.c12_ = t[15 * s]; g.c22_ = t[16 * s];
    g.detgamma_ = t[17 * s]; g.detg_ = t[18 * s];
    g.j00_ = g.j10_ = g.j11_ = 0.0;
    return g;
  }
};

template <int GEOM>
struct MetricOf;
template <>
struct MetricOf<PHX_GEOM_MINKOWSKI> { using type = MinkowskiMetric; };
template <>
struct MetricOf<PHX_GEOM_SPHERICAL_KS> { using type = SphericalKSMetric; };
template <>
struct MetricOf<PHX_GEOM_FMKS> { using type = FMKSMetric; };

// the cached geometry plane: 19 metric numbers, then (centres only) the x1 and x2
// derivatives of the 10 g_{mu nu} (upper triangle, row by row), d ln alpha / dX1, dX2, and
// the 10 g^{mu nu}
constexpr int kGeoMetric = 19, kGeoCentre = 19 + 20 + 2 + 10;
constexpr int kGeoPlanes = kGeoCentre + 2 * kGeoMetric;

// The metric of cell-plane point q = j * ni + i of block b at location loc (0 cell centre,
// 1 x1 face, 2 x2 face; an x3 face has the centre's metric when it does not depend on X3):
// from the block's cached plane when it has one, else evaluated at X.  CACHED = true
// promises the plane (packs of FMKS blocks always have one), so no evaluation code is
// compiled in.
template <int GEOM, bool CACHED = false>
__device__ __forceinline__ typename MetricOf<GEOM>::type metric_at(const DevParams &p,
                                                                   const BlockGeom &b, int loc,
                                                                   long q, double X1, double X2,
                                                                   double X3) {
  using M = typename MetricOf<GEOM>::type;
  if constexpr (GEOM == PHX_GEOM_FMKS) {
    const long plane = (long)b.ni * b.nj;
    const long off = loc == 0 ? 0 : loc == 1 ? kGeoCentre : kGeoCentre + kGeoMetric;
    if constexpr (CACHED) {
      return M::load(b.geo + off * plane + q, plane);
    } else {
      if (b.geo) return M::load(b.geo + off * plane + q, plane);
      return M(p, X1, X2, X3);
    }
  } else {
    return M(p, X1, X2, X3);
  }
}

// ---- bounds: fixup.hpp:91-134, 188-213 ------------------------------------------------
__device__ __forceinline__ void get_floors(const DevBounds &b, double x1, double x2,
                                           double x3, double &rflr, double &sflr) {
  switch (b.floor_type) {
  case PHX_FLOOR_CONSTANT_RHO_SIE:
    rflr = b.r0;
    sflr = b.s0;
    break;
  case PHX_FLOOR_EXPX1_RHO_SIE:
    rflr = b.r0 * exp(b.ralpha * x1);
    sflr = b.s0 * exp(b.salpha * x1);
    break;
  case PHX_FLOOR_EXPX1_RHO_U: {
    const double scratch = b.r0 * exp(b.ralpha * x1);
    sflr = b.s0 * exp(b.salpha * x1) / smax(rflr, scratch);
    rflr = scratch;
  } break;
  case PHX_FLOOR_X1_RHO_SIE:
    rflr = b.r0 * smin(1.0, pow(x1, b.ralpha));
    sflr = b.s0 * smin(1.0, pow(x1, b.salpha));
    break;
  default: {
    const double r = sqrt(x1 * x1 + x2 * x2 + x3 * x3);
    rflr = b.r0 * smin(1.0, pow(r, b.ralpha));
    sflr = b.s0 * smin(1.0, pow(r, b.salpha));
  } break;
  }
  rflr = smax(rflr, b.rho_min_eos);
  sflr = smax(sflr, b.sie_min_eos);
}

// ---- ideal-gas EOS (singularity-eos IdealGas closed forms, unit scales 1) -------------
__device__ __forceinline__ double eos_p_rho_sie(const DevParams &p, double rho, double sie) {
  return p.gm1 * rho * sie;
}
__device__ __forceinline__ double eos_t_rho_sie(const DevParams &p, double sie) {
  return sie / p.cv;
}
__device__ __forceinline__ double eos_p_rho_t(const DevParams &p, double rho, double t) {
  return p.gm1 * rho * p.cv * t;
}
__device__ __forceinline__ double eos_bmod_rho_t(const DevParams &p, double rho, double t) {
  return (p.gm1 + 1.0) * p.gm1 * rho * p.cv * t;
}

template <bool MHD>
struct Layout {
  static constexpr int nprim = MHD ? 10 : 7;
  static constexpr int ncons = MHD ? 8 : 5;
  static constexpr int prs = MHD ? 8 : 5;
  static constexpr int gm1 = MHD ? 9 : 6;
};

// ---- riemann.hpp:88-209 FluxState::prim_to_flux (Valencia energy) ---------------------
template <bool MHD, class M>
__device__ __forceinline__ void prim_to_flux(const DevParams &p, const int dir, const M &g,
                                             const double X[3], const double *q,
                                             double &vm, double &vp, double *U, double *F) {
  using L = Layout<MHD>;
  double rho_floor = q[0], sie_floor;
  get_floors(p.bnd, X[0], X[1], X[2], rho_floor, sie_floor);
  const double gam_max = p.bnd.gam_max, sie_max = p.bnd.sie_max;
  const double rho = smax(q[0], rho_floor);
  double vpcon[3] = {q[1], q[2], q[3]};
  double vsq0 = 0.;
#pragma unroll
  for (int ii = 0; ii < 3; ++ii)
#pragma unroll
    for (int jj = 0; jj < 3; ++jj)
      if (M::nz_cov(ii + 1, jj + 1)) vsq0 += g.gcov(ii + 1, jj + 1) * vpcon[ii] * vpcon[jj];
  double W = fsqrt(1. + vsq0);
  if (W > gam_max) {
    const double rescale = fsqrt(fdiv(gam_max * gam_max - 1., W * W - 1.));
#pragma unroll
    for (int ii = 0; ii < 3; ++ii) vpcon[ii] *= rescale;
  }
  double vcon[3] = {fdiv(vpcon[0], W), fdiv(vpcon[1], W), fdiv(vpcon[2], W)};
  const double vel = vcon[dir];
  double Bcon[3] = {0.0, 0.0, 0.0};
  const double u = (q[4] / rho > sie_floor ? (q[4] / rho > sie_max ? rho * sie_max : q[4])
                                            : rho * sie_floor);
  const double P = smax(q[L::prs], 0.0);
  const double gamma1 = q[L::gm1];
  if constexpr (MHD) {
    Bcon[0] = q[5];
    Bcon[1] = q[6];
    Bcon[2] = q[7];
  }
  double BdotB = 0.0, Bdotv = 0.0, vsq = 0.0;
#pragma unroll
  for (int m = 0; m < 3; m++)
#pragma unroll
    for (int n = 0; n < 3; n++) {
      if (!M::nz_cov(m + 1, n + 1)) continue;
      vsq += g.gcov(m + 1, n + 1) * vcon[m] * vcon[n];
      if constexpr (MHD) {
        Bdotv += g.gcov(m + 1, n + 1) * Bcon[m] * vcon[n];
        BdotB += g.gcov(m + 1, n + 1) * Bcon[m] * Bcon[n];
      }
    }
  const double vsq_max = 1.0 - 1.0 / (gam_max * gam_max);
  if (vsq > vsq_max) {
    double scale = fdiv(vsq_max, vsq);
    vsq *= scale;
    scale = fsqrt(scale);
    Bdotv *= scale;
#pragma unroll
    for (int m = 0; m < 3; m++) vcon[m] *= scale;
  }
  W = frsqrt(1 - vsq);
  const double ialpha_beta_dir = M::nz_beta(0) || M::nz_beta(1) || M::nz_beta(2)
                                     ? fdiv(g.beta(dir), g.alpha()) : 0.0;
  const double vtil = vel - ialpha_beta_dir;

  U[0] = rho * W;
  F[0] = U[0] * vtil;
  const double b0 = W * Bdotv;
  const double bsq = fdiv(BdotB + b0 * b0, W * W);
  const double rhohWsq = (rho + u + P) * W * W;
  const double bsqWsq = bsq * W * W;
#pragma unroll
  for (int m = 0; m < 3; m++) {
    double bcovm = 0.0;
    if (MHD && M::nz_cov(m + 1, 0)) bcovm = fdiv(g.gcov(m + 1, 0) * b0, g.alpha());
    double vcovm = 0.0;
#pragma unroll
    for (int n = 1; n < 4; n++) {
      if (!M::nz_cov(m + 1, n)) continue;
      if constexpr (MHD) {
        const double sh = M::nz_beta(n - 1) ? fdiv(g.beta(n - 1), g.alpha()) : 0.0;
        bcovm += g.gcov(m + 1, n) * (fdiv(Bcon[n - 1], W) + b0 * (vcon[n - 1] - sh));
      }
      vcovm += g.gcov(m + 1, n) * vcon[n - 1];
    }
    U[1 + m] = (rhohWsq + bsqWsq) * vcovm - b0 * bcovm;
    F[1 + m] = U[1 + m] * vtil + (P + 0.5 * bsq) * (double)(dir == m) - fdiv(bcovm * Bcon[dir], W);
  }
  U[4] = rhohWsq + bsqWsq - (P + 0.5 * bsq) - b0 * b0 - U[0];
  F[4] = U[4] * vtil + (P + 0.5 * bsq) * vel - Bdotv * Bcon[dir];
  if constexpr (MHD) {
#pragma unroll
    for (int m = 5; m <= 7; m++) {
      const double sh = M::nz_beta(m - 5) ? fdiv(g.beta(m - 5), g.alpha()) : 0.0;
      U[m] = Bcon[m - 5];
      F[m] = U[m] * vtil - Bcon[dir] * (vcon[m - 5] - sh);
    }
  }
  const double vasq = fdiv(bsqWsq, make_positive(rhohWsq + bsqWsq));
  const double cssq = make_bounded(fdiv(gamma1 * P * W * W, make_positive(rhohWsq)), 0.0, 1.0);
  const double cmsq = make_bounded(cssq + vasq * (1.0 - cssq), 0.0, 1.0);
  const double vcoff = fdiv(g.alpha(), 1.0 - vsq * cmsq);
  const double v0 = vel * (1.0 - cmsq);
  const double vpm =
      fsqrt(make_positive(cmsq * (1.0 - vsq) * (g.gcon(dir, dir) * (1.0 - vsq * cmsq) - vel * v0)));
  const double bd = M::nz_beta(dir) ? g.beta(dir) : 0.0;
  vp = vcoff * (v0 + vpm) - bd;
  vm = vcoff * (v0 - vpm) - bd;
}

// riemann.hpp:257-301.  Writes ncons fluxes (times sqrt(-g) / sqrt(gamma) as the reference)
// and returns the face signal speed.
template <bool MHD, int RIEMANN, class M>
__device__ __forceinline__ double riemann_flux(const DevParams &p, int dir, const M &g,
                                               const double X[3], const double *ql,
                                               const double *qr, double *flux) {
  constexpr int NC = Layout<MHD>::ncons;
  double Ul[NC], Ur[NC], Fl[NC], Fr[NC], vml, vpl, vmr, vpr;
  prim_to_flux<MHD>(p, dir, g, X, ql, vml, vpl, Ul, Fl);
  prim_to_flux<MHD>(p, dir, g, X, qr, vmr, vpr, Ur, Fr);
  if constexpr (RIEMANN == PHX_RIEMANN_LLF) {
    const double cmax = smax(smax(-vml, vpl), smax(-vmr, vpr));
#pragma unroll
    for (int m = 0; m < NC; m++)
      flux[m] = 0.5 * ((Fl[m] + Fr[m]) * g.detg() - cmax * ((Ur[m] - Ul[m]) * g.detgamma()));
    return cmax;
  } else {
    const double cl = smin(smin(vml, vmr), 0.0);
    const double cr = smax(smax(vpl, vpr), 0.0);
    const double crcl = cr * cl;
    const double inv = fratio(1.0, cr - cl);
#pragma unroll
    for (int m = 0; m < NC; m++)
      flux[m] = ((cr * Fl[m] - cl * Fr[m]) * g.detg() + crcl * (Ur[m] - Ul[m]) * g.detgamma()) * inv;
    return smax(-cl, cr);
  }
}

// ---- prim2con.hpp:25-102 ---------------------------------------------------------------
template <class M>
__device__ __forceinline__ void signal_speeds(double rho, double u, double pr, double bsq,
                                              const double v[3], double vsq, double gam1,
                                              const M &g, double sig[3]) {
  const double rho_rel = rho + u + pr;
  const double vasq = bsq / (rho_rel + bsq);
  double cssq = gam1 * pr / rho_rel;
  cssq += vasq - cssq * vasq;
  const double vcoff = g.alpha() / (1.0 - vsq * cssq);
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const double vpm = sqrt(cssq * (1.0 - vsq) *
                            (g.gcon(m, m) * (1.0 - vsq * cssq) - v[m] * v[m] * (1.0 - cssq)));
    const double bm = M::nz_beta(m) ? g.beta(m) : 0.0;
    const double vp = vcoff * (v[m] * (1.0 - cssq) + vpm) - bm;
    const double vmn = vcoff * (v[m] * (1.0 - cssq) - vpm) - bm;
    sig[m] = smax(fabs(vp), fabs(vmn));
  }
}

template <bool MHD, class M>
__device__ __forceinline__ void p2c(const double rho, const double vp[3], const double b[3],
                                    const double u, const double pr, const double gam1,
                                    const M &g, double &D, double S[3], double bcons[3],
                                    double &tau, double sig[3]) {
  double vsq = 0.0, Bsq = 0.0, Bdotv = 0.0, vsq0 = 0.0;
#pragma unroll
  for (int ii = 0; ii < 3; ++ii)
#pragma unroll
    for (int jj = 0; jj < 3; ++jj)
      if (M::nz_cov(ii + 1, jj + 1)) vsq0 += g.gcov(ii + 1, jj + 1) * vp[ii] * vp[jj];
  const double W = sqrt(1. + vsq0);
  const double iW = 1.0 / W;
  const double v[3] = {vp[0] * iW, vp[1] * iW, vp[2] * iW};
#pragma unroll
  for (int ii = 0; ii < 3; ++ii)
#pragma unroll
    for (int jj = 0; jj < 3; ++jj) {
      if (!M::nz_cov(ii + 1, jj + 1)) continue;
      vsq += g.gcov(ii + 1, jj + 1) * v[ii] * v[jj];
      if constexpr (MHD) {
        Bsq += g.gcov(ii + 1, jj + 1) * b[ii] * b[jj];
        Bdotv += g.gcov(ii + 1, jj + 1) * b[ii] * v[jj];
      }
    }
  const double b0 = W * Bdotv;
  double bcon[4] = {0.0, 0.0, 0.0, 0.0};
  double bsq = 0.0;
  double bcov[3] = {0.0, 0.0, 0.0};
  if constexpr (MHD) {
    bcon[0] = W * Bdotv / g.alpha();
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const double sh = M::nz_beta(m) ? g.beta(m) / g.alpha() : 0.0;
      bcon[m + 1] = b[m] * iW + Bdotv * W * (v[m] - sh);
    }
    bsq = Bsq * iW * iW + Bdotv * Bdotv;
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int n = 0; n < 4; ++n)
        if (M::nz_cov(m + 1, n)) bcov[m] += g.gcov(m + 1, n) * bcon[n];
  }
  D = g.detgamma() * rho * W;
  const double rho_rel = (rho + u + pr + bsq) * W * W;
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    double vcov = 0.0;
#pragma unroll
    for (int n = 0; n < 3; ++n)
      if (M::nz_cov(m + 1, n + 1)) vcov += g.gcov(m + 1, n + 1) * v[n];
    S[m] = g.detgamma() * (rho_rel * vcov - b0 * bcov[m]);
    bcons[m] = g.detgamma() * b[m];
  }
  tau = g.detgamma() * (rho_rel - (pr + 0.5 * bsq) - b0 * b0) - D;
  signal_speeds(rho, u, pr, bsq, v, vsq, gam1, g, sig);
}

// ---- con2prim_robust.hpp:46-180 Residual ----------------------------------------------
struct Residual {
  const DevParams *p;
  double D, q, bsq, bsq_rpsq, rsq, rbsq, v0sq;
  double x1, x2, x3;
  double rho_floor, e_floor, e_max;
  bool used_density_floor, used_energy_floor, used_energy_max, used_gamma_max;

  __device__ __forceinline__ double x_mu(double mu) const { return ratio(1.0, 1.0 + mu * bsq); }
  __device__ __forceinline__ double rbarsq_mu(double mu, double x) const {
    return x * (x * rsq + mu * (1.0 + x) * rbsq);
  }
  __device__ __forceinline__ double qbar_mu(double mu, double x) const {
    const double mux = mu * x;
    return q - 0.5 * (bsq + mux * mux * bsq_rpsq);
  }
  __device__ __forceinline__ double vhatsq_mu(double mu, double rbarsq) {
    const double vsq_trial = mu * mu * rbarsq;
    used_gamma_max = vsq_trial > v0sq;
    return used_gamma_max ? v0sq : vsq_trial;
  }
  __device__ __forceinline__ double rhohat_mu(double iWhat) {
    const double rho_trial = D * iWhat;
    used_density_floor = rho_trial <= rho_floor;
    return used_density_floor ? rho_floor : rho_trial;
  }
  __device__ __forceinline__ double ehat_mu(double mu, double qbar, double rbarsq,
                                            double vhatsq, double What) {
    double rho = rhohat_mu(ratio(1.0, What));
    get_floors(p->bnd, x1, x2, x3, rho, e_floor);
    e_floor *= p->fsf;
    const double ehat_trial =
        What * (qbar - mu * rbarsq) + vhatsq * What * ratio(What, 1.0 + What);
    used_energy_floor = false;
    used_energy_max = false;
    if (ehat_trial <= e_floor) {
      used_energy_floor = true;
      return e_floor;
    } else if (ehat_trial > e_max) {
      used_energy_max = true;
      return e_max;
    }
    return ehat_trial;
  }
  __device__ __forceinline__ double operator()(double mu) {  // :121-141
    const double x = x_mu(mu);
    const double rbarsq = rbarsq_mu(mu, x);
    const double qbar = qbar_mu(mu, x);
    const double vhatsq = vhatsq_mu(mu, rbarsq);
    const double iWhat = sqrt(1.0 - vhatsq);
    const double What = ratio(1.0, iWhat);
    const double rhohat = rhohat_mu(iWhat);
    const double ehat = ehat_mu(mu, qbar, rbarsq, vhatsq, What);
    const double Phat = eos_p_rho_sie(*p, rhohat, ehat);
    const double hhat = rhohat * (1.0 + ehat) + Phat;
    const double ahat = Phat / make_positive(hhat - Phat);
    const double nua = (1.0 + ahat) * (1.0 + ehat) * iWhat;
    const double nub = (1.0 + ahat) * (1.0 + qbar - mu * rbarsq);
    const double nuhat = smax(nua, nub);
    const double muhat = 1.0 / (nuhat + mu * rbarsq);
    return mu - muhat;
  }
};

// root_find.hpp:41-58 (refine_bracket) and :109-167 (regula_falsi, Illinois variant).
// Returns false when the root is not bracketed (the reference aborts there).
template <class F>
__device__ __forceinline__ bool regula_falsi(F &f, double a, double b, double tol,
                                             double guess, int max_iter, double &root,
                                             int &iters) {
  double ya, yb;
  int it = 0;
  if (a <= guess && b >= guess) {
    ya = f(guess);
    yb = f(b);
    if (ya * yb > 0.0) {
      yb = ya;
      b = guess;
      ya = f(a);
    } else {
      a = guess;
    }
  } else {
    ya = f(a);
    yb = f(b);
  }
  if (ya * yb > 0.0) {
    iters = 0;
    root = guess;
    return false;
  }
  const double sign = (ya < 0 ? 1.0 : -1.0);
  ya *= sign;
  yb *= sign;
  int b1 = 0, b2 = 0;
  while (b - a > 2.0 * tol && it < max_iter) {
    const double c = (a * yb - b * ya) / (yb - ya);
    if (c == a) {
      b = a;
      it++;
      continue;
    } else if (c == b) {
      a = b;
      it++;
      continue;
    }
    const double yc = sign * f(c);
    if (yc > 0.0) {
      b = c;
      yb = yc;
      b1++;
      ya *= (b1 > 1 ? 0.5 : 1.0);
      b2 = 0;
    } else if (yc < 0.0) {
      a = c;
      ya = yc;
      b2++;
      yb *= (b2 > 1 ? 0.5 : 1.0);
      b1 = 0;
    } else {
      a = c;
      b = c;
    }
    it++;
  }
  iters = it;
  root = 0.5 * (a + b);
  return true;
}

// root_find.hpp:123-166 on a bracket [a, b] whose residuals ya, yb are already known
// (ya * yb <= 0); the loop and its iteration count are the reference's.
template <class F>
__device__ __forceinline__ double regula_falsi_bracketed(F &f, double a, double b, double ya,
                                                         double yb, double tol, int max_iter,
                                                         int &iters) {
  int it = 0;
  const double sign = (ya < 0 ? 1.0 : -1.0);
  ya *= sign;
  yb *= sign;
  int b1 = 0, b2 = 0;
  while (b - a > 2.0 * tol && it < max_iter) {
    const double c = (a * yb - b * ya) / (yb - ya);
    if (c == a) {
      b = a;
      it++;
      continue;
    } else if (c == b) {
      a = b;
      it++;
      continue;
    }
    const double yc = sign * f(c);
    if (yc > 0.0) {
      b = c;
      yb = yc;
      b1++;
      ya *= (b1 > 1 ? 0.5 : 1.0);
      b2 = 0;
    } else if (yc < 0.0) {
      a = c;
      ya = yc;
      b2++;
      yb *= (b2 > 1 ? 0.5 : 1.0);
      b1 = 0;
    } else {
      a = c;
      b = c;
    }
    it++;
  }
  iters = it;
  return 0.5 * (a + b);
}

// con2prim_robust.hpp:299-453 for one zone, in three pieces so that a bracket predictor
// can run between the set-up and the root find: c2p_setup (:303-367), the root find
// (:374-376), c2p_unwrap (:382-452).
struct C2PState {
  Residual res;
  double rcon[3], bu[3], bdotr;
};

// :303-367.  Returns false for the NaN early-out (the zone keeps every field).
template <bool MHD, class M>
__device__ __forceinline__ bool c2p_setup(const DevParams &p, const M &g, const double X[3],
                                          const double *cons, double *prim, C2PState &z) {
  if (isnan(cons[0]) + isnan(cons[1]) + isnan(cons[4]) > 0) return false;
  const double igdet = 1.0 / g.detgamma();
  double rhoflr = 0.0, epsflr;
  get_floors(p.bnd, X[0], X[1], X[2], rhoflr, epsflr);
  rhoflr *= p.fsf;
  const double gam_max = p.bnd.gam_max;
  const double D = cons[0] * igdet;
  const double tau = cons[4] * igdet;
  const double q = ratio(tau, D);
  double bsq = 0.0, bsq_rpsq = 0.0, rbsq = 0.0, rsq = 0.0, bdotr = 0.0;
  double rcov[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) rcov[i] = ratio(cons[1 + i], cons[0]);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    z.rcon[i] = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j)
      if (M::nz_con(i, j)) z.rcon[i] += g.gcon(i, j) * rcov[j];
    rsq += z.rcon[i] * rcov[i];
  }
  z.bu[0] = z.bu[1] = z.bu[2] = 0.0;
  if constexpr (MHD) {
#pragma unroll
    for (int i = 0; i < 3; ++i) prim[5 + i] = cons[5 + i] * igdet;
    const double sD = 1.0 / sqrt(smax(D, rhoflr));
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      z.bu[i] = prim[5 + i] * sD;
      bdotr += z.bu[i] * rcov[i];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        if (M::nz_cov(i + 1, j + 1)) bsq += g.gcov(i + 1, j + 1) * z.bu[i] * z.bu[j];
    bsq = smax(0.0, bsq);
    rbsq = bdotr * bdotr;
    bsq_rpsq = bsq * rsq - rbsq;
  }
  z.bdotr = bdotr;
  const double zsq = rsq / 1.0;
  const double v0sq = smin(zsq / (1.0 + zsq), 1.0 - 1.0 / (gam_max * gam_max));
  Residual &res = z.res;
  res.p = &p;
  res.D = D; res.q = q; res.bsq = bsq; res.bsq_rpsq = bsq_rpsq; res.rsq = rsq;
  res.rbsq = rbsq; res.v0sq = v0sq; res.x1 = X[0]; res.x2 = X[1]; res.x3 = X[2];
  {
    double garbage = 0.0;
    res.rho_floor = 0.0;
    get_floors(p.bnd, X[0], X[1], X[2], res.rho_floor, garbage);
    res.e_max = p.bnd.sie_max;
    res.rho_floor *= p.fsf;
  }
  res.e_floor = 0.0;
  res.used_density_floor = res.used_energy_floor = res.used_energy_max = res.used_gamma_max = false;
  return true;
}

// :382-452 after the root mu; returns the fail-flag value
template <bool MHD, class M>
__device__ __forceinline__ double c2p_unwrap(const DevParams &p, const M &g, C2PState &z,
                                             const double mu, double *cons, double *prim,
                                             double &temp, double sig[3]) {
  using L = Layout<MHD>;
  Residual &res = z.res;
  const double x = res.x_mu(mu);
  const double rbarsq = res.rbarsq_mu(mu, x);
  const double vsq = res.vhatsq_mu(mu, rbarsq);
  const double iW = sqrt(1.0 - vsq);
  prim[0] = res.rhohat_mu(iW);
  const double qbar = res.qbar_mu(mu, x);
  const double W = 1.0 / iW;
  prim[4] = res.ehat_mu(mu, qbar, rbarsq, vsq, W);
  temp = eos_t_rho_sie(p, prim[4]);
  prim[4] *= prim[0];
  prim[L::prs] = eos_p_rho_t(p, prim[0], temp);
  prim[L::gm1] = eos_bmod_rho_t(p, prim[0], temp) / prim[L::prs];
  double vel[3], bu[3] = {z.bu[0], z.bu[1], z.bu[2]};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    vel[i] = W * mu * x * (z.rcon[i] + mu * z.bdotr * bu[i]);
    prim[1 + i] = vel[i];
  }
  if constexpr (MHD) {
#pragma unroll
    for (int i = 0; i < 3; ++i) bu[i] = prim[5 + i];
  }
  if (res.used_density_floor || res.used_energy_floor || res.used_energy_max ||
      res.used_gamma_max) {
    double S[3], bcons[3];
    p2c<MHD>(prim[0], vel, bu, prim[4], prim[L::prs], prim[L::gm1], g, cons[0], S, bcons,
             cons[4], sig);
#pragma unroll
    for (int i = 0; i < 3; ++i) cons[1 + i] = S[i];
    if (isnan(cons[0]) + isnan(cons[1]) + isnan(cons[4]) > 0) return 0.0;
  } else {
#pragma unroll
    for (int m = 0; m < 3; ++m) vel[m] *= iW;
    double bsq2 = 0.0;
    if constexpr (MHD) {
      double bdotv = 0.0;
#pragma unroll
      for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int n = 0; n < 3; ++n) {
          if (!M::nz_cov(m + 1, n + 1)) continue;
          bsq2 += bu[m] * bu[n] * g.gcov(m + 1, n + 1);
          bdotv += bu[m] * vel[n] * g.gcov(m + 1, n + 1);
        }
      bsq2 = iW * iW * bsq2 + bdotv * bdotv;
    }
    signal_speeds(prim[0], prim[4], prim[L::prs], bsq2, vel, vsq, prim[L::gm1], g, sig);
  }
  // (the NaN test of the reference, :437-441, can only fire in the branch above: here the
  // conserved variables are those c2p_setup found free of NaNs)
  if (p.fail_on_floors && (res.used_density_floor || res.used_energy_floor)) return 0.0;
  if (p.fail_on_ceilings && res.used_gamma_max) return 0.0;
  return 1.0;
}

// the whole solve with the reference's root find: regula_falsi on [0, 1] split at the
// guess mu_io.  iters = -1 marks an unbracketed root.
template <bool MHD, class M>
__device__ __forceinline__ double con2prim_zone(const DevParams &p, const M &g,
                                                const double X[3], double *cons,
                                                double *prim, double &temp, double &mu_io,
                                                double sig[3], int &iters) {
  iters = 0;
  C2PState z;
  if (!c2p_setup<MHD>(p, g, X, cons, prim, z)) return 0.0;
  double mu;
  if (!regula_falsi(z.res, 0.0, 1.0, p.tol, mu_io, p.max_iter, mu, iters)) {
    iters = -1;
    return 0.0;
  }
  mu_io = mu;
  return c2p_unwrap<MHD>(p, g, z, mu, cons, prim, temp, sig);
}

// con2prim_zone on a zone's fields in memory (component planes, cell c; the zone-centre
// metric of geometry-plane point q at X), for conserved variables cs that passed the NaN
// test: it writes the magnetic primitives of the set-up, then, when the root is bracketed,
// the other primitives, the temperature and the root, and the conserved variables only
// where a floor or ceiling re-derived them (elsewhere they are unchanged).  The input
// primitives and temperature are never read (the ideal-gas inversion does not use them as a
// guess).  With CACHED the metric is read from the block's cached geometry for the set-up
// and again for the unwrap, so it is not held in registers across the solve.  solved tells
// whether sig holds new signal speeds; iters is -1 for an unbracketed root.
template <bool MHD, int GEOM, bool CACHED, class PP, class CP>
__device__ __forceinline__ double con2prim_fields(const DevParams &p, const BlockGeom &b, long q,
                                                  const double X[3], double *cs,
                                                  const PP &prim,
                                                  double *__restrict__ temp,
                                                  double *__restrict__ mu,
                                                  const CP &cons, long c,
                                                  double sig[3], int &iters, bool &solved) {
  using M = typename MetricOf<GEOM>::type;
  using L = Layout<MHD>;
  iters = 0;
  C2PState z;
  double pr[L::nprim];
  const M g = metric_at<GEOM, CACHED>(p, b, 0, q, X[0], X[1], X[2]);
  c2p_setup<MHD>(p, g, X, cs, pr, z);
  if constexpr (MHD) {
#pragma unroll
    for (int i = 5; i < 8; ++i) prim[i][c] = pr[i];
  }
  double root;
  solved = regula_falsi(z.res, 0.0, 1.0, p.tol, mu[c], p.max_iter, root, iters);
  if (!solved) {
    iters = -1;
    return 0.0;
  }
  mu[c] = root;
  double tt = 0.0, st;
  if constexpr (CACHED) {
    asm volatile("" ::: "memory");  // compiler-only: the metric is read again below
    const M g2 = metric_at<GEOM, true>(p, b, 0, q, X[0], X[1], X[2]);
    st = c2p_unwrap<MHD>(p, g2, z, root, cs, pr, tt, sig);
  } else {
    st = c2p_unwrap<MHD>(p, g, z, root, cs, pr, tt, sig);
  }
#pragma unroll
  for (int v = 0; v < L::nprim; ++v)
    if (!MHD || v < 5 || v > 7) prim[v][c] = pr[v];
  if (temp) temp[c] = tt;
  if (z.res.used_density_floor || z.res.used_energy_floor || z.res.used_energy_max ||
      z.res.used_gamma_max) {
#pragma unroll
    for (int v = 0; v < 5; ++v) cons[v][c] = cs[v];
  }
  return st;
}

// ---- wavefront / block minimum (64-lane wavefronts) ----------------------------------
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = smin(v, __shfl_xor(v, off, 64));
  return v;
}

__device__ __forceinline__ double block_min(double v) {
  __shared__ double part[16];
  v = wave_min(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) part[w] = v;
  __syncthreads();
  const int nw = (blockDim.x + 63) >> 6;
  v = threadIdx.x < nw ? part[threadIdx.x] : DBL_MAX;
  if (w == 0) v = wave_min(v);
  return v;
}

}  // namespace phx
