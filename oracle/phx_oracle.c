/*
 * phx_oracle.c -- CPU restatement of phoebus's per-stage fluid path.  TEST
 * INFRASTRUCTURE ONLY (parity checker and timed CPU baseline); never linked into the
 * product.  Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).
 *
 * Operation order follows the reference line by line where it affects rounding; the
 * reference's std::min/std::max are restated as (b < a ? b : a) / (a < b ? b : a).
 */
#include "phx_oracle.h"

#define ORC_PI 3.14159265358979323846 /* M_PI */

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- robust.hpp:10-40 ---------------------------------------------------------------- */
#define ORC_SMALL (10.0 * DBL_MIN)
#define ORC_EPS (10.0 * DBL_EPSILON)
static inline double smin(double a, double b) { return (b < a) ? b : a; } /* std::min */
static inline double smax(double a, double b) { return (a < b) ? b : a; } /* std::max */
double orc_ratio(double a, double b) {
  const double sgn = b >= 0 ? 1.0 : -1.0;
  return a / (b + sgn * ORC_SMALL);
}
static inline double make_positive(double v) { return smax(v, ORC_EPS); }
static inline double make_bounded(double v, double lo, double hi) {
  return smin(smax(v, lo + ORC_EPS), hi * (1.0 - ORC_EPS));
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ---- reconstruction.hpp:101-244 ------------------------------------------------------ */
static inline double mc_lim(double dm, double dp, double alpha) { /* :103-107 */
  const double dc = (dm * dp > 0.0) * 0.5 * (dm + dp);
  return copysign(smin(fabs(dc), alpha * smin(fabs(dm), fabs(dp))), dc);
}

void orc_plm(double qm, double q0, double qp, double *ql, double *qr) { /* :116-123 */
  double dq = qp - q0;
  dq = mc_lim(q0 - qm, dq, 1.5);
  *ql = q0 + 0.5 * dq;
  *qr = q0 - 0.5 * dq;
}

void orc_weno5z(double q0, double q1, double q2, double q3, double q4, double *ql,
                double *qr) { /* :125-183 */
  const double a00 = 1.0 / 3.0, a01 = -7.0 / 6.0, a02 = 11.0 / 6.0;
  const double a10 = -1.0 / 6.0, a11 = 5.0 / 6.0, a12 = 1.0 / 3.0;
  const double a20 = 1.0 / 3.0, a21 = 5.0 / 6.0, a22 = -1.0 / 6.0;
  const double g0 = 0.1, g1 = 0.6, g2 = 0.3;
  const double eps = 1e-100;
  const double tt = 13.0 / 3.0;
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
  beta0 = (beta0 + tau5) / beta0;
  beta1 = (beta1 + tau5) / beta1;
  beta2 = (beta2 + tau5) / beta2;

  double w0 = g0 * beta0 + eps;
  double w1 = g1 * beta1 + eps;
  double w2 = g2 * beta2 + eps;
  double wsum = 1.0 / (w0 + w1 + w2);
  double l = w0 * (a00 * q0 + a01 * q1 + a02 * q2);
  l += w1 * (a10 * q1 + a11 * q2 + a12 * q3);
  l += w2 * (a20 * q2 + a21 * q3 + a22 * q4);
  l *= wsum;
  const double alpha_l = 3.0 * wsum * w0 * w1 * w2 / (g2 * w0 * w1 + g1 * w0 * w2 + g0 * w1 * w2) + eps;

  w0 = g0 * beta2 + eps;
  w1 = g1 * beta1 + eps;
  w2 = g2 * beta0 + eps;
  wsum = 1.0 / (w0 + w1 + w2);
  double r = w0 * (a00 * q4 + a01 * q3 + a02 * q2);
  r += w1 * (a10 * q3 + a11 * q2 + a12 * q1);
  r += w2 * (a20 * q2 + a21 * q1 + a22 * q0);
  r *= wsum;
  const double alpha_r = 3.0 * wsum * w0 * w1 * w2 / (g2 * w0 * w1 + g1 * w0 * w2 + g0 * w1 * w2) + eps;

  double dq = q3 - q2;
  dq = mc_lim(q2 - q1, dq, 2.0);
  const double alpha_lin = 2.0 * alpha_l * alpha_r / (alpha_l + alpha_r);
  *ql = alpha_lin * l + (1.0 - alpha_lin) * (q2 + 0.5 * dq);
  *qr = alpha_lin * r + (1.0 - alpha_lin) * (q2 - 0.5 * dq);
}

static inline double minmod2(double a, double b) { /* MINMOD macro :186 */
  return a * b > 0.0 ? (fabs(a) < fabs(b) ? a : b) : 0.0;
}
static inline double median3(double a, double b, double c) { return a + minmod2(b - a, c - a); }

static double mp5_sub(double Fjm2, double Fjm1, double Fj, double Fjp1, double Fjp2) {
  /* reconstruction.hpp:189-235 */
  const double alpha = 4.0, epsm = 1.e-12;
  double f = 2.0 * Fjm2 - 13.0 * Fjm1 + 47.0 * Fj + 27.0 * Fjp1 - 3.0 * Fjp2;
  f /= 60.0;
  const double fMP = Fj + minmod2(Fjp1 - Fj, alpha * (Fj - Fjm1));
  if ((f - Fj) * (f - fMP) <= epsm) return f;
  const double d2m = Fjm2 + Fj - 2.0 * Fjm1;
  const double d2 = Fjm1 + Fjp1 - 2.0 * Fj;
  const double d2p = Fj + Fjp2 - 2.0 * Fjp1;
  double s1 = minmod2(4.0 * d2 - d2p, 4.0 * d2p - d2);
  double s2 = minmod2(d2, d2p);
  const double dMMp = minmod2(s1, s2);
  s1 = minmod2(4.0 * d2m - d2, 4.0 * d2 - d2m);
  s2 = minmod2(d2, d2m);
  const double dMMm = minmod2(s1, s2);
  const double fUL = Fj + alpha * (Fj - Fjm1);
  const double fAV = 0.5 * (Fj + Fjp1);
  const double fMD = fAV - 0.5 * dMMp;
  const double fLC = 0.5 * (3.0 * Fj - Fjm1) + 4.0 / 3.0 * dMMm;
  s1 = fmin(Fj, Fjp1);
  s1 = fmin(s1, fMD);
  s2 = fmin(Fj, fUL);
  s2 = fmin(s2, fLC);
  const double Fmin = fmax(s1, s2);
  s1 = fmax(Fj, Fjp1);
  s1 = fmax(s1, fMD);
  s2 = fmax(Fj, fUL);
  s2 = fmax(s2, fLC);
  const double Fmax = fmin(s1, s2);
  return median3(f, Fmin, Fmax);
}

void orc_mp5(double q0, double q1, double q2, double q3, double q4, double *ql, double *qr) {
  *ql = mp5_sub(q0, q1, q2, q3, q4); /* :240-244 */
  *qr = mp5_sub(q4, q3, q2, q1, q0);
}

void orc_recon(int type, const double *q, double *ql, double *qr) {
  switch (type) {
  case ORC_RECON_CONSTANT: *ql = q[2]; *qr = q[2]; break;
  case ORC_RECON_LINEAR: orc_plm(q[1], q[2], q[3], ql, qr); break;
  case ORC_RECON_WENO5: orc_weno5z(q[0], q[1], q[2], q[3], q[4], ql, qr); break;
  default: orc_mp5(q[0], q[1], q[2], q[3], q[4], ql, qr); break;
  }
}

/* ---- geometry ------------------------------------------------------------------------ */
static void zero4(double g[4][4]) { memset(g, 0, 16 * sizeof(double)); }

/* spherical Kerr-Schild at (r = |X1|, theta = X2): spherical_kerr_schild.hpp:42-243 */
static void ks_point(double a, double X1, double X2, orc_point_geom *o) {
  const double a2 = a * a;
  const double r = fabs(X1), th = X2, cth = cos(th), sth = sin(th);
  const double sth2 = sth * sth, rho2 = r * r + a2 * cth * cth;
  o->alpha = sqrt(orc_ratio(rho2, 2 * r + rho2));
  o->beta[0] = orc_ratio(2 * r, 2 * r + rho2);
  double(*G)[4] = o->gcov;
  G[0][1] = G[1][0] = orc_ratio(2 * r, rho2);
  G[0][0] = -1 + G[0][1];
  G[0][3] = G[3][0] = -orc_ratio(2 * a * r * sth2, rho2);
  G[1][1] = 1 + G[0][1];
  G[1][3] = G[3][1] = -a * (1 + G[0][1]) * sth2;
  G[2][2] = rho2;
  G[3][3] = (rho2 + a2 * (1 + G[0][1]) * sth2) * sth2;
  const double br = orc_ratio(2 * r, rho2);
  o->gam[0][0] = 1 + br;
  o->gam[0][2] = o->gam[2][0] = -a * (1 + br) * sth2;
  o->gam[1][1] = rho2;
  o->gam[2][2] = (rho2 + a2 * (1 + br) * sth2) * sth2;
  o->gcon[0][0] = orc_ratio(rho2, 2 * r + rho2) + orc_ratio(a2 * sth2, rho2);
  o->gcon[0][2] = o->gcon[2][0] = orc_ratio(a, rho2);
  o->gcon[1][1] = orc_ratio(1, rho2);
  o->gcon[2][2] = orc_ratio(1, rho2 * sth2);
  o->detgamma = sqrt(fabs(rho2 * (2 * r + rho2) * sth2));
  o->detg = sqrt(fabs(rho2 * rho2 * sth2));
}

static void ks_gcon4(double a, double X1, double X2, double gi[4][4]) {
  const double a2 = a * a;
  const double r = fabs(X1), cth = cos(X2), sth = sin(X2);
  const double sth2 = sth * sth, rho2 = r * r + a2 * cth * cth;
  gi[0][1] = gi[1][0] = orc_ratio(2 * r, rho2); /* spherical_kerr_schild.hpp:86-102 */
  gi[0][0] = -1 - gi[0][1];
  gi[1][1] = orc_ratio(-2 * r + rho2 + a2 * sth2, rho2);
  gi[1][3] = gi[3][1] = orc_ratio(a, rho2);
  gi[2][2] = orc_ratio(1, rho2);
  gi[3][3] = orc_ratio(1, rho2 * sth2);
}

/* mckinney_gammie_ryan.hpp:66-108 (operator()) and :143-197 (r_, th_, thG_, thJ_) */
static double mgr_norm(const orc_mgr *m) {
  return 0.5 * ORC_PI * 1. / (1. + 1. / (m->alpha + 1.) * 1. / pow(m->xt, m->alpha));
}
static double mgr_thG(const orc_mgr *m, double X2) {
  return ORC_PI * X2 + ((1. - m->h) / 2.) * sin(2. * ORC_PI * X2);
}
static void mgr_thJ(const orc_mgr *m, double X2, double *y, double *thJ) {
  *y = 2. * X2 - 1.;
  *thJ = mgr_norm(m) * *y * (1. + pow(*y / m->xt, m->alpha) / (m->alpha + 1.)) + 0.5 * ORC_PI;
}
static double mgr_r(const orc_mgr *m, double X1) {
  const double dr = X1 - m->hexp_br;
  return exp(X1 + (dr > 0.) * m->hexp_csq * pow(dr, m->hexp_nsq));
}
static double mgr_th(const orc_mgr *m, double X1, double X2, double *y) {
  const double thG = mgr_thG(m, X2);
  double th;
  if (m->derefine_poles) {
    double thJ;
    mgr_thJ(m, X2, y, &thJ);
    th = thG + exp(m->smooth * (m->x0 - X1)) * (thJ - thG);
  } else {
    *y = 0;
    th = thG;
  }
  const double SINGSMALL = 1.e-20; /* keep theta off the poles */
  if (fabs(th) < SINGSMALL) th = th >= 0. ? SINGSMALL : -SINGSMALL;
  if (fabs(ORC_PI - th) < SINGSMALL) th = th >= ORC_PI ? ORC_PI + SINGSMALL : ORC_PI - SINGSMALL;
  return th;
}

void orc_mgr_transform(const orc_mgr *m, double X1, double X2, double X3, double C[3],
                       double Jcov[3][3], double Jcon[3][3]) {
  double y;
  C[0] = mgr_r(m, X1);
  C[1] = mgr_th(m, X1, X2, &y);
  C[2] = X3;
  const double dr = X1 - m->hexp_br;
  const double drdX1 = exp(X1 + (dr > 0.) * m->hexp_csq * pow(dr, m->hexp_nsq)) *
                       (1. + (dr > 0.) * m->hexp_csq * m->hexp_nsq * pow(dr, m->hexp_nsq - 1.));
  const double dthGdX2 = ORC_PI + ORC_PI * (1 - m->h) * cos(2 * ORC_PI * X2);
  double dthdX1, dthdX2;
  if (m->derefine_poles) {
    double thJ;
    mgr_thJ(m, X2, &y, &thJ);
    const double thG = mgr_thG(m, X2);
    const double dthJdX2 = mgr_norm(m) * (1 + pow(y / m->xt, m->alpha)) * 2.;
    const double e = exp(m->smooth * (m->x0 - X1));
    dthdX1 = -m->smooth * (thJ - thG) * e;
    dthdX2 = dthGdX2 + e * (dthJdX2 - dthGdX2);
  } else {
    dthdX1 = 0.0;
    dthdX2 = dthGdX2;
  }
  memset(Jcov, 0, 9 * sizeof(double));
  memset(Jcon, 0, 9 * sizeof(double));
  Jcov[0][0] = drdX1;
  Jcov[1][0] = dthdX1;
  Jcov[1][1] = dthdX2;
  Jcov[2][2] = 1.;
  Jcon[0][0] = orc_ratio(1., drdX1);
  Jcon[1][0] = -orc_ratio(dthdX1, drdX1 * dthdX2);
  Jcon[1][1] = orc_ratio(1., dthdX2);
  Jcon[2][2] = 1.;
}

static double det3(const double A[3][3]) { /* linear_algebra.hpp:50-54 */
  return (A[0][0] * A[1][1] * A[2][2] + A[0][1] * A[1][2] * A[2][0] +
          A[0][2] * A[1][0] * A[2][1] - A[0][2] * A[1][1] * A[2][0] -
          A[0][1] * A[1][0] * A[2][2] - A[0][0] * A[1][2] * A[2][1]);
}
/* S2ST_ of modified_system.hpp:239-245: the spatial Jacobian as a 4x4 with J[0][0] = 1 */
static double s2st(const double A[3][3], int mu, int nu) {
  if (mu == 0 || nu == 0) return mu == nu ? 1 : 0;
  return A[mu - 1][nu - 1];
}

void orc_geom_point(const orc_geom *g, double X1, double X2, double X3, orc_point_geom *o) {
  memset(o, 0, sizeof(*o));
  if (g->kind == ORC_GEOM_MINKOWSKI) { /* minkowski.hpp:35-115 */
    o->alpha = 1.0;
    o->gcov[0][0] = -1.0;
    for (int d = 0; d < 3; ++d) {
      o->gcov[d + 1][d + 1] = 1.0;
      o->gcon[d][d] = 1.0;
      o->gam[d][d] = 1.0;
    }
    o->detgamma = 1.0;
    o->detg = 1.0;
    return;
  }
  if (g->kind == ORC_GEOM_SPHKS) {
    ks_point(g->a, X1, X2, o);
    return;
  }
  /* FMKS: modified_system.hpp:63-168 over the KS point at (r, theta) */
  double C[3], Jcov[3][3], Jcon[3][3];
  orc_mgr_transform(&g->mgr, X1, X2, X3, C, Jcov, Jcon);
  orc_point_geom k;
  memset(&k, 0, sizeof(k));
  ks_point(g->a, C[0], C[1], &k);
  o->alpha = k.alpha;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) o->beta[i] += Jcon[i][j] * k.beta[j];
  for (int mu = 0; mu < 4; ++mu)
    for (int nu = 0; nu < 4; ++nu)
      for (int lam = 0; lam < 4; ++lam)
        for (int kap = 0; kap < 4; ++kap)
          o->gcov[mu][nu] += k.gcov[lam][kap] * s2st(Jcov, lam, mu) * s2st(Jcov, kap, nu);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int kk = 0; kk < 3; ++kk)
        for (int l = 0; l < 3; ++l) {
          o->gam[i][j] += k.gam[kk][l] * Jcov[kk][i] * Jcov[l][j];
          o->gcon[i][j] += k.gcon[kk][l] * Jcon[i][kk] * Jcon[j][l];
        }
  const double detJ = det3((const double(*)[3])Jcov);
  o->detgamma = fabs(k.detgamma * detJ);
  o->detg = fabs(k.detg * detJ);
}

void orc_geom_gcon4(const orc_geom *g, double X1, double X2, double X3, double gi[4][4]) {
  zero4(gi);
  if (g->kind == ORC_GEOM_MINKOWSKI) {
    gi[0][0] = -1;
    gi[1][1] = gi[2][2] = gi[3][3] = 1;
    return;
  }
  if (g->kind == ORC_GEOM_SPHKS) {
    ks_gcon4(g->a, X1, X2, gi);
    return;
  }
  double C[3], Jcov[3][3], Jcon[3][3], g0[4][4];
  orc_mgr_transform(&g->mgr, X1, X2, X3, C, Jcov, Jcon);
  zero4(g0);
  ks_gcon4(g->a, C[0], C[1], g0);
  for (int mu = 0; mu < 4; ++mu)
    for (int nu = 0; nu < 4; ++nu)
      for (int lam = 0; lam < 4; ++lam)
        for (int kap = 0; kap < 4; ++kap)
          gi[mu][nu] += g0[lam][kap] * s2st(Jcon, mu, lam) * s2st(Jcon, nu, kap);
}

/* geometry_utils.hpp:150-187 (SetMetricGradientByFD): central differences of g_{mu nu}
 * with step max(|X_d|, 1) * eps^(1/3) */
static void fd_dg(const orc_geom *g, double X1, double X2, double X3, double dg[4][4][4]) {
  const double EPS = pow(DBL_EPSILON, 1.0 / 3.0);
  for (int d = 1; d < 4; ++d) {
    double L[3] = {X1, X2, X3}, R[3] = {X1, X2, X3};
    const double dx = smax(fabs(L[d - 1]), 1.0) * EPS;
    L[d - 1] -= dx;
    R[d - 1] += dx;
    orc_point_geom gl, gr;
    orc_geom_point(g, R[0], R[1], R[2], &gr);
    orc_geom_point(g, L[0], L[1], L[2], &gl);
    for (int mu = 0; mu < 4; ++mu)
      for (int nu = 0; nu < 4; ++nu)
        dg[mu][nu][d] = orc_ratio(gr.gcov[mu][nu] - gl.gcov[mu][nu], 2 * dx);
  }
}

void orc_geom_dg(const orc_geom *g, double X1, double X2, double X3, double dg[4][4][4]) {
  memset(dg, 0, 64 * sizeof(double));
  if (g->kind == ORC_GEOM_MINKOWSKI) return;
  if (g->kind == ORC_GEOM_FMKS) {
    fd_dg(g, X1, X2, X3, dg);
    return;
  }
  /* spherical_kerr_schild.hpp:151-194: dg[a][b][c] = d_c g_ab */
  const double a = g->a, a2 = a * a;
  const double r = fabs(X1), th = X2, cth = cos(th), sth = sin(th);
  const double sth2 = sth * sth, rho2 = r * r + a2 * cth * cth;
  const double r2 = r * r, r3 = r2 * r, r4 = r2 * r2, r5 = r4 * r, a4 = a2 * a2;
  const double rho4 = rho2 * rho2, cth2 = cth * cth;
  const double s2th = sin(2 * th), c2th = cos(2 * th);
  (void)r4;
  dg[0][0][1] = 2 * orc_ratio(rho2 - 2 * r2, rho4);
  dg[0][1][1] = dg[1][0][1] = dg[0][0][1];
  dg[0][3][1] = dg[3][0][1] = orc_ratio(2 * a * (r2 - a2 * cth2) * sth2, rho4);
  dg[1][1][1] = dg[0][0][1];
  dg[1][3][1] = dg[3][1][1] = dg[0][3][1];
  dg[2][2][1] = 2 * r;
  dg[3][3][1] = orc_ratio(2 * sth2 * (r5 + a4 * r * cth2 * cth2 - a2 * r2 * sth2 + cth2 * (2 * a2 * r3 + a4 * sth2)), rho4);
  dg[0][0][2] = orc_ratio(4 * r * (rho2 - r2) * sth, rho4 * cth);
  dg[0][1][2] = dg[1][0][2] = dg[0][0][2];
  dg[0][3][2] = dg[3][0][2] = -orc_ratio(2 * a * r * (a2 + r2) * s2th, rho4);
  dg[1][1][2] = dg[0][0][2];
  dg[1][3][2] = dg[3][1][2] = -orc_ratio(2 * a * cth * (r3 * (2 + r) + a4 * cth2 * cth2 + a2 * r * (2 + r + r * c2th)) * sth, rho4);
  dg[2][2][2] = -a2 * s2th;
  dg[3][3][2] = (a2 + r * (r - 2) + orc_ratio(2 * r * (a2 + r2) * (a2 + r2), rho4)) * s2th;
}

void orc_geom_gradlnalpha(const orc_geom *g, double X1, double X2, double X3, double da[4]) {
  da[0] = da[1] = da[2] = da[3] = 0.0;
  if (g->kind == ORC_GEOM_MINKOWSKI) return;
  if (g->kind == ORC_GEOM_FMKS) { /* geometry_utils.hpp:94-123 (SetGradLnAlphaByFD) */
    const double EPS = pow(DBL_EPSILON, 1.0 / 3.0);
    for (int d = 1; d < 4; ++d) {
      double M[3] = {X1, X2, X3}, P[3] = {X1, X2, X3};
      const double dx = smax(fabs(M[d - 1]), 1.0) * EPS;
      P[d - 1] += dx;
      M[d - 1] -= dx;
      orc_point_geom gm, gp;
      orc_geom_point(g, M[0], M[1], M[2], &gm);
      orc_geom_point(g, P[0], P[1], P[2], &gp);
      da[d] = orc_ratio(gp.alpha - gm.alpha, dx * (gm.alpha + gp.alpha));
    }
    return;
  }
  const double a = g->a, a2 = a * a; /* spherical_kerr_schild.hpp:196-205 */
  const double r = fabs(X1), cth = cos(X2), sth = sin(X2);
  const double rho2 = r * r + a2 * cth * cth, rho4 = rho2 * rho2;
  da[1] = orc_ratio(r, rho2) - orc_ratio(1 + r, 2 * r + rho2);
  da[2] = -orc_ratio(2 * a2 * r * cth * sth, 2 * r * rho2 + rho4);
}

/* ---- bounds: fixup.hpp:91-134 (Floors::GetFloors), 188-213 (Ceilings) ---------------- */
void orc_get_floors(const orc_bounds *b, double x1, double x2, double x3, double *rflr,
                    double *sflr) {
  switch (b->floor_type) {
  case ORC_FLOOR_CONSTANT:
    *rflr = b->r0;
    *sflr = b->s0;
    break;
  case ORC_FLOOR_EXPX1_RHO_SIE:
    *rflr = b->r0 * exp(b->ralpha * x1);
    *sflr = b->s0 * exp(b->salpha * x1);
    break;
  case ORC_FLOOR_EXPX1_RHO_U: {
    const double scratch = b->r0 * exp(b->ralpha * x1);
    *sflr = b->s0 * exp(b->salpha * x1) / smax(*rflr, scratch);
    *rflr = scratch;
  } break;
  case ORC_FLOOR_X1_RHO_SIE:
    *rflr = b->r0 * smin(1.0, pow(x1, b->ralpha));
    *sflr = b->s0 * smin(1.0, pow(x1, b->salpha));
    break;
  default: {
    const double r = sqrt(x1 * x1 + x2 * x2 + x3 * x3);
    *rflr = b->r0 * smin(1.0, pow(r, b->ralpha));
    *sflr = b->s0 * smin(1.0, pow(r, b->salpha));
  } break;
  }
  *rflr = smax(*rflr, b->rho_min_eos);
  *sflr = smax(*sflr, b->sie_min_eos);
}

/* ---- ideal-gas EOS (singularity-eos IdealGas closed forms, unit scales 1) ------------ */
static inline double eos_p_rho_sie(const orc_params *p, double rho, double sie) {
  return p->gm1 * rho * sie;
}
static inline double eos_t_rho_sie(const orc_params *p, double rho, double sie) {
  (void)rho;
  return sie / p->cv;
}
static inline double eos_p_rho_t(const orc_params *p, double rho, double t) {
  return p->gm1 * rho * p->cv * t;
}
static inline double eos_bmod_rho_t(const orc_params *p, double rho, double t) {
  return (p->gm1 + 1.0) * p->gm1 * rho * p->cv * t;
}

int orc_nprim(int mhd) { return mhd ? 10 : 7; }
int orc_ncons(int mhd) { return mhd ? 8 : 5; }

/* ---- riemann.hpp:88-209 FluxState::prim_to_flux -------------------------------------- */
static double lorentz(const double v[3], const double gcov[4][4]) { /* relativity_utils:54-59 */
  double vsq = 0.;
  for (int ii = 0; ii < 3; ++ii)
    for (int jj = 0; jj < 3; ++jj) vsq += gcov[ii + 1][jj + 1] * v[ii] * v[jj];
  return sqrt(1. + vsq);
}

void orc_prim_to_flux(const orc_params *p, int dir, const orc_point_geom *g, const double X[3],
                      const double *q, double sie_max, double gam_max, double *vm,
                      double *vp, double *U, double *F) {
  const int mhd = p->mhd;
  const int prs = mhd ? 8 : 5, gm1i = mhd ? 9 : 6;
  double rho_floor = q[0];
  double sie_floor;
  orc_get_floors(&p->bounds, X[0], X[1], X[2], &rho_floor, &sie_floor);
  const double rho = smax(q[0], rho_floor);
  double vpcon[3] = {q[1], q[2], q[3]};
  double W = lorentz(vpcon, g->gcov);
  if (W > gam_max) {
    const double rescale = sqrt((gam_max * gam_max - 1.) / (W * W - 1.));
    for (int ii = 0; ii < 3; ++ii) vpcon[ii] *= rescale;
  }
  double vcon[3] = {vpcon[0] / W, vpcon[1] / W, vpcon[2] / W};
  const double vel = vcon[dir];
  double Bcon[3] = {0.0, 0.0, 0.0};
  const double u = (q[4] / rho > sie_floor ? (q[4] / rho > sie_max ? rho * sie_max : q[4])
                                            : rho * sie_floor);
  const double P = smax(q[prs], 0.0);
  const double gamma1 = q[gm1i];
  if (mhd)
    for (int m = 0; m < 3; ++m) Bcon[m] = q[5 + m];

  double BdotB = 0.0, Bdotv = 0.0, vsq = 0.0;
  for (int m = 0; m < 3; m++) {
    for (int n = 0; n < 3; n++) {
      vsq += g->gcov[m + 1][n + 1] * vcon[m] * vcon[n];
      Bdotv += g->gcov[m + 1][n + 1] * Bcon[m] * vcon[n];
      BdotB += g->gcov[m + 1][n + 1] * Bcon[m] * Bcon[n];
    }
  }
  const double vsq_max = 1.0 - 1.0 / (gam_max * gam_max);
  if (vsq > vsq_max) {
    double scale = vsq_max / vsq;
    vsq *= scale;
    scale = sqrt(scale);
    Bdotv *= scale;
    for (int m = 0; m < 3; m++) vcon[m] *= scale;
  }
  W = 1.0 / sqrt(1 - vsq);
  const double vtil = vel - g->beta[dir] / g->alpha;

  U[0] = rho * W;
  F[0] = U[0] * vtil;

  const double b0 = W * Bdotv;
  const double bsq = (BdotB + b0 * b0) / (W * W);
  const double rhohWsq = (rho + u + P) * W * W;
  const double bsqWsq = bsq * W * W;
  for (int m = 0; m < 3; m++) {
    double bcovm = g->gcov[m + 1][0] * b0 / g->alpha;
    double vcovm = 0.0;
    for (int n = 1; n < 4; n++) {
      bcovm += g->gcov[m + 1][n] * (Bcon[n - 1] / W + b0 * (vcon[n - 1] - g->beta[n - 1] / g->alpha));
      vcovm += g->gcov[m + 1][n] * vcon[n - 1];
    }
    U[1 + m] = (rhohWsq + bsqWsq) * vcovm - b0 * bcovm;
    F[1 + m] = U[1 + m] * vtil + (P + 0.5 * bsq) * (double)(dir == m) - bcovm * Bcon[dir] / W;
  }
  U[4] = rhohWsq + bsqWsq - (P + 0.5 * bsq) - b0 * b0 - U[0]; /* Valencia :172-174 */
  F[4] = U[4] * vtil + (P + 0.5 * bsq) * vel - Bdotv * Bcon[dir];
  if (mhd) {
    for (int m = 5; m <= 7; m++) {
      U[m] = Bcon[m - 5];
      F[m] = U[m] * vtil - Bcon[dir] * (vcon[m - 5] - g->beta[m - 5] / g->alpha);
    }
  }
  const double vasq = bsqWsq / make_positive(rhohWsq + bsqWsq);
  const double cssq = make_bounded(gamma1 * P * W * W / make_positive(rhohWsq), 0.0, 1.0);
  const double cmsq = make_bounded(cssq + vasq * (1.0 - cssq), 0.0, 1.0);
  const double vcoff = g->alpha / (1.0 - vsq * cmsq);
  const double v0 = vel * (1.0 - cmsq);
  const double vpm = sqrt(make_positive(cmsq * (1.0 - vsq) * (g->gcon[dir][dir] * (1.0 - vsq * cmsq) - vel * v0)));
  *vp = vcoff * (v0 + vpm) - g->beta[dir];
  *vm = vcoff * (v0 - vpm) - g->beta[dir];
}

/* riemann.hpp:257-301 */
double orc_face_flux(const orc_params *p, int dir, const double X[3], const double *ql,
                     const double *qr, double *flux) {
  orc_point_geom g;
  orc_geom_point(&p->geom, X[0], X[1], X[2], &g);
  const double gam_max = p->bounds.gam_max, sie_max = p->bounds.sie_max;
  double Ul[8], Ur[8], Fl[8], Fr[8], vml, vpl, vmr, vpr;
  orc_prim_to_flux(p, dir, &g, X, ql, sie_max, gam_max, &vml, &vpl, Ul, Fl);
  orc_prim_to_flux(p, dir, &g, X, qr, sie_max, gam_max, &vmr, &vpr, Ur, Fr);
  const int nc = orc_ncons(p->mhd);
  if (p->riemann == ORC_RIEMANN_LLF) {
    const double cmax = smax(smax(-vml, vpl), smax(-vmr, vpr));
    for (int m = 0; m < nc; m++)
      flux[m] = 0.5 * ((Fl[m] + Fr[m]) * g.detg - cmax * ((Ur[m] - Ul[m]) * g.detgamma));
    return cmax;
  }
  const double cl = smin(smin(vml, vmr), 0.0);
  const double cr = smax(smax(vpl, vpr), 0.0);
  const double crcl = cr * cl;
  const double inv = orc_ratio(1.0, cr - cl);
  for (int m = 0; m < nc; m++)
    flux[m] = ((cr * Fl[m] - cl * Fr[m]) * g.detg + crcl * (Ur[m] - Ul[m]) * g.detgamma) * inv;
  return smax(-cl, cr);
}

/* ---- prim2con.hpp:25-102 ------------------------------------------------------------- */
void orc_signal_speeds(double rho, double u, double pr, double bsq, const double v[3],
                       double vsq, double gam1, double alpha, const double beta[3],
                       const double gcon[3][3], double sig[3]) {
  const double rho_rel = rho + u + pr;
  const double vasq = bsq / (rho_rel + bsq);
  double cssq = gam1 * pr / rho_rel;
  cssq += vasq - cssq * vasq;
  const double vcoff = alpha / (1.0 - vsq * cssq);
  for (int m = 0; m < 3; ++m) {
    const double vpm = sqrt(cssq * (1.0 - vsq) * (gcon[m][m] * (1.0 - vsq * cssq) - v[m] * v[m] * (1.0 - cssq)));
    const double vp = vcoff * (v[m] * (1.0 - cssq) + vpm) - beta[m];
    const double vm = vcoff * (v[m] * (1.0 - cssq) - vpm) - beta[m];
    sig[m] = smax(fabs(vp), fabs(vm));
  }
}

static void p2c_core(const double rho, const double vp[3], const double b[3], const double u,
                     const double pr, const double gam1, const double gcov[4][4],
                     const double gcon[3][3], const double beta[3], const double alpha,
                     const double gdet, double *D, double S[3], double bcons[3], double *tau,
                     double sig[3]) {
  double vsq = 0.0, Bsq = 0.0, Bdotv = 0.0;
  const double W = lorentz(vp, gcov);
  const double iW = 1.0 / W;
  double v[3] = {vp[0] * iW, vp[1] * iW, vp[2] * iW};
  for (int ii = 0; ii < 3; ++ii)
    for (int jj = 0; jj < 3; ++jj) {
      vsq += gcov[ii + 1][jj + 1] * v[ii] * v[jj];
      Bsq += gcov[ii + 1][jj + 1] * b[ii] * b[jj];
      Bdotv += gcov[ii + 1][jj + 1] * b[ii] * v[jj];
    }
  const double b0 = W * Bdotv;
  double bcon[4] = {W * Bdotv / alpha, 0.0, 0.0, 0.0};
  for (int m = 0; m < 3; ++m) bcon[m + 1] = b[m] * iW + Bdotv * W * (v[m] - beta[m] / alpha);
  const double bsq = Bsq * iW * iW + Bdotv * Bdotv;
  double bcov[3] = {0.0, 0.0, 0.0};
  for (int m = 0; m < 3; ++m)
    for (int n = 0; n < 4; ++n) bcov[m] += gcov[m + 1][n] * bcon[n];
  *D = gdet * rho * W;
  const double rho_rel = (rho + u + pr + bsq) * W * W;
  for (int m = 0; m < 3; ++m) {
    double vcov = 0.0;
    for (int n = 0; n < 3; ++n) vcov += gcov[m + 1][n + 1] * v[n];
    S[m] = gdet * (rho_rel * vcov - b0 * bcov[m]);
    bcons[m] = gdet * b[m];
  }
  *tau = gdet * (rho_rel - (pr + 0.5 * bsq) - b0 * b0) - *D;
  if (sig) orc_signal_speeds(rho, u, pr, bsq, v, vsq, gam1, alpha, beta, gcon, sig);
}

void orc_p2c(const orc_params *p, const orc_point_geom *g, const double *prim, double *cons,
             double sig[3]) {
  const int mhd = p->mhd, prs = mhd ? 8 : 5, gm1i = mhd ? 9 : 6;
  const double vel[3] = {prim[1], prim[2], prim[3]};
  double bp[3] = {0.0, 0.0, 0.0};
  if (mhd)
    for (int d = 0; d < 3; ++d) bp[d] = prim[5 + d];
  double S[3], bc[3];
  p2c_core(prim[0], vel, bp, prim[4], prim[prs], prim[gm1i], g->gcov, g->gcon, g->beta,
           g->alpha, g->detgamma, &cons[0], S, bc, &cons[4], sig);
  for (int d = 0; d < 3; ++d) cons[1 + d] = S[d];
  if (mhd)
    for (int d = 0; d < 3; ++d) cons[5 + d] = bc[d];
}

/* ---- root_find.hpp:41-58, 109-167 ----------------------------------------------------- */
double orc_regula_falsi(orc_fn f, void *ctx, double a, double b, double tol, double guess,
                        int max_iter, int *iters, int *bracketed) {
  double ya, yb;
  int it = 0;
  if (a <= guess && b >= guess) { /* refine_bracket */
    ya = f(guess, ctx);
    yb = f(b, ctx);
    if (ya * yb > 0.0) {
      yb = ya;
      b = guess;
      ya = f(a, ctx);
    } else {
      a = guess;
    }
  } else {
    ya = f(a, ctx);
    yb = f(b, ctx);
  }
  if (ya * yb > 0.0) { /* check_bracket: the reference aborts here (PARTHENON_FAIL) */
    *bracketed = 0;
    *iters = 0;
    return guess;
  }
  *bracketed = 1;
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
    const double yc = sign * f(c, ctx);
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
  *iters = it;
  return 0.5 * (a + b);
}

/* ---- con2prim_robust.hpp:46-180 Residual ---------------------------------------------- */
typedef struct {
  const orc_params *p;
  double D, q, bsq, bsq_rpsq, rsq, rbsq, v0sq;
  double x1, x2, x3, fsf;
  double rho_floor, e_floor, gam_max, e_max;
  int used_density_floor, used_energy_floor, used_energy_max, used_gamma_max;
} orc_residual;

static void res_init(orc_residual *r, const orc_params *p, double D, double q, double bsq,
                     double bsq_rpsq, double rsq, double rbsq, double v0sq, double x1,
                     double x2, double x3) {
  r->p = p;
  r->D = D; r->q = q; r->bsq = bsq; r->bsq_rpsq = bsq_rpsq; r->rsq = rsq; r->rbsq = rbsq;
  r->v0sq = v0sq; r->x1 = x1; r->x2 = x2; r->x3 = x3;
  r->fsf = p->c2p_floor_scale_fac;
  /* :56-61.  rho_floor_ enters GetFloors uninitialised in the reference; its input value
   * only matters for the ExpX1RhoU floor's energy part, which is discarded here. */
  double garbage = 0.0;
  r->rho_floor = 0.0;
  orc_get_floors(&p->bounds, x1, x2, x3, &r->rho_floor, &garbage);
  r->gam_max = p->bounds.gam_max;
  r->e_max = p->bounds.sie_max;
  r->rho_floor *= r->fsf;
  r->e_floor = 0.0;
  r->used_density_floor = r->used_energy_floor = r->used_energy_max = r->used_gamma_max = 0;
}
static inline double res_x(const orc_residual *r, double mu) { return orc_ratio(1.0, 1.0 + mu * r->bsq); }
static inline double res_rbarsq(const orc_residual *r, double mu, double x) {
  return x * (x * r->rsq + mu * (1.0 + x) * r->rbsq);
}
static inline double res_qbar(const orc_residual *r, double mu, double x) {
  const double mux = mu * x;
  return r->q - 0.5 * (r->bsq + mux * mux * r->bsq_rpsq);
}
static inline double res_vhatsq(orc_residual *r, double mu, double rbarsq) {
  const double vsq_trial = mu * mu * rbarsq;
  if (vsq_trial > r->v0sq) {
    r->used_gamma_max = 1;
    return r->v0sq;
  }
  r->used_gamma_max = 0;
  return vsq_trial;
}
static inline double res_rhohat(orc_residual *r, double iWhat) {
  const double rho_trial = r->D * iWhat;
  if (rho_trial <= r->rho_floor) {
    r->used_density_floor = 1;
    return r->rho_floor;
  }
  r->used_density_floor = 0;
  return rho_trial;
}
static inline double res_ehat(orc_residual *r, double mu, double qbar, double rbarsq,
                              double vhatsq, double What) {
  double rho = res_rhohat(r, orc_ratio(1.0, What));
  orc_get_floors(&r->p->bounds, r->x1, r->x2, r->x3, &rho, &r->e_floor);
  r->e_floor *= r->fsf;
  const double ehat_trial = What * (qbar - mu * rbarsq) + vhatsq * What * orc_ratio(What, 1.0 + What);
  r->used_energy_floor = 0;
  r->used_energy_max = 0;
  if (ehat_trial <= r->e_floor) {
    r->used_energy_floor = 1;
    return r->e_floor;
  } else if (ehat_trial > r->e_max) {
    r->used_energy_max = 1;
    return r->e_max;
  }
  return ehat_trial;
}
static double res_eval(double mu, void *ctx) { /* :121-141 */
  orc_residual *r = (orc_residual *)ctx;
  const double x = res_x(r, mu);
  const double rbarsq = res_rbarsq(r, mu, x);
  const double qbar = res_qbar(r, mu, x);
  const double vhatsq = res_vhatsq(r, mu, rbarsq);
  const double iWhat = sqrt(1.0 - vhatsq);
  const double What = orc_ratio(1.0, iWhat);
  const double rhohat = res_rhohat(r, iWhat);
  const double ehat = res_ehat(r, mu, qbar, rbarsq, vhatsq, What);
  const double Phat = eos_p_rho_sie(r->p, rhohat, ehat);
  double hhat = rhohat * (1.0 + ehat) + Phat;
  const double ahat = Phat / make_positive(hhat - Phat);
  hhat /= make_positive(rhohat);
  (void)hhat;
  const double nua = (1.0 + ahat) * (1.0 + ehat) * iWhat;
  const double nub = (1.0 + ahat) * (1.0 + qbar - mu * rbarsq);
  const double nuhat = smax(nua, nub);
  const double muhat = 1.0 / (nuhat + mu * rbarsq);
  return mu - muhat;
}

/* set-up shared by the solver and the residual probe (con2prim_robust.hpp:305-367) */
static void c2p_setup(const orc_params *p, const orc_point_geom *g, const double X[3],
                      const double *cons, orc_residual *res, double rcon[3], double rcov[3],
                      double bu[3], double *bdotr_out, double primB[3]) {
  const double igdet = 1.0 / g->detgamma;
  double rhoflr = 0.0, epsflr;
  orc_get_floors(&p->bounds, X[0], X[1], X[2], &rhoflr, &epsflr);
  rhoflr *= p->c2p_floor_scale_fac;
  const double gam_max = p->bounds.gam_max;
  const double D = cons[0] * igdet;
  const double tau = cons[4] * igdet;
  const double q = orc_ratio(tau, D);
  double bsq = 0.0, bsq_rpsq = 0.0, rbsq = 0.0, rsq = 0.0, bdotr = 0.0;
  for (int i = 0; i < 3; ++i) rcov[i] = orc_ratio(cons[1 + i], cons[0]);
  for (int i = 0; i < 3; ++i) {
    rcon[i] = 0.0;
    for (int j = 0; j < 3; ++j) rcon[i] += g->gcon[i][j] * rcov[j];
    rsq += rcon[i] * rcov[i];
  }
  bu[0] = bu[1] = bu[2] = 0.0;
  primB[0] = primB[1] = primB[2] = 0.0;
  if (p->mhd) {
    for (int i = 0; i < 3; ++i) primB[i] = cons[5 + i] * igdet;
    const double sD = 1.0 / sqrt(smax(D, rhoflr));
    for (int i = 0; i < 3; ++i) {
      bu[i] = primB[i] * sD;
      bdotr += bu[i] * rcov[i];
    }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) bsq += g->gcov[i + 1][j + 1] * bu[i] * bu[j];
    bsq = smax(0.0, bsq);
    rbsq = bdotr * bdotr;
    bsq_rpsq = bsq * rsq - rbsq;
  }
  const double zsq = rsq / 1.0; /* h0sq_ = 1 (:251) */
  const double v0sq = smin(zsq / (1.0 + zsq), 1.0 - 1.0 / (gam_max * gam_max));
  res_init(res, p, D, q, bsq, bsq_rpsq, rsq, rbsq, v0sq, X[0], X[1], X[2]);
  *bdotr_out = bdotr;
}

double orc_c2p_residual(const orc_params *p, const orc_point_geom *g, const double X[3],
                        const double *cons, double mu) {
  orc_residual res;
  double rcon[3], rcov[3], bu[3], bdotr, primB[3];
  c2p_setup(p, g, X, cons, &res, rcon, rcov, bu, &bdotr, primB);
  return res_eval(mu, &res);
}

double orc_c2p(const orc_params *p, const orc_point_geom *g, const double X[3], double *cons,
               double *prim, double *temp, double *mu_io, double sig[3], int *iters) {
  const int mhd = p->mhd, prs = mhd ? 8 : 5, gm1i = mhd ? 9 : 6;
  *iters = 0;
  if (isnan(cons[0]) + isnan(cons[1]) + isnan(cons[4]) > 0) return 0.0; /* :303-304 */
  orc_residual res;
  double rcon[3], rcov[3], bu[3], bdotr, primB[3];
  c2p_setup(p, g, X, cons, &res, rcon, rcov, bu, &bdotr, primB);
  if (mhd)
    for (int i = 0; i < 3; ++i) prim[5 + i] = primB[i];

  int bracketed = 1;
  const double mu = orc_regula_falsi(res_eval, &res, 0.0, 1.0, p->c2p_tol, *mu_io,
                                     p->c2p_max_iter, iters, &bracketed);
  if (!bracketed) { /* reference aborts; reported here as a failure with -1 iterations */
    *iters = -1;
    return 0.0;
  }
  *mu_io = mu;

  /* unwrap (:382-403) */
  const double x = res_x(&res, mu);
  const double rbarsq = res_rbarsq(&res, mu, x);
  double vsq = res_vhatsq(&res, mu, rbarsq);
  const double iW = sqrt(1.0 - vsq);
  prim[0] = res_rhohat(&res, iW);
  const double qbar = res_qbar(&res, mu, x);
  const double W = 1.0 / iW;
  prim[4] = res_ehat(&res, mu, qbar, rbarsq, vsq, W);
  *temp = eos_t_rho_sie(p, prim[0], prim[4]);
  prim[4] *= prim[0];
  prim[prs] = eos_p_rho_t(p, prim[0], *temp);
  prim[gm1i] = eos_bmod_rho_t(p, prim[0], *temp) / prim[prs];

  double vel[3];
  for (int i = 0; i < 3; ++i) {
    vel[i] = W * mu * x * (rcon[i] + mu * bdotr * bu[i]);
    prim[1 + i] = vel[i];
  }
  if (mhd)
    for (int i = 0; i < 3; ++i) bu[i] = prim[5 + i];

  const int used_bounds = res.used_density_floor || res.used_energy_floor ||
                          res.used_energy_max || res.used_gamma_max;
  if (used_bounds) { /* :409-419 */
    double S[3], bcons[3];
    p2c_core(prim[0], vel, bu, prim[4], prim[prs], prim[gm1i], g->gcov, g->gcon, g->beta,
             g->alpha, g->detgamma, &cons[0], S, bcons, &cons[4], sig);
    for (int i = 0; i < 3; ++i) cons[1 + i] = S[i];
  } else { /* :420-435 */
    for (int m = 0; m < 3; ++m) vel[m] *= iW;
    double bsq = 0.0;
    if (mhd) {
      double bdotv = 0.0;
      for (int m = 0; m < 3; ++m)
        for (int n = 0; n < 3; ++n) {
          bsq += bu[m] * bu[n] * g->gcov[m + 1][n + 1];
          bdotv += bu[m] * vel[n] * g->gcov[m + 1][n + 1];
        }
      bsq = iW * iW * bsq + bdotv * bdotv;
    }
    orc_signal_speeds(prim[0], prim[4], prim[prs], bsq, vel, vsq, prim[gm1i], g->alpha,
                      g->beta, g->gcon, sig);
  }
  if (isnan(cons[0]) + isnan(cons[1]) + isnan(cons[4]) > 0) return 0.0;
  if (p->fail_on_floors && (res.used_density_floor || res.used_energy_floor)) return 0.0;
  if (p->fail_on_ceilings && res.used_gamma_max) return 0.0;
  return 1.0;
}

/* ---- tmunu.hpp:35-137 and fluid.cpp:566-657 ------------------------------------------ */
static void tmunu_full(const orc_params *p, double X1, double X2, double X3,
                       const orc_point_geom *g, const double *prim, double T[4][4]) {
  const int mhd = p->mhd, prs = mhd ? 8 : 5;
  double u[4], b[4], gi[4][4];
  double Bdotv = 0.0, Bsq = 0.0;
  const double vp[3] = {prim[1], prim[2], prim[3]};
  double bp[3] = {0.0, 0.0, 0.0};
  if (mhd)
    for (int d = 0; d < 3; ++d) bp[d] = prim[5 + d];
  double vsq = 0.;
  for (int ii = 0; ii < 3; ++ii)
    for (int jj = 0; jj < 3; ++jj) vsq += g->gam[ii][jj] * vp[ii] * vp[jj];
  const double W = sqrt(1. + vsq);
  for (int ii = 0; ii < 3; ++ii)
    for (int jj = 0; jj < 3; ++jj) {
      Bdotv += bp[ii] * (vp[jj] / W) * g->gam[ii][jj];
      Bsq += bp[ii] * bp[jj] * g->gam[ii][jj];
    }
  const double iW = orc_ratio(1., W);
  const double alpha = g->alpha;
  u[0] = orc_ratio(W, fabs(alpha));
  b[0] = u[0] * Bdotv;
  for (int l = 1; l < 4; ++l) {
    u[l] = vp[l - 1] - u[0] * g->beta[l - 1];
    b[l] = iW * (bp[l - 1] + alpha * b[0] * u[l]);
  }
  const double bsq = Bsq * iW * iW + Bdotv * Bdotv;
  orc_geom_gcon4(&p->geom, X1, X2, X3, gi);
  const double A = prim[0] + prim[4] + prim[prs] + bsq;
  const double B = prim[prs] + 0.5 * bsq;
  for (int mu = 0; mu < 4; ++mu)
    for (int nu = 0; nu < 4; ++nu) T[mu][nu] = A * u[mu] * u[nu] + B * gi[mu][nu] - b[mu] * b[nu];
}

void orc_tmunu(const orc_params *p, double X1, double X2, double X3, const double *prim,
               double T[4][4]) {
  orc_point_geom g;
  orc_geom_point(&p->geom, X1, X2, X3, &g);
  tmunu_full(p, X1, X2, X3, &g, prim, T);
}

void orc_sources(const orc_params *p, double X1, double X2, double X3, const double *prim,
                 double src[5]) {
  orc_point_geom g;
  orc_geom_point(&p->geom, X1, X2, X3, &g);
  double T[4][4], dg[4][4][4], gam[4][4][4];
  tmunu_full(p, X1, X2, X3, &g, prim, T);
  const double gdet = g.detg;
  orc_geom_dg(&p->geom, X1, X2, X3, dg);
  for (int a = 0; a < 4; ++a) /* geometry_utils.hpp:72-78 */
    for (int b = 0; b < 4; ++b)
      for (int c = 0; c < 4; ++c) gam[a][b][c] = 0.5 * (dg[b][a][c] + dg[c][a][b] - dg[b][c][a]);
  src[0] = 0.0;
  for (int l = 0; l < 3; ++l) {
    double s = 0.0;
    for (int m = 0; m < 4; ++m)
      for (int n = 0; n < 4; ++n) s += T[m][n] * (dg[n][l + 1][m] - gam[l + 1][n][m]);
    src[1 + l] = gdet * s;
  }
  const double alpha = g.alpha;
  const double ia2 = orc_ratio(1, alpha * alpha);
  const double gcon0[4] = {-ia2, ia2 * g.beta[0], ia2 * g.beta[1], ia2 * g.beta[2]};
  double TGam = 0.0;
  for (int m = 0; m < 4; m++)
    for (int n = 0; n < 4; n++) {
      double gam0 = 0;
      for (int r = 0; r < 4; r++) gam0 += gcon0[r] * gam[r][m][n];
      TGam += T[m][n] * gam0;
    }
  double da[4], Ta = 0.0;
  orc_geom_gradlnalpha(&p->geom, X1, X2, X3, da);
  for (int m = 0; m < 4; m++) Ta += T[m][0] * da[m];
  src[4] = gdet * alpha * (Ta - TGam);
}

/* ---- block helpers ------------------------------------------------------------------- */
void orc_block_dims(const orc_block *b, int n[3]) {
  for (int d = 0; d < 3; ++d) n[d] = b->nx[d] + (d < b->ndim ? 2 * b->ng : 0);
}
static inline double xf(const orc_block *b, int d, int idx) {
  const int is = d < b->ndim ? b->ng : 0;
  return b->xmin[d] + (idx - is) * b->dx[d];
}
static inline double xc(const orc_block *b, int d, int idx) {
  const int is = d < b->ndim ? b->ng : 0;
  return b->xmin[d] + (idx - is + 0.5) * b->dx[d];
}

/* fluid.cpp:659-773.  Face (k,j,i) of direction d lies at the lower edge of cell (k,j,i);
 * its left state is the upper-edge reconstruction of the cell below, its right state the
 * lower-edge reconstruction of cell (k,j,i).  Faces span k in [ks-dk, ke+dk],
 * j in [js-dj, je+dj], i in [is-1, ie+1] for every direction, as in the reference. */
void orc_calculate_fluxes(const orc_params *p, const orc_block *b, const double *prim,
                          double *flux, double *face_sig) {
  int n[3];
  orc_block_dims(b, n);
  const long nc = (long)n[0] * n[1] * n[2];
  const int np = orc_nprim(p->mhd), ncn = orc_ncons(p->mhd);
  const int ng = b->ng, ndim = b->ndim;
  const int dj = ndim > 1, dk = ndim > 2;
  const int is = ng, ie = ng + b->nx[0] - 1;
  const int js = dj ? ng : 0, je = js + b->nx[1] - 1;
  const int ks = dk ? ng : 0, ke = ks + b->nx[2] - 1;
  const long strides[3] = {1, n[0], (long)n[0] * n[1]};
  for (int d = 0; d < ndim; ++d) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = ks - dk; k <= ke + dk; ++k)
      for (int j = js - dj; j <= je + dj; ++j)
        for (int i = is - 1; i <= ie + 1; ++i) {
          const long c = ((long)k * n[1] + j) * n[0] + i;
          const long st = strides[d];
          double ql[10], qr[10], f[8];
          for (int v = 0; v < np; ++v) {
            const double *pv = prim + v * nc;
            double s[5], lo, hi;
            /* cell below: stencil centred at c - st */
            for (int o = 0; o < 5; ++o) s[o] = pv[c - st + (o - 2) * st];
            orc_recon(p->recon, s, &lo, &hi);
            ql[v] = lo; /* upper edge of the cell below */
            for (int o = 0; o < 5; ++o) s[o] = pv[c + (o - 2) * st];
            orc_recon(p->recon, s, &lo, &hi);
            qr[v] = hi; /* lower edge of this cell */
          }
          double X[3] = {xc(b, 0, i), xc(b, 1, j), xc(b, 2, k)};
          const int idx[3] = {i, j, k};
          X[d] = xf(b, d, idx[d]);
          const double sg = orc_face_flux(p, d, X, ql, qr, f);
          for (int m = 0; m < ncn; ++m) flux[((long)d * ncn + m) * nc + c] = f[m];
          face_sig[(long)d * nc + c] = sg;
        }
  }
}

/* fluid.cpp:775-841 */
void orc_flux_ct(const orc_params *p, const orc_block *b, double *flux, double *emf) {
  if (!p->mhd || b->ndim == 1) return;
  int n[3];
  orc_block_dims(b, n);
  const long nc = (long)n[0] * n[1] * n[2];
  const int ncn = orc_ncons(p->mhd);
  const int ng = b->ng;
  const int is = ng, ie = ng + b->nx[0] - 1, js = ng, je = ng + b->nx[1] - 1;
  const int ks = b->ndim > 2 ? ng : 0, ke = ks + b->nx[2] - 1;
#define FX(d, comp, k, j, i) flux[((long)(d) * ncn + 5 + (comp)) * nc + (((long)(k) * n[1] + (j)) * n[0] + (i))]
#define EMF(c, k, j, i) emf[(long)(c) * nc + (((long)(k) * n[1] + (j)) * n[0] + (i))]
  if (b->ndim == 2) {
    for (int k = ks; k <= ke; ++k)
      for (int j = js; j <= je + 1; ++j)
        for (int i = is; i <= ie + 1; ++i)
          EMF(0, k, j, i) = 0.25 * (FX(0, 1, k, j, i) + FX(0, 1, k, j - 1, i) - FX(1, 0, k, j, i) - FX(1, 0, k, j, i - 1));
    for (int k = ks; k <= ke; ++k)
      for (int j = js; j <= je + 1; ++j)
        for (int i = is; i <= ie + 1; ++i) {
          FX(0, 0, k, j, i) = 0.0;
          FX(0, 1, k, j, i) = 0.5 * (EMF(0, k, j, i) + EMF(0, k, j + 1, i));
          FX(1, 0, k, j, i) = -0.5 * (EMF(0, k, j, i) + EMF(0, k, j, i + 1));
          FX(1, 1, k, j, i) = 0.0;
        }
    return;
  }
#pragma omp parallel for collapse(2)
  for (int k = ks; k <= ke + 1; ++k)
    for (int j = js; j <= je + 1; ++j)
      for (int i = is; i <= ie + 1; ++i) {
        EMF(0, k, j, i) = 0.25 * (FX(1, 2, k, j, i) + FX(1, 2, k - 1, j, i) - FX(2, 1, k, j, i) - FX(2, 1, k, j - 1, i));
        EMF(1, k, j, i) = -0.25 * (FX(0, 2, k, j, i) + FX(0, 2, k - 1, j, i) - FX(2, 0, k, j, i) - FX(2, 0, k, j, i - 1));
        EMF(2, k, j, i) = 0.25 * (FX(0, 1, k, j, i) + FX(0, 1, k, j - 1, i) - FX(1, 0, k, j, i) - FX(1, 0, k, j, i - 1));
      }
#pragma omp parallel for collapse(2)
  for (int k = ks; k <= ke + 1; ++k)
    for (int j = js; j <= je + 1; ++j)
      for (int i = is; i <= ie + 1; ++i) {
        FX(0, 0, k, j, i) = 0.0;
        FX(0, 1, k, j, i) = 0.5 * (EMF(2, k, j, i) + EMF(2, k, j + 1, i));
        FX(0, 2, k, j, i) = -0.5 * (EMF(1, k, j, i) + EMF(1, k + 1, j, i));
        FX(1, 0, k, j, i) = -0.5 * (EMF(2, k, j, i) + EMF(2, k, j, i + 1));
        FX(1, 1, k, j, i) = 0.0;
        FX(1, 2, k, j, i) = 0.5 * (EMF(0, k, j, i) + EMF(0, k + 1, j, i));
        FX(2, 0, k, j, i) = 0.5 * (EMF(1, k, j, i) + EMF(1, k, j, i + 1));
        FX(2, 1, k, j, i) = -0.5 * (EMF(0, k, j, i) + EMF(0, k, j + 1, i));
        FX(2, 2, k, j, i) = 0.0;
      }
#undef FX
#undef EMF
}

/* parthenon Update::FluxDivergence (uniform coordinates: FaceArea_1 = dx2*dx3, ...,
 * CellVolume = dx1*dx2*dx3), SumData (dudt += src for S_i, tau), AverageIndependentData
 * (U_s = beta*U_s + (1-beta)*U_base), UpdateIndependentData (U_out = U_s + beta*dt*dudt),
 * in the order of phoebus_driver.cpp:560-576. */
void orc_flux_divergence_update(const orc_params *p, const orc_block *b, const double *flux,
                                const double *src, double *u_stage, const double *u_base,
                                double beta, double dt, double *u_out) {
  int n[3];
  orc_block_dims(b, n);
  const long nc = (long)n[0] * n[1] * n[2];
  const int ncn = orc_ncons(p->mhd), ndim = b->ndim, ng = b->ng;
  const int is = ng, ie = ng + b->nx[0] - 1;
  const int js = ndim > 1 ? ng : 0, je = js + b->nx[1] - 1;
  const int ks = ndim > 2 ? ng : 0, ke = ks + b->nx[2] - 1;
  const double a1 = b->dx[1] * b->dx[2], a2 = b->dx[0] * b->dx[2], a3 = b->dx[0] * b->dx[1];
  const double vol = b->dx[0] * b->dx[1] * b->dx[2];
  const double bdt = beta * dt;
#pragma omp parallel for collapse(2)
  for (int k = ks; k <= ke; ++k)
    for (int j = js; j <= je; ++j)
      for (int i = is; i <= ie; ++i) {
        const long c = ((long)k * n[1] + j) * n[0] + i;
        for (int l = 0; l < ncn; ++l) {
          const double *F1 = flux + ((long)0 * ncn + l) * nc;
          double du = a1 * F1[c + 1] - a1 * F1[c];
          if (ndim >= 2) {
            const double *F2 = flux + ((long)1 * ncn + l) * nc;
            du += a2 * F2[c + n[0]] - a2 * F2[c];
          }
          if (ndim == 3) {
            const double *F3 = flux + ((long)2 * ncn + l) * nc;
            du += a3 * F3[c + (long)n[0] * n[1]] - a3 * F3[c];
          }
          double dudt = -du / vol;
          if (src && l >= 1 && l <= 4) dudt = dudt + src[(long)l * nc + c];
          const double us = beta * u_stage[(long)l * nc + c] + (1 - beta) * u_base[(long)l * nc + c];
          u_stage[(long)l * nc + c] = us;
          u_out[(long)l * nc + c] = us + bdt * dudt;
        }
      }
}

void orc_con2prim_block(const orc_params *p, const orc_block *b, const int lo[3],
                        const int hi[3], double *cons, double *prim, double *temp,
                        double *mu, double *cell_sig, double *fail, int *iters) {
  int n[3];
  orc_block_dims(b, n);
  const long nc = (long)n[0] * n[1] * n[2];
  const int np = orc_nprim(p->mhd), ncn = orc_ncons(p->mhd);
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
  for (int k = lo[2]; k <= hi[2]; ++k)
    for (int j = lo[1]; j <= hi[1]; ++j)
      for (int i = lo[0]; i <= hi[0]; ++i) {
        const long c = ((long)k * n[1] + j) * n[0] + i;
        double X[3] = {xc(b, 0, i), xc(b, 1, j), xc(b, 2, k)};
        orc_point_geom g;
        orc_geom_point(&p->geom, X[0], X[1], X[2], &g);
        double cs[8], pr[10], sig[3], t = temp ? temp[c] : 0.0, m = mu[c];
        for (int v = 0; v < ncn; ++v) cs[v] = cons[v * nc + c];
        for (int v = 0; v < np; ++v) pr[v] = prim[v * nc + c];
        for (int d = 0; d < 3; ++d) sig[d] = d < b->ndim ? cell_sig[d * nc + c] : 0.0;
        int it = 0;
        const double st = orc_c2p(p, &g, X, cs, pr, &t, &m, sig, &it);
        for (int v = 0; v < ncn; ++v) cons[v * nc + c] = cs[v];
        for (int v = 0; v < np; ++v) prim[v * nc + c] = pr[v];
        for (int d = 0; d < b->ndim; ++d) cell_sig[d * nc + c] = sig[d];
        if (temp) temp[c] = t;
        mu[c] = m;
        fail[c] = st;
        if (iters) iters[c] = it;
      }
}

void orc_prim_to_con_block(const orc_params *p, const orc_block *b, const int lo[3],
                           const int hi[3], const double *prim, double *cons,
                           double *cell_sig) {
  int n[3];
  orc_block_dims(b, n);
  const long nc = (long)n[0] * n[1] * n[2];
  const int np = orc_nprim(p->mhd), ncn = orc_ncons(p->mhd);
#pragma omp parallel for collapse(2)
  for (int k = lo[2]; k <= hi[2]; ++k)
    for (int j = lo[1]; j <= hi[1]; ++j)
      for (int i = lo[0]; i <= hi[0]; ++i) {
        const long c = ((long)k * n[1] + j) * n[0] + i;
        orc_point_geom g;
        orc_geom_point(&p->geom, xc(b, 0, i), xc(b, 1, j), xc(b, 2, k), &g);
        double pr[10], cs[8], sig[3];
        for (int v = 0; v < np; ++v) pr[v] = prim[v * nc + c];
        orc_p2c(p, &g, pr, cs, sig);
        for (int v = 0; v < ncn; ++v) cons[v * nc + c] = cs[v];
        if (cell_sig)
          for (int d = 0; d < b->ndim; ++d) cell_sig[d * nc + c] = sig[d];
      }
}

void orc_source_terms_block(const orc_params *p, const orc_block *b, const double *prim,
                            double *src) {
  int n[3];
  orc_block_dims(b, n);
  const long nc = (long)n[0] * n[1] * n[2];
  const int np = orc_nprim(p->mhd), ndim = b->ndim, ng = b->ng;
  const int is = ng, ie = ng + b->nx[0] - 1;
  const int js = ndim > 1 ? ng : 0, je = js + b->nx[1] - 1;
  const int ks = ndim > 2 ? ng : 0, ke = ks + b->nx[2] - 1;
#pragma omp parallel for collapse(2)
  for (int k = ks; k <= ke; ++k)
    for (int j = js; j <= je; ++j)
      for (int i = is; i <= ie; ++i) {
        const long c = ((long)k * n[1] + j) * n[0] + i;
        double pr[10], s[5];
        for (int v = 0; v < np; ++v) pr[v] = prim[v * nc + c];
        orc_sources(p, xc(b, 0, i), xc(b, 1, j), xc(b, 2, k), pr, s);
        for (int l = 1; l <= 4; ++l) src[(long)l * nc + c] = s[l];
      }
}

double orc_estimate_dt(const orc_params *p, const orc_block *b, const double *cell_sig,
                       const double *face_sig, double cfl) {
  (void)p;
  int n[3];
  orc_block_dims(b, n);
  const long nc = (long)n[0] * n[1] * n[2];
  const int ndim = b->ndim, ng = b->ng;
  const int is = ng, ie = ng + b->nx[0] - 1;
  const int js = ndim > 1 ? ng : 0, je = js + b->nx[1] - 1;
  const int ks = ndim > 2 ? ng : 0, ke = ks + b->nx[2] - 1;
  const long st[3] = {1, n[0], (long)n[0] * n[1]};
  double min_dt = DBL_MAX;
  for (int k = ks; k <= ke; ++k)
    for (int j = js; j <= je; ++j)
      for (int i = is; i <= ie; ++i) {
        const long c = ((long)k * n[1] + j) * n[0] + i;
        double ldt = 0.0;
        for (int d = 0; d < ndim; d++) {
          double s = cell_sig[d * nc + c];
          if (face_sig) s = smax(s, smax(face_sig[d * nc + c], face_sig[d * nc + c + st[d]]));
          ldt += s / b->dx[d];
        }
        min_dt = smin(min_dt, 1.0 / (ldt + 1.e-50));
      }
  return cfl * min_dt;
}

/* Physical boundaries for conserved variables (bc_vars = conserved), one face after the
 * other in the order x1 inner, x1 outer, x2 inner, ...: outflow copies the first interior
 * layer scaled by ratio(detgamma(ghost), detgamma(ref)) (phoebus_boundaries.cpp:118-150);
 * periodic copies the opposite interior layers (a block that is its own neighbour).  The
 * ghost slab of face d spans the entire block in the other dimensions. */
void orc_apply_bcs(const orc_params *p, const orc_block *b, int nvar, double *q,
                   const int kinds[6]) {
  int n[3];
  orc_block_dims(b, n);
  const long nc = (long)n[0] * n[1] * n[2];
  for (int d = 0; d < b->ndim; ++d) {
    const int is = b->ng, ie = b->ng + b->nx[d] - 1, nint = b->nx[d];
    for (int side = 0; side < 2; ++side) {
      const int periodic = kinds[2 * d + side] == ORC_BC_PERIODIC;
      const int g0 = side == 0 ? 0 : ie + 1, g1 = side == 0 ? is - 1 : n[d] - 1;
      for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
          for (int i = 0; i < n[0]; ++i) {
            int idx[3] = {i, j, k};
            if (idx[d] < g0 || idx[d] > g1) continue;
            int ref[3] = {i, j, k};
            if (periodic)
              ref[d] = side == 0 ? idx[d] + nint : idx[d] - nint;
            else
              ref[d] = side == 0 ? is : ie;
            double gratio = 1.0;
            if (!periodic) {
              orc_point_geom g, gr;
              orc_geom_point(&p->geom, xc(b, 0, idx[0]), xc(b, 1, idx[1]), xc(b, 2, idx[2]), &g);
              orc_geom_point(&p->geom, xc(b, 0, ref[0]), xc(b, 1, ref[1]), xc(b, 2, ref[2]), &gr);
              gratio = orc_ratio(g.detgamma, gr.detgamma);
            }
            const long c = ((long)idx[2] * n[1] + idx[1]) * n[0] + idx[0];
            const long cr = ((long)ref[2] * n[1] + ref[1]) * n[0] + ref[0];
            for (int l = 0; l < nvar; ++l) q[l * nc + c] = gratio * q[l * nc + cr];
          }
    }
  }
}

/* ---- batched con2prim on spherical Kerr-Schild points (r, theta, phi = 0) ------------- */
void orc_c2p_batch(const orc_params *p, long n, const double *coords, double *cons,
                   double *prim, double *mu, double *fail, int *iters) {
  const int np = orc_nprim(p->mhd), ncn = orc_ncons(p->mhd);
#pragma omp parallel for schedule(dynamic, 256)
  for (long s = 0; s < n; ++s) {
    const double X[3] = {coords[s], coords[n + s], 0.0};
    orc_point_geom g;
    orc_geom_point(&p->geom, X[0], X[1], X[2], &g);
    double cs[8], pr[10] = {0}, sig[3], t = 0.0, m = mu[s];
    for (int v = 0; v < ncn; ++v) cs[v] = cons[v * n + s];
    int it = 0;
    const double st = orc_c2p(p, &g, X, cs, pr, &t, &m, sig, &it);
    for (int v = 0; v < ncn; ++v) cons[v * n + s] = cs[v];
    for (int v = 0; v < np; ++v) prim[v * n + s] = pr[v];
    mu[s] = m;
    fail[s] = st;
    if (iters) iters[s] = it;
  }
}

void orc_p2c_batch(const orc_params *p, long n, const double *coords, const double *prim,
                   double *cons) {
  const int np = orc_nprim(p->mhd), ncn = orc_ncons(p->mhd);
#pragma omp parallel for
  for (long s = 0; s < n; ++s) {
    orc_point_geom g;
    orc_geom_point(&p->geom, coords[s], coords[n + s], 0.0, &g);
    double pr[10], cs[8], sig[3];
    for (int v = 0; v < np; ++v) pr[v] = prim[v * n + s];
    orc_p2c(p, &g, pr, cs, sig);
    for (int v = 0; v < ncn; ++v) cons[v * n + s] = cs[v];
  }
}

/* ---- fixups: fixup_c2p.cpp:44-356 and fixup.cpp:285-530 ------------------------------- */
static inline double eos_bmod_rho_sie(const orc_params *p, double rho, double sie) {
  return (p->gm1 + 1.0) * p->gm1 * rho * sie;
}

/* effective switches after fixup::Initialize's forcing rules (fixup.cpp:102-135) */
static void fixup_flags(const orc_params *p, const orc_fixup *fx, int *c2p_fixup, int *floors,
                        int *mhd_floors) {
  *c2p_fixup = fx->enable_c2p_fixup;
  *mhd_floors = fx->enable_mhd_floors && p->mhd;
  *floors = fx->enable_floors || *c2p_fixup || *mhd_floors;
}

/* ConservedToPrimitiveFixupImpl (fixup_c2p.cpp:44-356) for the fluid, interior cells.
 * Failed cells take the average of their valid face neighbours' rho, v, u (or the
 * near-vacuum state), the gamma ceiling with the reference's rescale
 * sqrt(gmax^2 - 1) / (W^2 - 1) (:269-275), then T, P, Gamma_1 and cons (not B) and
 * the cell signal speeds by p2c.  A neighbour that is itself a failed interior cell
 * contributes exactly 0: its weight fail = 0 multiplies a value that the same pass
 * rewrites, which the reference reads racily (only non-finite values make it differ). */
void orc_c2p_fixup_block(const orc_params *p, const orc_fixup *fx, const orc_block *b,
                         double *prim, double *cons, double *temp, double *cell_sig,
                         const double *fail) {
  int c2p_fixup, floors, mhd_floors;
  fixup_flags(p, fx, &c2p_fixup, &floors, &mhd_floors);
  if (!c2p_fixup) return; /* :120-122 */
  int n[3];
  orc_block_dims(b, n);
  const long nc = (long)n[0] * n[1] * n[2];
  const int mhd = p->mhd, prs = mhd ? 8 : 5, gm1i = mhd ? 9 : 6;
  int lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    lo[d] = d < b->ndim ? b->ng : 0;
    hi[d] = lo[d] + b->nx[d] - 1;
  }
  const long st[3] = {1, n[0], (long)n[0] * n[1]};
  const int pv[5] = {0, 1, 2, 3, 4}; /* rho, v1..3, u (:229-234) */
#pragma omp parallel for collapse(2)
  for (int k = lo[2]; k <= hi[2]; ++k)
    for (int j = lo[1]; j <= hi[1]; ++j)
      for (int i = lo[0]; i <= hi[0]; ++i) {
        const long c = ((long)k * n[1] + j) * n[0] + i;
        if (fail[c] != 0.0) continue; /* FailFlags::fail == 0 */
        const int idx[3] = {i, j, k};
        /* neighbour (d, side) at q = 2d + side: weight = its fail flag; a failed
         * interior neighbour is skipped (it is rewritten by this same pass) */
        double w[6] = {0, 0, 0, 0, 0, 0};
        long nb[6] = {c, c, c, c, c, c};
        int skip[6] = {1, 1, 1, 1, 1, 1};
        for (int d = 0; d < b->ndim; ++d)
          for (int s = 0; s < 2; ++s) {
            const int q = 2 * d + s, ni = idx[d] + (s == 0 ? -1 : 1);
            nb[q] = c + (s == 0 ? -st[d] : st[d]);
            w[q] = fail[nb[q]];
            skip[q] = w[q] == 0.0 && ni >= lo[d] && ni <= hi[d];
          }
        double num_valid = w[0] + w[1]; /* :218-221 */
        if (b->ndim > 1) num_valid += w[2] + w[3];
        if (b->ndim == 3) num_valid += w[4] + w[5];
        double pr[10];
        for (int v = 0; v < orc_nprim(mhd); ++v) pr[v] = prim[v * nc + c];
        if (num_valid > 0.5 && fx->failure_strategy == ORC_FAILURE_INTERPOLATE) {
          const double norm = 1.0 / num_valid;
          for (int m = 0; m < 5; ++m) {
            const int iv = pv[m];
            double term[6];
            for (int q = 0; q < 6; ++q) term[q] = skip[q] ? 0.0 : w[q] * prim[iv * nc + nb[q]];
            double acc = term[0] + term[1];
            if (b->ndim > 1) {
              acc += term[2] + term[3];
              if (b->ndim == 3) acc += term[4] + term[5];
            }
            pr[iv] = norm * acc;
          }
          pr[0] = smax(pr[0], 100. * ORC_SMALL);
          pr[4] = smax(pr[4], 100. * ORC_SMALL);
        } else {
          pr[0] = 100. * ORC_SMALL;
          pr[4] = 100. * ORC_SMALL;
          for (int m = 0; m < 3; ++m) pr[1 + m] = 0.;
        }
        const double X[3] = {xc(b, 0, i), xc(b, 1, j), xc(b, 2, k)};
        orc_point_geom g;
        orc_geom_point(&p->geom, X[0], X[1], X[2], &g);
        const double gamma_max = p->bounds.gam_max;
        double vpcon[3] = {pr[1], pr[2], pr[3]};
        double W = lorentz(vpcon, g.gcov);
        if (W > gamma_max) { /* :269-275 */
          const double rescale = sqrt(gamma_max * gamma_max - 1.) / (W * W - 1.);
          for (int m = 0; m < 3; ++m) vpcon[m] *= rescale;
          for (int m = 0; m < 3; ++m) pr[1 + m] = vpcon[m];
          W = gamma_max;
        }
        const double t = eos_t_rho_sie(p, pr[0], orc_ratio(pr[4], pr[0])); /* :277-287 */
        pr[prs] = eos_p_rho_t(p, pr[0], t);
        pr[gm1i] = orc_ratio(eos_bmod_rho_t(p, pr[0], t), pr[prs]);
        double bp[3] = {0.0, 0.0, 0.0};
        if (mhd)
          for (int m = 0; m < 3; ++m) bp[m] = pr[5 + m];
        double D, S[3], bc[3], tau, sig[3];
        p2c_core(pr[0], vpcon, bp, pr[4], pr[prs], pr[gm1i], g.gcov, g.gcon, g.beta, g.alpha,
                 g.detgamma, &D, S, bc, &tau, sig); /* :290-312 */
        for (int v = 0; v < orc_nprim(mhd); ++v) prim[v * nc + c] = pr[v];
        if (temp) temp[c] = t;
        cons[c] = D;
        for (int m = 0; m < 3; ++m) cons[(1 + m) * nc + c] = S[m];
        cons[4 * nc + c] = tau;
        for (int d = 0; d < b->ndim; ++d) cell_sig[d * nc + c] = sig[d];
      }
}

/* ApplyFloorsImpl (fixup.cpp:285-530) for the fluid over cells lo..hi (the reference's
 * default domain is the entire block).  Density/energy floors (and the MHD b^2 ceilings)
 * are injected as a fluid at rest in the normal-observer frame, added to the conserved
 * state and re-inverted by the robust con2prim (falling back to the floor state when it
 * fails); then the gamma / sie ceilings with the rescale quirk of :510; then T, P,
 * Gamma_1.  mu and cell_sig are the con2prim warm start and outputs. */
void orc_apply_floors_block(const orc_params *p, const orc_fixup *fx, const orc_block *b,
                            const int lo[3], const int hi[3], double *prim, double *cons,
                            double *temp, double *mu, double *cell_sig) {
  int c2p_fixup, floors, mhd_floors;
  fixup_flags(p, fx, &c2p_fixup, &floors, &mhd_floors);
  if (!floors) return; /* :309 */
  int n[3];
  orc_block_dims(b, n);
  const long nc = (long)n[0] * n[1] * n[2];
  const int mhd = p->mhd, prs = mhd ? 8 : 5, gm1i = mhd ? 9 : 6;
  const int np = orc_nprim(mhd), ncn = orc_ncons(mhd);
#pragma omp parallel for collapse(2)
  for (int k = lo[2]; k <= hi[2]; ++k)
    for (int j = lo[1]; j <= hi[1]; ++j)
      for (int i = lo[0]; i <= hi[0]; ++i) {
        const long c = ((long)k * n[1] + j) * n[0] + i;
        const double X[3] = {xc(b, 0, i), xc(b, 1, j), xc(b, 2, k)};
        double rho_floor, sie_floor;
        orc_get_floors(&p->bounds, X[0], X[1], X[2], &rho_floor, &sie_floor);
        const double gamma_max = p->bounds.gam_max, e_max = p->bounds.sie_max;
        const double bsqorho_max = fx->bsqorho_max, bsqou_max = fx->bsqou_max;
        double rho_floor_max = rho_floor;
        double u_floor_max = rho_floor * sie_floor;
        int floor_applied = 0, ceiling_applied = 0;
        orc_point_geom g;
        orc_geom_point(&p->geom, X[0], X[1], X[2], &g);
        double pr[10], cs[8];
        for (int v = 0; v < np; ++v) pr[v] = prim[v * nc + c];
        for (int v = 0; v < ncn; ++v) cs[v] = cons[v * nc + c];
        double t = temp ? temp[c] : 0.0;
        if (mhd_floors) { /* :411-428 */
          double Bsq = 0.0, Bdotv = 0.0;
          const double vp[3] = {pr[1], pr[2], pr[3]};
          const double bp[3] = {pr[5], pr[6], pr[7]};
          const double W = lorentz(vp, g.gcov);
          const double iW = 1.0 / W;
          for (int ii = 0; ii < 3; ++ii)
            for (int jj = 0; jj < 3; ++jj) {
              Bsq += g.gcov[ii + 1][jj + 1] * bp[ii] * bp[jj];
              Bdotv += g.gcov[ii + 1][jj + 1] * bp[ii] * vp[jj];
            }
          const double bcon0 = W * Bdotv / g.alpha;
          const double bsq = (Bsq + g.alpha * g.alpha * bcon0 * bcon0) * iW * iW;
          rho_floor_max = smax(rho_floor_max, bsq / bsqorho_max);
          u_floor_max = smax(u_floor_max, bsq / bsqou_max);
          rho_floor_max = smax(rho_floor_max, smax(pr[4], u_floor_max) / e_max);
        }
        double drho = rho_floor_max - pr[0];
        double du = u_floor_max - pr[4];
        if (drho > 0. || du > 0.) { /* :434-440 */
          floor_applied = 1;
          drho = smax(drho, du / sie_floor);
          du = smax(du, sie_floor * drho);
        }
        double bp[3] = {0.0, 0.0, 0.0};
        if (mhd)
          for (int m = 0; m < 3; ++m) bp[m] = pr[5 + m];
        double dS[3], dB[3];
        if (floor_applied) { /* :448-498 */
          const double vzero[3] = {0.0, 0.0, 0.0};
          const double dprs = eos_p_rho_sie(p, drho, orc_ratio(du, drho));
          const double dgm1 = orc_ratio(eos_bmod_rho_sie(p, drho, orc_ratio(du, drho)), dprs);
          double dcrho, dtau;
          p2c_core(drho, vzero, bp, du, dprs, dgm1, g.gcov, g.gcon, g.beta, g.alpha,
                   g.detgamma, &dcrho, dS, dB, &dtau, NULL);
          cs[0] += dcrho;
          for (int m = 0; m < 3; ++m) cs[1 + m] += dS[m];
          cs[4] += dtau;
          double sig[3], m_io = mu[c];
          for (int d = 0; d < 3; ++d) sig[d] = d < b->ndim ? cell_sig[d * nc + c] : 0.0;
          int it = 0;
          const double status = orc_c2p(p, &g, X, cs, pr, &t, &m_io, sig, &it);
          mu[c] = m_io;
          for (int d = 0; d < b->ndim; ++d) cell_sig[d * nc + c] = sig[d];
          if (status == 0.0) { /* :473-497 */
            pr[0] = drho;
            for (int m = 0; m < 3; ++m) pr[1 + m] = vzero[m];
            pr[4] = du;
            t = eos_t_rho_sie(p, pr[0], pr[4] / pr[0]);
            double tau;
            p2c_core(drho, vzero, bp, du, dprs, dgm1, g.gcov, g.gcon, g.beta, g.alpha,
                     g.detgamma, &cs[0], dS, dB, &tau, NULL);
            cs[4] = tau;
            for (int m = 0; m < 3; ++m) cs[1 + m] = dS[m];
          }
        }
        double vpcon[3] = {pr[1], pr[2], pr[3]}; /* :501-507 */
        const double W = lorentz(vpcon, g.gcov);
        if (W > gamma_max || pr[4] / pr[0] > e_max) ceiling_applied = 1;
        if (ceiling_applied) { /* :509-526 */
          const double rescale = sqrt(gamma_max * gamma_max - 1.) / (W * W - 1.);
          for (int m = 0; m < 3; ++m) vpcon[m] *= rescale;
          for (int m = 0; m < 3; ++m) pr[1 + m] = vpcon[m];
          double tau;
          p2c_core(pr[0], vpcon, bp, pr[4], pr[prs], pr[gm1i], g.gcov, g.gcon, g.beta,
                   g.alpha, g.detgamma, &cs[0], dS, dB, &tau, NULL);
          cs[4] = tau;
          for (int m = 0; m < 3; ++m) cs[1 + m] = dS[m];
        }
        if (floor_applied || ceiling_applied) { /* :528-539 */
          t = eos_t_rho_sie(p, pr[0], orc_ratio(pr[4], pr[0]));
          pr[prs] = eos_p_rho_t(p, pr[0], t);
          pr[gm1i] = orc_ratio(eos_bmod_rho_t(p, pr[0], t), pr[prs]);
          for (int v = 0; v < np; ++v) prim[v * nc + c] = pr[v];
          for (int v = 0; v < ncn; ++v) cons[v * nc + c] = cs[v];
          if (temp) temp[c] = t;
        }
      }
}

/* ConvertBoundaryConditions' outer-x1 inflow check (phoebus_boundaries.cpp:356-380) on
 * cells lo..hi: where u^1 = W v^1 - W beta^1 / alpha < 0 the normal-observer velocity's
 * x1 component is replaced by beta^1 / alpha and W recomputed. */
void orc_inflow_check_block(const orc_params *p, const orc_block *b, const int lo[3],
                            const int hi[3], double *prim) {
  int n[3];
  orc_block_dims(b, n);
  const long nc = (long)n[0] * n[1] * n[2];
  for (int k = lo[2]; k <= hi[2]; ++k)
    for (int j = lo[1]; j <= hi[1]; ++j)
      for (int i = lo[0]; i <= hi[0]; ++i) {
        const long c = ((long)k * n[1] + j) * n[0] + i;
        orc_point_geom g;
        orc_geom_point(&p->geom, xc(b, 0, i), xc(b, 1, j), xc(b, 2, k), &g);
        double vcon[3] = {prim[nc + c], prim[2 * nc + c], prim[3 * nc + c]};
        double W = lorentz(vcon, g.gcov);
        const double ucon1 = vcon[0] - W * g.beta[0] / g.alpha;
        if (ucon1 < 0) {
          for (int ii = 0; ii < 3; ++ii) vcon[ii] /= W;
          vcon[0] = g.beta[0] / g.alpha;
          double vsq = 0.;
          for (int ii = 0; ii < 3; ++ii)
            for (int jj = 0; jj < 3; ++jj) vsq += g.gcov[ii + 1][jj + 1] * vcon[ii] * vcon[jj];
          W = 1. / sqrt(1. - vsq);
          for (int ii = 0; ii < 3; ++ii) prim[(1 + ii) * nc + c] = W * vcon[ii];
        }
      }
}
