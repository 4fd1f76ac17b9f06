/*
 * phx_oracle.h -- CPU restatement (plain C99 + OpenMP) of phoebus's per-RK-stage fluid
 * path, used ONLY as test infrastructure: the parity checker of the HIP path and the
 * timed CPU baseline.  Nothing in the product links or calls this code.
 *
 * Every function cites the reference file:line it restates.  Arithmetic follows the
 * reference's operation order and is compiled with -ffp-contract=off so that, for
 * +,-,*,/,sqrt (all correctly rounded), results are bit-reproducible against a HIP
 * build with the same contraction setting.
 *
 * Pinning (how each part is checked; see tests/test_oracle_*.py):
 *   recon       closed forms (constant/linear data), mirror symmetry, convergence order
 *   geometry    reference unit-test known answers (tst/unit/geometry/test_geometry.cpp
 *               :400-470 spherical KS at r=6, a=0.9), finite differences of the metric
 *   p2c         independent T^{mu nu} contraction; known Lorentz factor
 *               (tst/unit/phoebus_utils/test_relativity_utils.cpp:90-97)
 *   fluxes      U_L = U_R reduces HLL and LLF to the physical flux (independent T^{mu nu}
 *               form); static state; sound speed closed form
 *   c2p         p2c -> c2p round trips; B=0,v=0 closed form mu = 1/h; forced failures
 *   divergence  telescoping sums over periodic blocks (conserved totals to round-off)
 *   FluxCT      corner div B unchanged to round-off
 *   dt          brute-force minimum
 *   EOS         parity unpinned at the ulp level: singularity-eos (IdealGas in a
 *               UnitSystem wrapper, version not recorded in the reference snapshot) is
 *               absent; restated as its published closed forms with unit scales of 1.
 *   parthenon Update::FluxDivergence / AverageIndependentData / UpdateIndependentData:
 *               op order unpinned at the ulp level (parthenon absent); restated from its
 *               published algorithm, pinned by telescoping and SSP-RK2 closed forms.
 */
#ifndef PHX_ORACLE_H_
#define PHX_ORACLE_H_

#ifdef __cplusplus
extern "C" {
#endif

/* ---- geometry ---------------------------------------------------------------------- */
enum { ORC_GEOM_MINKOWSKI = 0, ORC_GEOM_SPHKS = 1, ORC_GEOM_FMKS = 2 };

/* FMKS = spherical Kerr-Schild under the McKinney-Gammie-Ryan transformation
 * (mckinney_gammie_ryan.hpp:45-209, modified_system.hpp:52-248): r = exp(X1) (with the
 * hyper-exponential stretch beyond hexp_br), theta(X1, X2) with derefined poles. */
typedef struct {
  int derefine_poles;
  double h, xt, alpha, x0, smooth, hexp_br, hexp_nsq, hexp_csq;
} orc_mgr;

typedef struct {
  int kind;
  double a; /* spin for Kerr-Schild */
  orc_mgr mgr;
} orc_geom;

/* metric quantities at one point (Analytic<...> interface, analytic_system.hpp:157-446) */
typedef struct {
  double alpha, beta[3], gcov[4][4], gcon[3][3], gam[3][3], detgamma, detg;
} orc_point_geom;

void orc_geom_point(const orc_geom *g, double X1, double X2, double X3, orc_point_geom *out);
/* the MGR transformation: C = (r, theta, phi), Jcov = d(r,th,ph)/dX, Jcon = its inverse */
void orc_mgr_transform(const orc_mgr *m, double X1, double X2, double X3, double C[3],
                       double Jcov[3][3], double Jcon[3][3]);
void orc_geom_gcon4(const orc_geom *g, double X1, double X2, double X3, double gi[4][4]);
void orc_geom_dg(const orc_geom *g, double X1, double X2, double X3, double dg[4][4][4]);
void orc_geom_gradlnalpha(const orc_geom *g, double X1, double X2, double X3, double da[4]);

/* ---- bounds (fixup/fixup.hpp:72-327) ------------------------------------------------ */
enum { ORC_FLOOR_CONSTANT = 0, ORC_FLOOR_EXPX1_RHO_SIE = 1, ORC_FLOOR_EXPX1_RHO_U = 2,
       ORC_FLOOR_X1_RHO_SIE = 3, ORC_FLOOR_R_RHO_SIE = 4 };

typedef struct {
  int floor_type;
  double r0, s0, ralpha, salpha;  /* floor parameters */
  double rho_min_eos, sie_min_eos; /* SetEOSBnds (fixup.hpp:136-142) */
  double gam_max, sie_max;         /* ConstantGamSie ceiling */
} orc_bounds;

void orc_get_floors(const orc_bounds *b, double x1, double x2, double x3, double *rflr,
                    double *sflr);

/* ---- fluid parameters ---------------------------------------------------------------- */
enum { ORC_RECON_CONSTANT = 0, ORC_RECON_LINEAR = 1, ORC_RECON_WENO5 = 2, ORC_RECON_MP5 = 3 };
enum { ORC_RIEMANN_LLF = 0, ORC_RIEMANN_HLL = 1 };

typedef struct {
  double gm1, cv; /* ideal gas: Gamma - 1, specific heat */
  int mhd;
  int recon, riemann;
  double c2p_tol;
  int c2p_max_iter;
  double c2p_floor_scale_fac;
  int fail_on_floors, fail_on_ceilings;
  orc_bounds bounds;
  orc_geom geom;
} orc_params;

/* ---- point functions ------------------------------------------------------------------ */
double orc_ratio(double a, double b);
void orc_recon(int type, const double *q /* 5-point stencil q[-2..2] at q[2] */,
               double *ql, double *qr);
void orc_plm(double qm, double q0, double qp, double *ql, double *qr);
void orc_weno5z(double q0, double q1, double q2, double q3, double q4, double *ql, double *qr);
void orc_mp5(double q0, double q1, double q2, double q3, double q4, double *ql, double *qr);

/* prim layout for point functions and blocks: rho, v1..3, u, [B1..3], P, gamma1 */
int orc_nprim(int mhd);
int orc_ncons(int mhd);

/* riemann.hpp:88-209.  q = prim vector, face geometry from orc_geom_point at the face. */
void orc_prim_to_flux(const orc_params *p, int dir, const orc_point_geom *g, const double X[3],
                      const double *q, double sie_max, double gam_max, double *vm,
                      double *vp, double *U, double *F);
/* riemann.hpp:257-301: returns the face signal speed, writes ncons fluxes */
double orc_face_flux(const orc_params *p, int dir, const double X[3], const double *ql,
                     const double *qr, double *flux);

/* prim2con.hpp:25-102 */
void orc_p2c(const orc_params *p, const orc_point_geom *g, const double *prim, double *cons,
             double sig[3]);
void orc_signal_speeds(double rho, double u, double pr, double bsq, const double v[3],
                       double vsq, double gam1, double alpha, const double beta[3],
                       const double gcon[3][3], double sig[3]);

/* con2prim_robust.hpp:299-453.  cons: D, S1..3, tau, [B1..3] (densitized).
 * prim out: rho, v1..3, u, [B1..3], P, gamma1.  Returns 1.0 (success) / 0.0 (failure),
 * the reference's FailFlags encoding (con2prim_robust.hpp:41-44).  *mu in: guess, out:
 * root; *iters out: RootFind::iteration_count.  cons is rewritten when bounds were hit.
 * bracket_lo/hi: the initial bracket (0,1 in the reference). */
double orc_c2p(const orc_params *p, const orc_point_geom *g, const double X[3], double *cons,
               double *prim, double *temp, double *mu, double sig[3], int *iters);

/* Residual (con2prim_robust.hpp:121-141), exposed for tests */
double orc_c2p_residual(const orc_params *p, const orc_point_geom *g, const double X[3],
                        const double *cons, double mu);

/* root_find.hpp:109-167 on a C callback; returns root, writes iteration count */
typedef double (*orc_fn)(double x, void *ctx);
double orc_regula_falsi(orc_fn f, void *ctx, double a, double b, double tol, double guess,
                        int max_iter, int *iters, int *bracketed);

/* T^{mu nu} (tmunu.hpp:35-137) and the geometric source terms (fluid.cpp:593-654) */
void orc_tmunu(const orc_params *p, double X1, double X2, double X3, const double *prim,
               double T[4][4]);
void orc_sources(const orc_params *p, double X1, double X2, double X3, const double *prim,
                 double src[5]);

/* ---- block functions -------------------------------------------------------------------
 * A block holds (nk, nj, ni) cells including ng ghosts in every active dimension.
 * Arrays are SoA, i fastest: var[n][k][j][i].  Logical coordinates follow parthenon's
 * uniform coordinates: Xf_d(idx) = xmin_d + (idx - is_d)*dx_d, Xc_d = xmin_d +
 * (idx - is_d + 0.5)*dx_d with is_d = ng in active dimensions and 0 otherwise. */
typedef struct {
  int ndim, nx[3], ng;
  double xmin[3], dx[3];
} orc_block;

void orc_block_dims(const orc_block *b, int n[3] /* ni, nj, nk */);

/* fluid.cpp:659-773 (reconstruction + Riemann, all directions): flux[d][m][cells],
 * face_sig[d][cells]. */
void orc_calculate_fluxes(const orc_params *p, const orc_block *b, const double *prim,
                          double *flux, double *face_sig);
/* fluid.cpp:775-841 */
void orc_flux_ct(const orc_params *p, const orc_block *b, double *flux, double *emf);
/* parthenon Update::FluxDivergence + SumData + AverageIndependentData +
 * UpdateIndependentData as sequenced at phoebus_driver.cpp:560-576 (interior cells) */
void orc_flux_divergence_update(const orc_params *p, const orc_block *b, const double *flux,
                                const double *src, double *u_stage, const double *u_base,
                                double beta, double dt, double *u_out);
/* fluid.cpp:478-517 over a cell range [lo, hi] (inclusive, per dim ijk) */
void orc_con2prim_block(const orc_params *p, const orc_block *b, const int lo[3],
                        const int hi[3], double *cons, double *prim, double *temp,
                        double *mu, double *cell_sig, double *fail, int *iters);
/* fluid.cpp:359-449 */
void orc_prim_to_con_block(const orc_params *p, const orc_block *b, const int lo[3],
                           const int hi[3], const double *prim, double *cons,
                           double *cell_sig);
/* fluid.cpp:566-657: src has ncons slots per cell, only S_i and tau written */
void orc_source_terms_block(const orc_params *p, const orc_block *b, const double *prim,
                            double *src);
/* fluid.cpp:899-944; face_sig NULL selects the first-call form */
double orc_estimate_dt(const orc_params *p, const orc_block *b, const double *cell_sig,
                       const double *face_sig, double cfl);

/* phoebus_boundaries.cpp:118-150 (outflow, conserved) and single-block periodic fills,
 * applied face by face; kinds[2*d + side] */
enum { ORC_BC_OUTFLOW = 0, ORC_BC_PERIODIC = 1 };
void orc_apply_bcs(const orc_params *p, const orc_block *b, int nvar, double *q,
                   const int kinds[6]);

/* batched con2prim over independent states with spherical-KS geometry at (r, theta)
 * per state: cons[8][n] in, prim[nprim][n] out */
void orc_c2p_batch(const orc_params *p, long n, const double *coords /* [2][n] */,
                   double *cons, double *prim, double *mu, double *fail, int *iters);
void orc_p2c_batch(const orc_params *p, long n, const double *coords, const double *prim,
                   double *cons);

/* ---- fixups (fixup_c2p.cpp:44-356, fixup.cpp:285-530) -------------------------------
 * Switches as the fixup package reads them (fixup.cpp:63-135, 217-243); enable_floors is
 * forced on by enable_c2p_fixup or enable_mhd_floors, and MHD floors need MHD. */
enum { ORC_FAILURE_INTERPOLATE = 0, ORC_FAILURE_FLOORS = 1 };
typedef struct {
  int enable_c2p_fixup, failure_strategy;
  int enable_floors, enable_mhd_floors;
  double bsqorho_max, bsqou_max; /* MHDCeilings ConstantBsqRat */
} orc_fixup;

/* interior cells; fail is the con2prim flag field */
void orc_c2p_fixup_block(const orc_params *p, const orc_fixup *fx, const orc_block *b,
                         double *prim, double *cons, double *temp, double *cell_sig,
                         const double *fail);
/* cells lo..hi (inclusive, i j k) */
void orc_apply_floors_block(const orc_params *p, const orc_fixup *fx, const orc_block *b,
                            const int lo[3], const int hi[3], double *prim, double *cons,
                            double *temp, double *mu, double *cell_sig);

/* the outer-x1 inflow check of ConvertBoundaryConditions (phoebus_boundaries.cpp:356-380)
 * on the primitives of cells lo..hi */
void orc_inflow_check_block(const orc_params *p, const orc_block *b, const int lo[3],
                            const int hi[3], double *prim);

int orc_num_threads(void);
void orc_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
