/*
 * phx.h -- C ABI of the MI355X-native phoebus fluid-stage library (libphx.so).
 *
 * This is the drop-in boundary for phoebus's per-RK-stage fluid path.  Each entry point
 * replaces one task of the reference (file:line cited at each declaration).  The ABI
 * contains plain pointers and sizes only.
 *
 * Conventions
 *  - All array arguments are DEVICE pointers, caller-owned, FP64, structure-of-arrays
 *    with the cell index i fastest: a field with nv components over a block is laid out
 *    as [nv][nk][nj][ni], where (ni, nj, nk) include ng ghost layers in every active
 *    dimension (ng = 4 for WENO5/MP5, as in the reference).
 *  - Primitive ordering is the reference's reconstruction pack (fluid.cpp:272-286):
 *      MHD   : rho, v1, v2, v3, u, B1, B2, B3, P, gamma1     (nprim = 10)
 *      hydro : rho, v1, v2, v3, u, P, gamma1                 (nprim = 7)
 *    where v^i is the phoebus primitive velocity (W times the normal-observer 3-velocity)
 *    and u = rho * eps.  Conserved ordering: D, S1, S2, S3, tau, [B1, B2, B3]
 *    (ncons = 8 or 5), densitized by sqrt(gamma); tau is the Valencia energy.
 *  - Face fluxes: flux[d][m][k][j][i] is the flux through the LOWER face of cell (k,j,i)
 *    in direction d, as in the reference (riemann.hpp:296, fluid.cpp:753-770).
 *  - Every call is asynchronous on the given stream (NULL = the context's stream) and
 *    returns PHX_OK or a negative error code.  No call aborts or throws; per-zone
 *    failures are data (the fail field, 1.0 = success, 0.0 = failure, as
 *    con2prim_robust.hpp:41-44).
 *  - No entry point allocates device memory after phx_create; scratch lives in the
 *    context.  A context is bound to one device and is not thread-safe.
 *  - The context's scratch (the CFL partial minima written by phx_fluid_stage and
 *    phx_estimate_dt when dt_out is set) is one buffer per context: calls that use it
 *    must be ordered, i.e. issued on one stream, or on streams the caller orders with
 *    events.  Two such calls in flight at once on different streams race on it.
 */
#ifndef PHX_H_
#define PHX_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct phx_ctx phx_ctx;
typedef void *phx_stream; /* hipStream_t */

enum {
  PHX_OK = 0,
  PHX_ERR_INVALID = -1,     /* bad argument or configuration (reference: PARTHENON_THROW) */
  PHX_ERR_UNSUPPORTED = -2, /* configuration the library does not implement */
  PHX_ERR_DEVICE = -3,      /* HIP runtime error */
  PHX_ERR_NOMEM = -4
};

enum { PHX_GEOM_MINKOWSKI = 0, PHX_GEOM_SPHERICAL_KS = 1 };
enum { PHX_RECON_CONSTANT = 0, PHX_RECON_LINEAR = 1, PHX_RECON_WENO5Z = 2, PHX_RECON_MP5 = 3 };
enum { PHX_RIEMANN_LLF = 0, PHX_RIEMANN_HLL = 1 };
enum {
  PHX_FLOOR_CONSTANT_RHO_SIE = 0,
  PHX_FLOOR_EXPX1_RHO_SIE = 1,
  PHX_FLOOR_EXPX1_RHO_U = 2,
  PHX_FLOOR_X1_RHO_SIE = 3,
  PHX_FLOOR_R_RHO_SIE = 4
};

/* Analytic metric selection (reference: compile-time Geometry::CoordSysMeshBlock,
 * cmake/Geometry.cmake:17-41).  Minkowski uses Cartesian X; spherical Kerr-Schild uses
 * X = (r, theta, phi) and spin a (spherical_kerr_schild.hpp:42-243); FMKS uses
 * X = (log r, x2 in [0, 1], phi): spherical Kerr-Schild under the McKinney-Gammie-Ryan
 * transformation (mckinney_gammie_ryan.hpp:45-209, modified_system.hpp:52-248) with the
 * parameters of fmks.cpp:37-52 (coordinates/derefine_poles, hslope, poly_xt,
 * poly_alpha, smooth, hexp_*; x0 = min(x1min, max(0, x1min - nghost dx1))).  FMKS metric
 * derivatives are central finite differences as geometry_utils.hpp:94-187. */
enum { PHX_GEOM_FMKS = 2 };
typedef struct {
  int derefine_poles;
  double h, xt, alpha, x0, smooth;
  double hexp_br, hexp_nsq, hexp_csq;
} phx_mgr;

typedef struct {
  int kind;
  double a;
  phx_mgr mgr; /* FMKS only */
} phx_geometry;

/* fixup::Bounds (fixup.hpp:72-327): floors of floor_type with (rho0, sie0, rho_exp,
 * sie_exp), clamped below by the EOS minima; constant gamma / sie ceilings.  Disabled
 * floors are rho0 = sie0 = -DBL_MAX; disabled ceilings are DBL_MAX. */
typedef struct {
  int floor_type;
  double rho0, sie0, rho_exp, sie_exp;
  double rho_min_eos, sie_min_eos;
  double gam_max, sie_max;
} phx_bounds;

/* Arithmetic of the flux sweeps (reconstruction + Riemann flux; no reference counterpart).
 * EXACT: IEEE division and square root without contraction, rounding as the reference
 * (fluxes bit-identical to the reference's operation order).  FAST: fused multiply-adds
 * and Newton-refined hardware reciprocals and reciprocal square roots in the sweeps only
 * (fluxes within about 1e-14 relative of EXACT, well inside a 1e-12 flux tolerance); every
 * other kernel is the same in both. */
enum { PHX_FLUX_ARITH_EXACT = 0, PHX_FLUX_ARITH_FAST = 1 };

/* The fluid package parameters the path reads (fluid.cpp:44-136) plus the ideal-gas EOS
 * (eos_phoebus.cpp:73-95).  Ye is not supported (PHX_ERR_UNSUPPORTED). */
typedef struct {
  double gamma, cv;
  int mhd;
  int recon, riemann;
  double c2p_tol;
  int c2p_max_iter;
  double c2p_floor_scale_fac;
  int c2p_fail_on_floors, c2p_fail_on_ceilings;
  phx_bounds bounds;
  phx_geometry geom;
  int flux_arith; /* PHX_FLUX_ARITH_EXACT (0) or PHX_FLUX_ARITH_FAST */
} phx_fluid_params;

/* One meshblock: interior cells nx[3] (nx = 1 in inactive dimensions), ng ghost layers
 * in active dimensions, uniform logical coordinates Xf_d(idx) = xmin_d + (idx - is_d)
 * * dx_d with is_d = ng (active) or 0 (inactive), as parthenon's uniform coordinates. */
typedef struct {
  int ndim, nx[3], ng;
  double xmin[3], dx[3];
} phx_block;

/* Per-component device pointers of a block's multi-component fields, for hosts whose
 * variables are allocated separately: parthenon gives every variable, and every direction
 * of its flux, an allocation of its own (the recon and flux packs of CalculateFluxes,
 * PackVariablesAndFluxes at riemann.hpp:223-250, indexed per variable at fluid.cpp:686-707).
 * A non-NULL entry is the base of that component's [nk][nj][ni] array and is used in place
 * of the component's slice of the contiguous field (prim + v * cells, flux + (d * ncons +
 * m) * cells, ...); NULL entries keep the contiguous slice, so a field may be passed as a
 * table alone (its contiguous pointer NULL).  The components of one parthenon vector
 * variable sit in its one allocation: component c at base + c * cells.  The kernels take
 * these bases as launch arguments and index them directly; nothing is gathered or copied. */
typedef struct {
  double *prim[10];   /* rho, v1, v2, v3, u, B1, B2, B3, P, gamma1 (hydro: rho, v1..3, u, P,
                         gamma1 in entries 0..6) */
  double *cons[8];    /* D, S1, S2, S3, tau, B1, B2, B3 */
  double *flux[3][8]; /* flux[d][m]: direction d, conserved component m */
  double *src[8];     /* geometric sources, conserved order */
} phx_block_vars;

/* Device fields of one block (any may be NULL where the call does not use it). */
typedef struct {
  double *prim;     /* [nprim][cells]                                   */
  double *cons;     /* [ncons][cells]                                   */
  double *temp;     /* [cells] temperature                              */
  double *c2p_mu;   /* [cells] con2prim root (warm start, fluid.cpp:224) */
  double *fail;     /* [cells] 1.0 success / 0.0 failure                */
  double *cell_sig; /* [ndim][cells] cell signal speeds                  */
  double *face_sig; /* [ndim][cells] face signal speeds                  */
  double *flux;     /* [ndim][ncons][cells]                              */
  double *emf;      /* [3][cells] corner EMFs (MHD, ndim > 1)            */
  double *src;      /* [ncons][cells] geometric sources (S_i, tau slots) */
  const phx_block_vars *vars; /* per-component overrides of prim, cons, flux, src; nullable */
} phx_block_fields;

/* ---- context ---------------------------------------------------------------------- */
int phx_create(phx_ctx **ctx, int device);
int phx_destroy(phx_ctx *ctx);
const char *phx_last_error(const phx_ctx *ctx);

/* ---- per-block task entry points --------------------------------------------------- */

/* fluid::PrimitiveToConservedRegion (fluid.cpp:359-449) over cells lo..hi (inclusive,
 * (i, j, k) order): cons and cell_sig from prim. */
int phx_prim_to_con(phx_ctx *ctx, const phx_fluid_params *p, const phx_block *b,
                    const phx_block_fields *f, const int lo[3], const int hi[3],
                    phx_stream stream);

/* fluid::CalculateFluxes (fluid.cpp:659-773): reconstruction (ReconLoop,
 * reconstruction.hpp:246-271) fused with the HLL/LLF flux (riemann.hpp:257-301).  Writes
 * f->flux and f->face_sig on faces k in [ks-dk, ke+dk], j in [js-dj, je+dj],
 * i in [is-1, ie+1] for every direction (the reference's ranges).  The face states are
 * never materialized (no ql/qr arrays). */
int phx_calculate_fluxes(phx_ctx *ctx, const phx_fluid_params *p, const phx_block *b,
                         const phx_block_fields *f, phx_stream stream);

/* fluid::FluxCT (fluid.cpp:775-841): corner EMFs into f->emf, then B fluxes.  No-op for
 * hydro or 1-D. */
int phx_flux_ct(phx_ctx *ctx, const phx_fluid_params *p, const phx_block *b,
                const phx_block_fields *f, phx_stream stream);

/* fluid::CalculateFluidSourceTerms (fluid.cpp:566-657): geometric sources into f->src
 * (S_i and tau slots, interior cells).  Identically zero for Minkowski. */
int phx_fluid_source_terms(phx_ctx *ctx, const phx_fluid_params *p, const phx_block *b,
                           const phx_block_fields *f, phx_stream stream);

/* parthenon Update::FluxDivergence + SumData(src) + AverageIndependentData(beta) +
 * UpdateIndependentData(beta*dt) as sequenced at phoebus_driver.cpp:560-576, interior
 * cells: u_stage <- beta*u_stage + (1-beta)*u_base; u_out <- u_stage + beta*dt*dU/dt.
 * src may be NULL.  u_out may alias u_base (the rk2 second stage does). */
int phx_flux_divergence_update(phx_ctx *ctx, const phx_fluid_params *p, const phx_block *b,
                               const double *flux, const double *src, double *u_stage,
                               const double *u_base, double beta, double dt, double *u_out,
                               phx_stream stream);

/* fluid::ConservedToPrimitiveRobust (fluid.cpp:478-517) with con2prim_robust::ConToPrim
 * (con2prim_robust.hpp:224-453) over cells lo..hi.  Reads f->cons and f->c2p_mu (guess);
 * writes f->prim, f->temp, f->c2p_mu, f->cell_sig, f->fail and, for zones that hit a
 * bound, f->cons.  iters (nullable, int32 per cell) receives RootFind::iteration_count;
 * an unbracketed root (the reference aborts there, root_find.hpp:115-121) is reported
 * as a failure with iters = -1. */
int phx_con2prim(phx_ctx *ctx, const phx_fluid_params *p, const phx_block *b,
                 const phx_block_fields *f, const int lo[3], const int hi[3], int32_t *iters,
                 phx_stream stream);

/* fluid::EstimateTimestepBlock (fluid.cpp:899-944): dt_out[0] (device) = cfl * min over
 * interior cells.  use_face_speeds = 0 selects the reference's first-call form
 * (cell speeds only, :930-939). */
int phx_estimate_dt(phx_ctx *ctx, const phx_fluid_params *p, const phx_block *b,
                    const phx_block_fields *f, int use_face_speeds, double cfl,
                    double *dt_out, phx_stream stream);

/* ---- one RK stage --------------------------------------------------------------------
 * The fluid tasks of one stage of PhoebusDriver::RungeKuttaStage (phoebus_driver.cpp:
 * 330-722) for one block, fused into five launches: CalculateFluxes (:344), FluxCT's
 * corner EMFs (:345), CalculateFluidSourceTerms (:322, curved metrics only), then
 * FluxCT's rewrite of the B fluxes, FluxDivergence + SumData + AverageIndependentData +
 * UpdateIndependentData (:560-576) fused with the robust ConservedToPrimitive
 * (FillDerived, :588) and the CFL reduction (:720-722) in one kernel.
 * con2prim runs on the interior cells: the ghost zones are refilled by the boundary
 * exchange and re-inverted there (:646-668), which this call does not cover.
 *
 * Reads f->prim (the stage's primitives) and f->c2p_mu (guess); f->flux, f->face_sig,
 * f->emf and f->src (curved metrics) are written as scratch and keep this stage's values.
 * With CT (MHD, 2-D and 3-D) the update forms FluxCT's B fluxes from the corner EMFs in
 * f->emf (the same bits as the rewritten arrays), so f->flux keeps the B fluxes of
 * CalculateFluxes; phx_flux_ct on f afterwards gives the rewritten ones.
 * f->cons is U_s (averaged in place); writes a->u_out (U_{s+1}, after any con2prim
 * re-synchronisation), a->prim_out (may alias f->prim), f->temp, f->c2p_mu, f->cell_sig,
 * f->fail and, when non-NULL, a->iters and a->dt_out (device; cfl * min over the block,
 * with face speeds as fluid.cpp:913-928).  Bit-identical to the separate task entry
 * points called in the same order. */
typedef struct {
  double beta, dt;       /* stage weights (LowStorageIntegrator beta, step dt) */
  const double *u_base;  /* [ncons][cells] U^n                                  */
  double *u_out;         /* [ncons][cells] U_{s+1}; may alias u_base            */
  double *prim_out;      /* [nprim][cells]; may alias f->prim                   */
  int32_t *iters;        /* [cells] con2prim iterations, nullable               */
  double cfl;
  double *dt_out;        /* [1] device, nullable                                */
  /* per-component overrides (as phx_block_fields::vars), nullable: base_vars->cons for
   * u_base; out_vars->cons for u_out and out_vars->prim for prim_out */
  const phx_block_vars *base_vars, *out_vars;
} phx_stage_args;

int phx_fluid_stage(phx_ctx *ctx, const phx_fluid_params *p, const phx_block *b,
                    const phx_block_fields *f, const phx_stage_args *a, phx_stream stream);

/* ---- the same stage without a context, allocating nothing ---------------------------
 * No counterpart in the reference.  For callers that cannot allocate device memory or
 * create a context on the calling path (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The same launches and results as phx_fluid_stage,
 * bit for bit, with the built-in default sweep shapes (no autotuned shape: there is no
 * context).  The CFL partial minima go to a file-scope device array of 65536 doubles
 * (512 KiB) that each device holds from its first use until the process ends; its address
 * is looked up by phx_create, or by the first call on a device with that device current.
 * Call with the stream's device current (stream NULL = that device's null stream).
 * Limit: with a->dt_out set, at most 65536 * 256 = 16,777,216 interior cells per call;
 * larger blocks are refused with PHX_ERR_INVALID before anything is launched.  The array
 * keeps the last call's partials; every call writes the entries it reads before reading
 * them.  Calls with dt_out set that overlap on one device (different streams, or threads)
 * race on the one array: order them.  On failure phx_last_error(NULL) returns the calling
 * thread's message. */
int phx_fluid_stage_noalloc(const phx_fluid_params *p, const phx_block *b,
                            const phx_block_fields *f, const phx_stage_args *a,
                            phx_stream stream);

/* ---- batched con2prim (independent states) ---------------------------------------
 * The per-zone solve of con2prim_robust.hpp:299-453 over n states at spherical
 * Kerr-Schild points coords = [r[n], theta[n]] (phi = 0), or Minkowski when
 * p->geom.kind == PHX_GEOM_MINKOWSKI (coords may then be NULL).  cons [ncons][n] (in,
 * rewritten where a bound was hit), prim [nprim][n] out, mu [n] in (guess) / out,
 * fail [n] out, iters [n] out (nullable). */
int phx_con2prim_batch(phx_ctx *ctx, const phx_fluid_params *p, int64_t n,
                       const double *coords, double *cons, double *prim, double *mu,
                       double *fail, int32_t *iters, phx_stream stream);

/* prim2con::p2c (prim2con.hpp:49-102) over the same batched layout. */
int phx_prim_to_con_batch(phx_ctx *ctx, const phx_fluid_params *p, int64_t n,
                          const double *coords, const double *prim, double *cons,
                          phx_stream stream);

/* ---- tile-shape autotuning (no reference counterpart: Kokkos picks team sizes) -----
 * phx_autotune times every admissible y/z flux sweep shape on the caller's fields --
 * overwriting f->flux and f->face_sig -- and keeps the fastest in the context for this
 * (recon, riemann, mhd, geometry, block size); phx_calculate_fluxes and phx_fluid_stage
 * then use it.  A shape (ti, ts) with ts > 0 is a tile of ti cells along i by ts along
 * the sweep; with ts < 0, one 64-lane wave of ti columns (a power of two) by 64 / ti rows
 * marching along the sweep in -ts segments of the faces.  Results do not depend on the
 * shape.  reps = timed launches per candidate.  phx_tile_shape reports the shape in use
 * for dir 1 (y) or 2 (z); phx_set_tile_shape fixes one (ts = 0: the default, marching
 * waves of 8 x 8 in segments of at least 24 faces where that makes >= 1024 waves, else
 * tiles of 8 x up to 64). */
int phx_autotune(phx_ctx *ctx, const phx_fluid_params *p, const phx_block *b,
                 const phx_block_fields *f, int reps, phx_stream stream);
int phx_tile_shape(phx_ctx *ctx, const phx_fluid_params *p, const phx_block *b, int dir,
                   int *ti, int *ts);
int phx_set_tile_shape(phx_ctx *ctx, const phx_fluid_params *p, const phx_block *b, int dir,
                       int ti, int ts);

#ifdef __cplusplus
}
#endif
#endif /* PHX_H_ */
