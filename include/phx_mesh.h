/*
 * phx_mesh.h -- a uniform 3-D mesh of equal meshblocks on one device, advanced through the
 * RK2 cycle of PhoebusDriver (phoebus_driver.cpp:204-738) with everything the fluid path
 * needs between blocks: the ghost-zone exchange, the physical boundaries, and the fixup
 * package.
 *
 * Per stage, on all blocks at once (each kernel is one launch over a work list of the
 * blocks): CalculateFluidSourceTerms (curved metrics), CalculateFluxes, FluxCT,
 * FixFluxes (:323-366; fixup.cpp:728-1078, gr_outflow faces), FluxDivergence + SumData +
 * AverageIndependentData + UpdateIndependentData (:560-576), FillDerived con2prim +
 * ConservedToPrimitiveFixup + ApplyFloors (:586-597), the boundary exchange of every
 * FillGhost field (:655-659; with bc_vars = primitive, fluid.cpp:159-165: the primitives,
 * temperature, con2prim root, fail flag and conserved variables), ApplyBoundaryConditions
 * (gr_outflow = a copy of the first interior layer, phoebus_boundaries.cpp:118-300),
 * ConvertBoundaryConditions (:303-448: the outer-x1 inflow check of FMKS, then p2c on the
 * ghost slabs of every face that is not a block boundary), and FillDerived + fixups again
 * (:662-676); after the last stage, EstimateTimestep over all blocks.
 *
 * The first FillDerived of a stage inverts, besides the interior, only the ghost cells
 * the c2p fixup reads (the face-adjacent ghost layer); floors there act on the interior.
 * Every other ghost value it would produce is overwritten by the exchange and the
 * boundary conditions before anything reads it, so the results are those of the
 * reference's IndexDomain::entire passes.
 *
 * Blocks are numbered gid = lx1 + nb1 * (lx2 + nb2 * lx3), x1 fastest.  Block coordinates:
 * xmin_b = xmin + (xmax - xmin) * lx / nb per dimension, dx = (xmax - xmin) / nx (the
 * reference's mesh generator lives in parthenon, which is absent; its rounding of block
 * edges is not reproduced bit for bit).
 */
#ifndef PHX_MESH_H_
#define PHX_MESH_H_

#include <stdint.h>

#include "phx.h"
#include "phx_driver.h"
#include "phx_fixup.h"
#include "phx_mlp.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct phx_mesh phx_mesh;

typedef struct {
  int nx[3];              /* <parthenon/mesh> nx1..nx3 (all > 1: 3-D)                    */
  int block_nx[3];        /* <parthenon/meshblock> nx1..nx3; each divides nx              */
  int ng;                 /* nghost                                                       */
  double xmin[3], xmax[3];
  int bc[6];              /* ix1, ox1, ix2, ox2, ix3, ox3: PHX_BC_OUTFLOW (gr_outflow with
                             bc_vars = primitive) or PHX_BC_PERIODIC (both sides)         */
  int flux_fixup;         /* fixup/enable_flux_fixup                                      */
  int ox1_inflow_check;   /* fixup/enable_ox1_fmks_inflow_check (FMKS only)               */
  double cfl;             /* fluid/cfl                                                    */
  double dt_init_fact;    /* parthenon/time/dt_init_fact                                  */
} phx_mesh_desc;

/* Allocates every block's fields on the current device (two conserved containers). */
int phx_mesh_create(phx_ctx *ctx, const phx_fluid_params *p, const phx_fixup_params *fx,
                    const phx_mesh_desc *d, phx_mesh **out);
int phx_mesh_destroy(phx_mesh *m);
int phx_mesh_nblocks(const phx_mesh *m, int *nblocks, int64_t *interior_cells);
/* geometry of block gid */
int phx_mesh_block(const phx_mesh *m, int gid, phx_block *b);
/* host primitives [nprim][cells] over the entire block gid, as a problem generator sets
 * them (ghost zones too) */
int phx_mesh_set_prim(phx_mesh *m, int gid, const double *host_prim);
/* p2c, floors, exchange, boundaries, conversion, con2prim; the first time step
 * (PostInitializationCommunication, phoebus_driver.cpp:147-202) */
int phx_mesh_initialize(phx_mesh *m);
/* one RK2 step of size dt (clipped to the tlim of the last execute call) */
int phx_mesh_step(phx_mesh *m);
/* steps until time >= tlim or max_steps steps were taken */
int phx_mesh_execute(phx_mesh *m, double tlim, int max_steps);
int phx_mesh_time(const phx_mesh *m, double *time, double *dt, int64_t *ncycle);
/* a field (PHX_DRIVER_PRIM, _CONS (base container), _FAIL, _C2P_MU) of block gid to host */
int phx_mesh_get(const phx_mesh *m, int gid, int what, double *host);
/* con2prim iteration counts [cells] of block gid (int32): on the interior those of the
 * last stage's inversion of U_{s+1}, on the ghosts those of the last FillDerived */
int phx_mesh_get_iters(const phx_mesh *m, int gid, int32_t *host);
/* the learned bracket for every con2prim of the cycle of the cells with no previous root
 * (c2p_mu <= 0; the others start from it, as in the plain solver) (mlp non-NULL; it must
 * outlive its use), or the plain solver with the c2p_mu warm start everywhere (NULL, the
 * default); as phx_pack_set_learned */
int phx_mesh_set_learned(phx_mesh *m, const phx_mlp *mlp);
/* bracket misses and network-bracketed solves since the last reset */
int phx_mesh_learned_stats(phx_mesh *m, int reset, int64_t *misses, int64_t *solves);
/* tunes the y/z flux tile shapes on the mesh's own fields (overwriting fluxes) */
int phx_mesh_autotune(phx_mesh *m, int reps);

#ifdef __cplusplus
}
#endif
#endif /* PHX_MESH_H_ */
