/*
 * phx_pack.h -- the fluid stage over a pack of meshblocks of different sizes.
 *
 * Replaces the per-MeshData tasks of one RK stage (phoebus_driver.cpp:323-576: sources,
 * CalculateFluxes, FluxCT, FluxDivergence + SumData + AverageIndependentData +
 * UpdateIndependentData, then FillDerived con2prim on the interior and the dt partials)
 * for all blocks of a MeshData partition at once: each kernel of the stage is ONE launch
 * whose workgroups walk a work list over the blocks (a prefix sum of per-block workgroup
 * counts, searched by each workgroup), so the launch count per stage does not grow with
 * the number of blocks.  A stage, or a sequence of stages, can be captured into a HIP
 * graph and replayed; the time step can be read from device memory so that replays need
 * no host value.  Blocks must be 3-D and MHD (the torus configurations); results per block
 * are those of phx_fluid_stage on that block (FluxCT's B fluxes are formed in the update
 * from the corner EMFs there too, so each block's flux keeps the B fluxes of
 * CalculateFluxes).
 *
 * Containers follow parthenon's rk2 LowStorageIntegrator (phoebus_driver.cpp:289-299,
 * 560-576): stage 1 reads U_s = U^n = f.cons and writes U_1 = cons1; stage 2 averages U_1
 * with U^n (beta = 1/2) and writes U^{n+1} into f.cons.  Primitives are updated in place.
 */
#ifndef PHX_PACK_H_
#define PHX_PACK_H_

#include <stdint.h>

#include "phx.h"
#include "phx_fixup.h"
#include "phx_mlp.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct phx_pack phx_pack;
typedef struct phx_graph phx_graph;

typedef struct {
  phx_block block;
  phx_block_fields f; /* prim, cons (U^n), temp, c2p_mu, fail, cell_sig, face_sig, flux,
                         emf, src (curved metrics) */
  double *cons1;      /* the "1" container */
  int32_t *iters;     /* con2prim iterations, nullable */
} phx_pack_block;

typedef struct {
  int stage;            /* 1 or 2 (rk2) */
  double dt;            /* used when dt_dev is NULL */
  const double *dt_dev; /* device scalar read at run time, nullable */
  double cfl;
  double *dt_out;       /* device: cfl * min over the pack of the CFL times, nullable */
} phx_pack_stage_args;

/* Copies the block table to the device; the fields stay caller-owned. */
int phx_pack_create(phx_ctx *ctx, const phx_fluid_params *p, int nblocks,
                    const phx_pack_block *blocks, phx_pack **out);
/* PHX_ERR_INVALID (and nothing freed) while graphs captured from the pack exist. */
int phx_pack_destroy(phx_pack *pack);
int phx_pack_nblocks(const phx_pack *pack, int *n, int64_t *interior_cells);
/* Enables (fx non-NULL) or disables (NULL) fixup::ConservedToPrimitiveFixup and
 * fixup::ApplyFloors after each stage's con2prim (phoebus_driver.cpp:586-597), on the
 * interior cells of every block (include/phx_fixup.h for the semantics); the stage's
 * dt_out then sees the signal speeds after the fixups.  Off by default. */
int phx_pack_set_fixup(phx_ctx *ctx, phx_pack *pack, const phx_fixup_params *fx);
/* Selects the learned con2prim bracket (include/phx_mlp.h) for the stage's inversion
 * (mlp non-NULL; the network must outlive its use) or the plain solver (NULL, the
 * default).  With the network the fused update inverts the cells that have a previous
 * root (c2p_mu > 0) from it, as the plain stage does, and queues the others; one more
 * launch brackets the queued cells with the network (its MFMA forward pass fused with the
 * bracketed solve) and ends at once when the queue is empty.  The first call with a
 * network allocates the queue (8 bytes per cell of the pack; PHX_ERR_NOMEM if that
 * fails), kept until phx_pack_destroy. */
int phx_pack_set_learned(phx_ctx *ctx, phx_pack *pack, const phx_mlp *mlp);
/* bracket misses (the plain [0, 1] solve ran) and network-bracketed solves since the last
 * reset */
int phx_pack_learned_stats(phx_pack *pack, int reset, int64_t *misses, int64_t *solves);
/* One RK stage over the pack (4-10 launches). */
int phx_pack_stage(phx_ctx *ctx, phx_pack *pack, const phx_pack_stage_args *a,
                   phx_stream stream);
/* Times every admissible y/z tile shape on the pack's own fields (overwriting fluxes) and
 * keeps the fastest for this pack (results do not depend on it).  Tuning rewrites the
 * pack's device work lists, which captured graphs reference with the grid sizes and
 * shapes of their capture: it returns PHX_ERR_INVALID while any graph of the pack exists
 * (tune first, then capture; or destroy the graphs, tune, and capture again). */
int phx_pack_autotune(phx_ctx *ctx, phx_pack *pack, int reps, phx_stream stream);
int phx_pack_tile_shape(const phx_pack *pack, int dir, int *ti, int *ts);
/* Captures nstages stages (in order) on a non-default stream into a replayable graph.  A
 * graph replays the stages as configured at capture (tile shapes, fixups, learned
 * bracket); later phx_pack_set_fixup / phx_pack_set_learned calls affect new captures
 * and eager stages only.  The pack must outlive its graphs. */
int phx_pack_graph_create(phx_ctx *ctx, phx_pack *pack, int nstages,
                          const phx_pack_stage_args *stages, phx_stream stream,
                          phx_graph **out);
int phx_graph_launch(phx_graph *g, phx_stream stream);
int phx_graph_destroy(phx_graph *g);

#ifdef __cplusplus
}
#endif
#endif /* PHX_PACK_H_ */
