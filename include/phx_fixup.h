/*
 * phx_fixup.h -- the fixup package's repair of con2prim failures and its floors and
 * ceilings, the two tasks that follow every FillDerived con2prim of a stage
 * (phoebus_driver.cpp:586-597 after the update, :664-676 after the boundary exchange).
 *
 * Switches are those of fixup::Initialize (fixup.cpp:63-135, 217-243) and are normalised
 * the same way here: enable_floors is forced on by enable_c2p_fixup or enable_mhd_floors,
 * and MHD floors are off for hydro.  Floors and the gamma / sie ceilings are the ones of
 * phx_fluid_params.bounds (the values con2prim uses); the b^2 ceilings of the MHD floors
 * are MHDCeilings' ConstantBsqRat (bsqorho0_ceiling, bsqou0_ceiling; defaults 50, 2500).
 * Radiation floors and the radiation fixup are out of scope.
 */
#ifndef PHX_FIXUP_H_
#define PHX_FIXUP_H_

#include "phx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* fixup/fluid_c2p_failure_strategy (fixup.hpp:28-31) */
enum { PHX_FAILURE_INTERPOLATE = 0, PHX_FAILURE_FLOORS = 1 };

typedef struct {
  int enable_c2p_fixup;  /* fixup/enable_c2p_fixup        */
  int failure_strategy;  /* PHX_FAILURE_*                 */
  int enable_floors;     /* fixup/enable_floors           */
  int enable_mhd_floors; /* fixup/enable_mhd_floors       */
  double bsqorho_max;    /* fixup/bsqorho0_ceiling        */
  double bsqou_max;      /* fixup/bsqou0_ceiling          */
} phx_fixup_params;

/* fixup::ConservedToPrimitiveFixup (fixup_c2p.cpp:44-356), fluid part, interior cells.
 * Reads f->fail (1.0 success / 0.0 failure) on the interior and the first ghost layer;
 * for each failed cell writes f->prim (rho, v, u averaged over the valid face neighbours,
 * or the near-vacuum state 100*SMALL at rest; gamma ceiling with the rescale
 * sqrt(gmax^2-1)/(W^2-1) of fixup_c2p.cpp:271; then T, P, Gamma_1), f->temp, f->cons
 * (D, S_i, tau; B untouched) and f->cell_sig.  A failed neighbour contributes exactly
 * zero (weight fail = 0; the reference multiplies a value that the same pass may be
 * rewriting, which differs only for non-finite values).  No-op unless enable_c2p_fixup. */
int phx_c2p_fixup(phx_ctx *ctx, const phx_fluid_params *p, const phx_fixup_params *fx,
                  const phx_block *b, const phx_block_fields *f, phx_stream stream);

/* fixup::ApplyFloors (fixup.cpp:285-530), fluid part, cells lo..hi (inclusive, (i, j, k)
 * order; the reference's domain is the entire block).  Floors (and the MHD b^2 ceilings)
 * are injected at rest in the normal-observer frame into f->cons and re-inverted with the
 * robust con2prim (warm start f->c2p_mu; f->prim, f->temp, f->c2p_mu, f->cell_sig
 * written), falling back to the floor state if that fails; the gamma / sie ceiling
 * rescales v by sqrt(gmax^2-1)/(W^2-1) (fixup.cpp:510) and rewrites cons; T, P, Gamma_1
 * follow.  f->fail is not written (as the reference).  No-op unless floors are enabled. */
int phx_apply_floors(phx_ctx *ctx, const phx_fluid_params *p, const phx_fixup_params *fx,
                     const phx_block *b, const phx_block_fields *f, const int lo[3],
                     const int hi[3], phx_stream stream);

#ifdef __cplusplus
}
#endif
#endif /* PHX_FIXUP_H_ */
